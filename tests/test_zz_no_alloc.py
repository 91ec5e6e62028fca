"""The allocation-free fused render + L2 forward
(ngt_render_density_l2_fwd_noalloc: its scratch is a static device array)
called through the C ABI against ngt_render_density_l2_fwd with a
caller-allocated workspace: identical images and loss bit for bit (the same
kernels in the same order), on two calls in a row with the second on a
smaller grid and image, with and without a caller-made brick copy and with
the gradient record; and sizes beyond the static array are rejected."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from ngt import ops, _lib, synth
    from ngt.renderer import Renderer
    return ops, _lib, _lib.load(), synth, Renderer


def _scene(synth, Renderer, dims, img, depth, views, seed):
    grid_t, cams = synth.ring_setup(dims, img, depth, views)
    M, V, P, F = Renderer()._transform_params([grid_t], cams)
    rho = torch.from_numpy(synth.blobs(dims, 6, seed=seed)[None]).cuda()
    tgt = torch.rand((1, views) + tuple(img) + (3,), device="cuda", generator=torch.Generator("cuda").manual_seed(seed))
    return rho, tgt, (M, V, P, F), float(cams[0].depth_step), (depth,) + tuple(img)


def _call(env, rho, tgt, cams, ds, out_dhw, noalloc, use_brick, use_record):
    ops, _lib, lib, _, _ = env
    c = ops._Cams(*cams)
    lp, _, keep = _lib.host_floats((0.9, 0.8, 0.7), 3)
    s_d = _lib.shape5(rho)
    od = _lib.dhw(out_dhw)
    st = _lib.stream(rho.device)
    img = torch.full((1, c.nv) + tuple(out_dhw[1:]) + (4,), float("nan"), device="cuda")
    loss = torch.full((), float("nan"), device="cuda")
    brk = None
    if use_brick:
        brk = torch.empty(lib.ngt_bricked_grid_bytes(s_d), dtype=torch.uint8, device="cuda")
        _lib.check(lib.ngt_bricked_grid(_lib.ptr(rho), s_d, 0, _lib.ptr(brk), st))
    rec, rbytes = None, 0
    if use_record:
        rbytes = lib.ngt_render_density_l2_record_size(s_d, c.m, c.v, c.p, c.f, c.nv, od)
        rec = torch.zeros(rbytes, dtype=torch.uint8, device="cuda")
    bp = None if brk is None else _lib.ptr(brk)
    rp = None if rec is None else _lib.ptr(rec)
    if noalloc:
        rc = lib.ngt_render_density_l2_fwd_noalloc(_lib.ptr(rho), s_d, lp, ds, c.m, c.v, c.p, c.f, c.nv, od, 1, 0, 0,
                                                   _lib.ptr(tgt), 0.5, bp, rp, rbytes, _lib.ptr(img), _lib.ptr(loss),
                                                   st)
    else:
        ws, nb = _lib.workspace(lib.ngt_render_density_l2_fwd_workspace_size(s_d, 1, c.nv, od), rho.device)
        rc = lib.ngt_render_density_l2_fwd(_lib.ptr(rho), s_d, lp, ds, c.m, c.v, c.p, c.f, c.nv, od, 1, 0, 0,
                                           _lib.ptr(tgt), 0.5, bp, rp, rbytes, _lib.ptr(img), _lib.ptr(loss),
                                           _lib.ptr(ws), nb, st)
    _lib.check(rc)
    torch.cuda.synchronize()
    return img, loss, rec


@pytest.mark.parametrize("use_brick,use_record", [(False, False), (True, False), (True, True)])
def test_noalloc_matches_workspace_entry(env, use_brick, use_record):
    ops, _lib, lib, synth, Renderer = env
    for dims, img_hw, depth, views, seed in (((96, 80, 112), (90, 160), 96, 3, 1), ((24, 30, 20), (36, 40), 24, 2, 2)):
        print("case: noalloc dims=%s img=%s views=%d brick=%s record=%s" % (dims, img_hw, views, use_brick,
                                                                              use_record), flush=True)
        rho, tgt, cams, ds, out_dhw = _scene(synth, Renderer, dims, img_hw, depth, views, seed)
        a_img, a_loss, a_rec = _call(env, rho, tgt, cams, ds, out_dhw, True, use_brick, use_record)
        b_img, b_loss, b_rec = _call(env, rho, tgt, cams, ds, out_dhw, False, use_brick, use_record)
        assert torch.equal(a_img, b_img)
        assert torch.equal(a_loss, b_loss)
        assert bool(torch.isfinite(a_loss)) and float(a_loss) > 0
        if use_record:
            assert torch.equal(a_rec, b_rec)


def test_noalloc_rejects_sizes_beyond_the_static_scratch(env):
    ops, _lib, lib, synth, Renderer = env
    cap = lib.ngt_static_scratch_bytes()
    assert cap > 0
    dims = (256, 256, 256)  # the brick copy alone needs ~148 MB
    rho = torch.zeros((1,) + dims + (1,), device="cuda")
    grid_t, cams = synth.ring_setup(dims, (16, 16), 8, 1)
    M, V, P, F = Renderer()._transform_params([grid_t], cams)
    c = ops._Cams(M, V, P, F)
    s_d = _lib.shape5(rho)
    od = _lib.dhw((8, 16, 16))
    assert lib.ngt_render_density_l2_fwd_workspace_size(s_d, 1, 1, od) > cap
    lp, _, keep = _lib.host_floats((1.0, 1.0, 1.0), 3)
    img = torch.empty((1, 1, 16, 16, 4), device="cuda")
    tgt = torch.zeros((1, 1, 16, 16, 3), device="cuda")
    loss = torch.empty((), device="cuda")
    rc = lib.ngt_render_density_l2_fwd_noalloc(_lib.ptr(rho), s_d, lp, 0.01, c.m, c.v, c.p, c.f, 1, od, 1, 0, 0,
                                               _lib.ptr(tgt), 1.0, None, None, 0, _lib.ptr(img), _lib.ptr(loss),
                                               _lib.stream(rho.device))
    assert rc == 3 and b"static" in lib.ngt_last_error()
