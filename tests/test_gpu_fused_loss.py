"""GPU parity of the fused render + L2 image loss (render_density_fused_l2):
the loss and images against the FP64 oracle and the unfused composition,
and the density gradient (image gradient formed inside the render adjoint)
against the oracle's render adjoint of 2 * w * g * (image - target).

Tolerances: loss and images 1e-5, gradients 1e-4, normwise relative.
"""
import numpy as np
import pytest
import torch

from oracle import BORDER, CLAMP, LINEAR
from oracle import blend as B
from oracle import raymarch as RM

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
GRAD_TOL = 1e-4


@pytest.fixture(scope="module")
def ops():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from ngt import ops as o, _lib
    _lib.load()
    return o


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float32))).to("cuda:0")


def host(t):
    return t.detach().cpu().double().numpy()


def rel(got, exp):
    return float(np.max(np.abs(got - exp))) / max(float(np.max(np.abs(exp))), 1e-30)


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def _scene(dims=(12, 11, 13), hw=(10, 14), depth=16, views=3):
    from ngt import synth
    from ngt.renderer import Renderer
    grid, cams = synth.ring_setup(dims, hw, depth, views)
    grid.set_translation([0.0123, -0.0217, 0.0311])
    M, V, P, F = Renderer()._transform_params([grid], cams)
    return [np.asarray(a, np.float32) for a in (M, V, P, F)], float(cams[0].depth_step), (depth,) + hw


@pytest.mark.parametrize("bound,gather", [(BORDER, True), (BORDER, False), (CLAMP, True)])
@pytest.mark.parametrize("mode", [B.BEER_LAMBERT, B.ADDITIVE])
def test_fused_l2_loss_and_gradient(ops, bound, gather, mode):
    rng = np.random.default_rng(61 + bound + 2 * gather + mode)
    dims = (12, 11, 13)
    (M, V, P, F), ds, out_dhw = _scene(dims)
    rgb = np.array([0.9, 0.7, 0.5])
    rho = f32(rng.uniform(0.0, 0.5, (1,) + dims + (1,)))
    tgt = f32(rng.uniform(0.0, 0.3, (1, 3) + out_dhw[1:] + (3,)))
    w = 0.37
    bname = {BORDER: "BORDER", CLAMP: "CLAMP"}[bound]
    mname = {B.BEER_LAMBERT: "BEER_LAMBERT", B.ADDITIVE: "ADDITIVE"}[mode]
    old = ops.use_gather_adjoint
    ops.use_gather_adjoint = gather
    try:
        rt = dev(rho).requires_grad_()
        loss, img = ops.render_density_fused_l2(rt, dev(tgt), rgb, ds, M, V, P, F, out_dhw, loss_weight=w,
                                                boundary=bname, blending_mode=mname)
        (loss * 3.0).backward()
    finally:
        ops.use_gather_adjoint = old
    cat = np.concatenate([rho * rgb, rho], -1) * ds
    eimg = RM.raymarch_grid_transform(cat, M, V, P, F, out_dhw, LINEAR, bound, mode)
    assert rel(host(img), eimg) < FWD_TOL
    eloss = w * float(np.sum((eimg[..., :3] - tgt) ** 2))
    assert abs(float(loss) - eloss) / abs(eloss) < FWD_TOL
    gimg = np.zeros_like(eimg)
    gimg[..., :3] = 2.0 * w * 3.0 * (eimg[..., :3] - tgt)
    g_cat = RM.raymarch_grid_transform_grad(cat, eimg, gimg, M, V, P, F, out_dhw, LINEAR, bound, mode)
    eg = ds * (g_cat[..., :3] @ rgb + g_cat[..., 3])[..., None]
    assert rel(host(rt.grad), eg) < GRAD_TOL


@pytest.mark.parametrize("dims", [(1, 11, 13), (12, 1, 13), (12, 11, 1)])
def test_fused_l2_grid_of_extent_one(ops, dims):
    """Grids one cell thick along an axis (the bricked walks' validity bound
    there is the float below 0.5, not 0.5): images and loss against the
    oracle, the recording walk's image against the image walk's bit for bit,
    and the density gradient against the oracle adjoint."""
    rng = np.random.default_rng(64 + sum(dims))
    (M, V, P, F), ds, out_dhw = _scene(dims)
    rho = f32(rng.uniform(0.0, 0.5, (1,) + dims + (1,)))
    tgt = f32(rng.uniform(0.0, 0.3, (1, 3) + out_dhw[1:] + (3,)))
    print("case: fused L2 on a %s grid" % (dims,), flush=True)
    rt = dev(rho).requires_grad_()
    loss, img = ops.render_density_fused_l2(rt, dev(tgt), (1, 1, 1), ds, M, V, P, F, out_dhw)
    loss.backward()
    img2 = ops.render_density_fused(dev(rho), (1, 1, 1), ds, M, V, P, F, out_dhw)
    assert torch.equal(img.detach(), img2.detach())
    cat = np.concatenate([rho, rho, rho, rho], -1) * ds
    eimg = RM.raymarch_grid_transform(cat, M, V, P, F, out_dhw, LINEAR, BORDER, B.BEER_LAMBERT)
    if dims[1] > 1:  # a slab one cell high lies edge-on to the horizontal ring of cameras
        assert np.max(np.abs(eimg[..., 3])) > 0
    assert rel(host(img), eimg) < FWD_TOL
    eloss = float(np.sum((eimg[..., :3] - tgt) ** 2))
    assert abs(float(loss) - eloss) / abs(eloss) < FWD_TOL
    gimg = np.zeros_like(eimg)
    gimg[..., :3] = 2.0 * (eimg[..., :3] - tgt)
    g_cat = RM.raymarch_grid_transform_grad(cat, eimg, gimg, M, V, P, F, out_dhw, LINEAR, BORDER, B.BEER_LAMBERT)
    eg = ds * np.sum(g_cat, axis=-1, keepdims=True)
    assert rel(host(rt.grad), eg) < GRAD_TOL


def test_fused_l2_matches_unfused_composition(ops):
    rng = np.random.default_rng(62)
    dims = (12, 11, 13)
    (M, V, P, F), ds, out_dhw = _scene(dims)
    rho = f32(rng.uniform(0.0, 0.5, (1,) + dims + (1,)))
    tgt = dev(rng.uniform(0.0, 0.3, (1, 3) + out_dhw[1:] + (3,)))
    r1, r2 = dev(rho).requires_grad_(), dev(rho).requires_grad_()
    loss, img = ops.render_density_fused_l2(r1, tgt, (1, 1, 1), ds, M, V, P, F, out_dhw)
    loss.backward()
    img2 = ops.render_density_fused(r2, (1, 1, 1), ds, M, V, P, F, out_dhw)
    loss2 = ((img2[..., :3] - tgt) ** 2).sum()
    loss2.backward()
    assert torch.equal(img, img2)
    assert abs(float(loss) - float(loss2)) <= 1e-5 * abs(float(loss2))
    assert rel(host(r1.grad), host(r2.grad)) < 1e-5


def test_fused_l2_rejects_bad_target(ops):
    (M, V, P, F), ds, out_dhw = _scene()
    rho = torch.zeros((1, 12, 11, 13, 1), device="cuda:0")
    with pytest.raises(ValueError, match="target must be"):
        ops.render_density_fused_l2(rho, torch.zeros((1, 3) + out_dhw[1:] + (4,), device="cuda:0"), (1, 1, 1), ds,
                                    M, V, P, F, out_dhw)


@pytest.mark.parametrize("views", [3, 9])
def test_fused_l2_shared_brick_copy_equals_own_copy(ops, views):
    """The op makes one haloed brick copy (ngt_bricked_grid) for the forward
    and the backward and has the forward record the per-sample gradients;
    the C ABI called without either (each call making its own copy, the
    backward walking the rays itself) gives bit-identical images and loss,
    and the same gradient up to rounding: the gather adjoint ranks samples
    within a bin by LDS atomics, so the order of its sums varies between
    runs, and the recording forms each sample's gradient as
    ds <light, g> (h_k - Q_N) from running sums instead of the adjoint walk's
    prefix sums (1e-5 here; both are checked against the oracle at 1e-4).  9 views take two camera launches, so the per-work-group
    loss partials and recorded rays of both count."""
    from ngt import _lib
    lib = _lib.load()
    rng = np.random.default_rng(70 + views)
    dims = (12, 11, 13)
    (M, V, P, F), ds, out_dhw = _scene(dims, views=views)
    rho = dev(rng.uniform(0.0, 0.5, (1,) + dims + (1,)))
    tgt = dev(rng.uniform(0.0, 0.3, (1, views) + out_dhw[1:] + (3,)))
    rt = rho.clone().requires_grad_()
    loss, img = ops.render_density_fused_l2(rt, tgt, (0.9, 0.8, 0.7), ds, M, V, P, F, out_dhw, loss_weight=0.5)
    (loss * 1.75).backward()
    cams = ops._Cams(M, V, P, F)
    lp, _, keep = _lib.host_floats((0.9, 0.8, 0.7), 3)
    s_d = _lib.shape5(rho)
    od = _lib.dhw(out_dhw)
    st = _lib.stream(rho.device)
    img2 = torch.empty_like(img)
    loss2 = torch.empty((), device="cuda:0")
    fws, fnb = _lib.workspace(lib.ngt_render_density_l2_fwd_workspace_size(s_d, 1, cams.nv, od), rho.device)
    _lib.check(lib.ngt_render_density_l2_fwd(_lib.ptr(rho), s_d, lp, float(ds), cams.m, cams.v, cams.p, cams.f,
                                             cams.nv, od, 1, 0, 0, _lib.ptr(tgt), 0.5, None, None, 0, _lib.ptr(img2),
                                             _lib.ptr(loss2), _lib.ptr(fws), fnb, st))
    g2 = torch.empty_like(rho)
    one = torch.full((), 1.75, device="cuda:0")
    bws, bnb = _lib.workspace(lib.ngt_render_density_l2_bwd_workspace_size(s_d, cams.m, cams.v, cams.p, cams.f,
                                                                           cams.nv, od, 1, 0), rho.device)
    _lib.check(lib.ngt_render_density_l2_bwd(_lib.ptr(rho), s_d, lp, float(ds), _lib.ptr(img2), _lib.ptr(tgt), 0.5,
                                             _lib.ptr(one), cams.m, cams.v, cams.p, cams.f, cams.nv, od, 1, 0, 0,
                                             None, 0, _lib.ptr(g2), _lib.ptr(bws), bnb, st))
    torch.cuda.synchronize()
    assert torch.equal(img, img2)
    assert torch.equal(loss, loss2)
    assert rel(host(rt.grad), host(g2)) < 1e-5
    # and the loss against the oracle
    cat = np.concatenate([host(rho) * np.array([0.9, 0.8, 0.7]), host(rho)], -1) * ds
    eimg = RM.raymarch_grid_transform(cat, M, V, P, F, out_dhw, LINEAR, BORDER, B.BEER_LAMBERT)
    eloss = 0.5 * float(np.sum((eimg[..., :3] - host(tgt)) ** 2))
    assert abs(float(loss) - eloss) / abs(eloss) < FWD_TOL


def test_bricked_grid_rejects_three_channels(ops):
    from ngt import _lib
    lib = _lib.load()
    x = torch.zeros((1, 4, 4, 4, 3), device="cuda:0")
    out = torch.empty(1 << 16, dtype=torch.uint8, device="cuda:0")
    rc = lib.ngt_bricked_grid(_lib.ptr(x), _lib.shape5(x), 0, _lib.ptr(out), _lib.stream(x.device))
    assert rc != 0 and b"Channel" in lib.ngt_last_error()


def _record_rays(ops, rho, tgt, light, ds, M, V, P, F, out_dhw):
    """The fused L2 forward with a record buffer, through the C ABI; returns
    the per-ray records as int32 [pairs * pixels, 12] (RayRec: steps at 8,
    the ray's own sample count at 9)."""
    from ngt import _lib
    lib = _lib.load()
    cams = ops._Cams(M, V, P, F)
    lp, _, keep = _lib.host_floats(light, 3)
    s_d = _lib.shape5(rho)
    od = _lib.dhw(out_dhw)
    st = _lib.stream(rho.device)
    brk = torch.empty(lib.ngt_bricked_grid_bytes(s_d), dtype=torch.uint8, device=rho.device)
    _lib.check(lib.ngt_bricked_grid(_lib.ptr(rho), s_d, 0, _lib.ptr(brk), st))
    rbytes = lib.ngt_render_density_l2_record_size(s_d, cams.m, cams.v, cams.p, cams.f, cams.nv, od)
    rec = torch.zeros(rbytes, dtype=torch.uint8, device=rho.device)
    fws, fnb = _lib.workspace(lib.ngt_render_density_l2_fwd_workspace_size(s_d, 1, cams.nv, od), rho.device)
    img = torch.empty((1, cams.nv) + tuple(out_dhw[1:]) + (4,), device=rho.device)
    loss = torch.empty((), device=rho.device)
    _lib.check(lib.ngt_render_density_l2_fwd(_lib.ptr(rho), s_d, lp, float(ds), cams.m, cams.v, cams.p, cams.f,
                                             cams.nv, od, 1, 0, 0, _lib.ptr(tgt), 1.0, _lib.ptr(brk), _lib.ptr(rec),
                                             rbytes, _lib.ptr(img), _lib.ptr(loss), _lib.ptr(fws), fnb, st))
    torch.cuda.synchronize()
    n = cams.nv * out_dhw[1] * out_dhw[2]
    return rec[:n * 48].view(torch.int32).reshape(n, 12).cpu().numpy()


def test_recorded_rays_are_never_cut(ops):
    """Every recorded ray keeps all its samples (recorded count == the ray's
    own count): the 128^3 five-view workload and off-axis, wide-angle and
    inside-the-grid cameras."""
    from ngt import synth
    from ngt.renderer import Renderer
    from oracle import geometry as G
    grid_t, cams = synth.ring_setup((128, 128, 128), (144, 256), 128, 5)
    M, V, P, F = Renderer()._transform_params([grid_t], cams)
    scenes = [((128, 128, 128), (M, V, P, F), (128, 144, 256))]
    dims = (24, 30, 20)
    Mo = G.op_matrix(G.grid_transform(dims, rotation_deg=(9.0, -31.0, 4.0), center=True,
                                      scale_=[1.2 / dims[2], 0.9 / dims[1], 1.0 / dims[0]])).astype(np.float32)
    views = []
    for dist, trbl, near, yaw, pitch in ((2.4, [0.9, 0.2, 0.1, -0.6], 1.4, 17.0, 0.0),
                                         (1.9, [0.05, 1.6, -0.7, -0.1], 0.6, 40.0, 25.0),
                                         (1.0, [0.8, 0.8, -0.8, -0.8], 0.3, -60.0, 10.0)):
        C = G.rotation_xyz_deg([pitch, yaw, 0.0]) @ G.translation([0.0, 0.0, dist])
        views.append((G.op_matrix(np.linalg.inv(C)).astype(np.float32),
                      G.op_matrix(G.projection(near, dist + 1.5, trbl)).astype(np.float32),
                      G.frustum_params(near, dist + 1.5, trbl).astype(np.float32)))
    scenes.append((dims, ([Mo], [v[0] for v in views], [v[1] for v in views], np.stack([v[2] for v in views])),
                   (40, 36, 48)))
    for dims, (M, V, P, F), out_dhw in scenes:
        print("case: record %s -> %d views %s" % (dims, len(V), out_dhw), flush=True)
        rho = torch.from_numpy(synth.blobs(dims, 8, seed=5)[None]).cuda()
        tgt = torch.zeros((1, len(V)) + tuple(out_dhw[1:]) + (3,), device="cuda:0")
        rec = _record_rays(ops, rho, tgt, (1.0, 1.0, 1.0), 0.01, M, V, P, F, out_dhw)
        assert rec[:, 9].max() > 0
        np.testing.assert_array_equal(rec[:, 8], rec[:, 9])


@pytest.mark.parametrize("path", ["record", "prepass", "scatter"])
def test_gradient_behind_opaque_media(ops, path):
    """A dense volume (per-sample optical depth about 1.4, so the far half of
    every ray is hidden): the exact BEER_LAMBERT adjoint forms each sample's
    suffix sum as a difference of FP32 prefix sums, which loses relative
    accuracy where the gradient is tiny.  Every path (the recording forward,
    the prepass walk with the gather, the scatter adjoint) stays within 1e-5
    of the oracle's largest gradient overall and within 2e-3 of the largest
    gradient among the occluded voxels (oracle |g| below 1e-3 of its max)."""
    from ngt import synth
    from ngt.renderer import Renderer
    dims, hw, depth = (16, 16, 16), (12, 12), 24
    grid, cams = synth.ring_setup(dims, hw, depth, 1)
    M, V, P, F = [np.asarray(a, np.float32) for a in Renderer()._transform_params([grid], cams)]
    ds = float(np.float32(cams[0].depth_step))
    rng = np.random.default_rng(20)
    rho = f32(20.0 * (0.8 + 0.4 * rng.uniform(size=(1,) + dims + (1,))))
    tgt = f32(rng.uniform(0.0, 0.3, (1, 1) + hw + (3,)))
    cat = np.concatenate([rho, rho, rho, rho], -1) * ds
    eimg = RM.raymarch_grid_transform(cat, M, V, P, F, (depth,) + hw, LINEAR, BORDER, B.BEER_LAMBERT)
    gimg = np.zeros_like(eimg)
    gimg[..., :3] = 2.0 * (eimg[..., :3] - tgt)
    g_cat = RM.raymarch_grid_transform_grad(cat, eimg, gimg, M, V, P, F, (depth,) + hw, LINEAR, BORDER,
                                            B.BEER_LAMBERT)
    eg = ds * np.sum(g_cat, -1, keepdims=True)
    gmax = float(np.max(np.abs(eg)))
    occluded = (np.abs(eg) < 1e-3 * gmax) & (np.abs(eg) > 0.0)
    assert np.sum(occluded) > 500
    rt = dev(rho).requires_grad_()
    if path == "record":
        loss, _ = ops.render_density_fused_l2(rt, dev(tgt), (1, 1, 1), ds, M, V, P, F, (depth,) + hw)
        loss.backward()
    else:
        old = ops.use_gather_adjoint
        ops.use_gather_adjoint = path == "prepass"
        try:
            img = ops.render_density_fused(rt, (1, 1, 1), ds, M, V, P, F, (depth,) + hw)
            ((img[..., :3] - dev(tgt)) ** 2).sum().backward()
        finally:
            ops.use_gather_adjoint = old
    err = np.abs(host(rt.grad) - eg)
    region = float(np.max(err[occluded])) / float(np.max(np.abs(eg[occluded])))
    print("%s: per-sample optical depth %.2f, %d occluded voxels; error %.2e of max |g|, %.2e of the occluded "
          "voxels' max |g|" % (path, 20.0 * ds, int(np.sum(occluded)), float(np.max(err)) / gmax, region), flush=True)
    assert float(np.max(err)) <= 1e-5 * gmax
    assert region <= 2e-3
