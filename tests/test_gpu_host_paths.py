"""GPU parity of the host-side callers built on the ops (Renderer.render_density
fused / light-grid / frustum paths, the ``sample_gradients`` gather backward,
VelocityGrid self-advection, DensityGrid.warped) against the FP64 oracle.

Tolerances: forward 1e-5, gradients 1e-4 normwise relative, as for the ops.
"""
import numpy as np
import pytest
import torch

from oracle import BORDER, LINEAR
from oracle import resample as R
from oracle import blend as B
from oracle import raymarch as RM
from oracle import advect as A

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
GRAD_TOL = 1e-4


@pytest.fixture(scope="module")
def ngt():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import ngt as pkg
    from ngt import _lib
    _lib.load()
    return pkg


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float32))).to("cuda:0")


def host(t):
    return t.detach().cpu().double().numpy()


def rel(got, exp):
    return float(np.max(np.abs(got - exp))) / max(float(np.max(np.abs(exp))), 1e-30)


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def smooth(rng, shape, amp, sigma=1.5):
    from scipy.ndimage import gaussian_filter
    f = gaussian_filter(rng.standard_normal(shape), sigma)
    return amp * f / np.max(np.abs(f))


def _scene(ngt, dims=(12, 12, 12), hw=(10, 14), depth=16, views=3):
    from ngt import synth
    grid, cams = synth.ring_setup(dims, hw, depth, views)
    # move the grid off the symmetric lattice (see test_gpu_parity.ring_args)
    grid.set_translation([0.0123, -0.0217, 0.0311])
    return grid, cams


def _op_args(ngt, grid, cams):
    from ngt.renderer import Renderer
    M, V, P, F = Renderer()._transform_params([grid], cams)
    return [np.asarray(m, np.float32) for m in M], [np.asarray(v, np.float32) for v in V], \
        [np.asarray(p, np.float32) for p in P], [np.asarray(f, np.float32) for f in F]


@pytest.mark.parametrize("path", ["fused_light", "light_grid", "frustum"])
def test_render_density_end_to_end(ngt, path):
    from ngt.renderer import Renderer
    from ngt.lighting import Light
    rng = np.random.default_rng(21)
    dims, hw, depth = (12, 12, 12), (10, 14), 16
    grid, cams = _scene(ngt, dims, hw, depth)
    rho = f32(rng.uniform(0.0, 0.6, (1,) + dims + (1,)))
    lights = [Light(color=(0.9, 0.8, 0.7), intensity=1.2), Light(color=(0.1, 0.2, 0.3), intensity=0.5)]
    rgb = sum(lt.rgb() for lt in lights)
    kw = {"fused_light": dict(fused=True), "light_grid": dict(fused=True, fused_lighting=False),
          "frustum": dict(fused=False)}[path]
    r = Renderer(filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="NONE", blend_mode="BEER_LAMBERT", **kw)
    rt = dev(rho).requires_grad_()
    grid.data = rt
    imgs = r.render_density(grid, lights, cams, cut_alpha=False)
    got = np.stack([host(i) for i in imgs], 1)  # [N, V, H, W, 4]; cut_alpha=False keeps the raw optical depth
    Ms, Vs, Ps, Fs = _op_args(ngt, grid, cams)
    ds = float(cams[0].depth_step)
    cat = np.concatenate([rho * rgb, rho], axis=-1) * ds
    out_dhw = (depth,) + hw
    if path == "frustum":
        fr = R.sample_grid_transform(cat, Ms, Vs, Ps, Fs, out_dhw, LINEAR, BORDER)
        exp = np.stack([B.reduce_grid_blend(fr[:, v], B.BEER_LAMBERT) for v in range(len(cams))], 1)
    else:
        exp = RM.raymarch_grid_transform(cat, Ms, Vs, Ps, Fs, out_dhw, LINEAR, BORDER, B.BEER_LAMBERT)
    assert rel(got, exp) < FWD_TOL
    # gradient of a loss on RGB and transmittance w.r.t. the density
    g = f32(rng.standard_normal(got.shape))
    loss = sum((imgs[v] * dev(g[:, v])).sum() for v in range(len(cams)))
    loss.backward()
    g_raw = g
    if path == "frustum":
        g_cat = np.zeros_like(cat)
        for v in range(len(cams)):
            gf = B.reduce_grid_blend_grad(g_raw[:, v], exp[:, v], fr[:, v], B.BEER_LAMBERT)
            g_cat += R.sample_grid_transform_grad(cat, gf[:, None], Ms, Vs[v:v + 1], Ps[v:v + 1], Fs[v:v + 1],
                                                  LINEAR, BORDER)
    else:
        g_cat = RM.raymarch_grid_transform_grad(cat, exp, g_raw, Ms, Vs, Ps, Fs, out_dhw, LINEAR, BORDER,
                                                B.BEER_LAMBERT)
    eg = ds * (g_cat[..., :3] @ rgb + g_cat[..., 3])[..., None]
    assert rel(host(rt.grad), eg) < GRAD_TOL


def test_sample_gradients_gather(ngt):
    """sample_gradients=True: d/dx of camera sampling is the reverse-mode
    (OS <- frustum) gather of dy summed over the views (renderer.py:296-323)."""
    from ngt.renderer import Renderer
    rng = np.random.default_rng(22)
    dims, hw, depth = (10, 11, 12), (9, 8), 7
    grid, cams = _scene(ngt, dims, hw, depth, views=2)
    x = f32(rng.standard_normal((1,) + dims + (2,)))
    r = Renderer(filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="NONE", sample_gradients=True)
    xt = dev(x).requires_grad_()
    y = r.sample_camera(xt, grid, cams, inverse=False, squeeze_batch=False)
    Ms, Vs, Ps, Fs = _op_args(ngt, grid, cams)
    out_dhw = (depth,) + hw
    exp = R.sample_grid_transform(x, Ms, Vs, Ps, Fs, out_dhw, LINEAR, BORDER)
    assert rel(host(y), exp) < FWD_TOL
    dy = f32(rng.standard_normal(exp.shape))
    y.backward(dev(dy))
    eg = np.zeros_like(x)
    for v in range(len(cams)):
        eg += R.sample_grid_transform(dy[:, v], Ms, Vs[v:v + 1], Ps[v:v + 1], Fs[v:v + 1], dims, LINEAR, BORDER,
                                      coordinate_mode=R.TRANSFORM_LINDEPTH_REVERSE)[:, 0]
    assert rel(host(xt.grad), eg) < GRAD_TOL


@pytest.mark.parametrize("order,clamp", [(1, "NONE"), (2, "MC_SMOOTH")])
@pytest.mark.parametrize("fused", [True, False])
def test_velocity_self_advect(ngt, order, clamp, fused):
    from ngt.data_structures import VelocityGrid
    from ngt.renderer import Renderer
    rng = np.random.default_rng(31 + order)
    D, H, W = 9, 10, 11
    comps = [f32(smooth(rng, s, 1.6))[..., None] for s in ((D, H, W + 1), (D, H + 1, W), (D + 1, H, W))]
    ts = [dev(c[None]).requires_grad_() for c in comps]
    r = Renderer(filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="NONE", fused=fused)
    vg = VelocityGrid([D, H, W], x=ts[0], y=ts[1], z=ts[2], warp_renderer=r)
    if not fused:
        vg._can_fuse = lambda: False
    outs = vg.warped(None, order=order, dt=1.0, clamp=clamp)
    exp = A.velocity_self_advect_forward(*comps, dt=1.0, order=order, clamp=clamp, boundary=BORDER)
    for o, e in zip(outs, exp):
        assert rel(host(o)[0], e) < FWD_TOL
    gs = [f32(rng.standard_normal(e.shape)) for e in exp]
    loss = sum((o * dev(g[None])).sum() for o, g in zip(outs, gs))
    loss.backward()
    eg = A.velocity_self_advect_backward(*comps, gs, dt=1.0, order=order, clamp=clamp, boundary=BORDER,
                                         semantics="exact")
    for t, e in zip(ts, eg):
        assert rel(host(t.grad)[0], e) < GRAD_TOL


@pytest.mark.parametrize("fused", [True, False])
def test_density_warped_staggered(ngt, fused):
    """DensityGrid.warped with a MAC VelocityGrid (MacCormack + MC_SMOOTH),
    fused advect op and the reference's LUT composition."""
    from ngt.data_structures import VelocityGrid, DensityGrid
    from ngt.renderer import Renderer
    rng = np.random.default_rng(41)
    D, H, W = 10, 9, 12
    rho = f32(np.clip(smooth(rng, (D, H, W), 1.0), 0.0, 1.0))[..., None]
    comps = [f32(smooth(rng, s, 1.7))[..., None] for s in ((D, H, W + 1), (D, H + 1, W), (D + 1, H, W))]
    ts = [dev(c[None]).requires_grad_() for c in comps]
    rt = dev(rho[None]).requires_grad_()
    r = Renderer(filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="NONE", fused=True)
    vg = VelocityGrid([D, H, W], x=ts[0], y=ts[1], z=ts[2], warp_renderer=r)
    dg = DensityGrid([D, H, W], d=rt)
    if fused:
        out = dg.warped(vg, order=2, dt=1.0, clamp="MC_SMOOTH")
    else:
        out = vg.warp(rt, order=2, dt=1.0, clamp="MC_SMOOTH", fused=False)
    vc = A.centered_from_staggered(*comps)
    exp = A.warp_forward(rho, vc, 1.0, 2, "MC_SMOOTH", BORDER)
    assert rel(host(out)[0], exp) < FWD_TOL
    g = f32(rng.standard_normal(exp.shape))
    out.backward(dev(g[None]))
    eg_rho, eg_v = A.warp_backward(rho, vc, g, 1.0, 2, "MC_SMOOTH", BORDER)
    assert rel(host(rt.grad)[0], eg_rho) < GRAD_TOL
    for t, e in zip(ts, A.centered_from_staggered_adjoint(eg_v)):
        assert rel(host(t.grad)[0], e) < GRAD_TOL


@pytest.mark.parametrize("semantics", ["exact", "reference"])
def test_inverse_raymarch_and_its_gradient(ngt, semantics):
    """Inverse raymarching (renderer.py:372-398): the ADDITIVE raymarch
    gradient op applied to images lifts them into the grid; its gradient is
    the ADDITIVE forward raymarch (ops_loader.py:98-118)."""
    from ngt.renderer import Renderer
    from ngt import ops
    rng = np.random.default_rng(51)
    dims, hw, depth = (12, 11, 13), (10, 14), 16
    grid, cams = _scene(ngt, dims, hw, depth)
    imgs = f32(rng.uniform(0.0, 1.0, (1, len(cams)) + hw + (2,)))
    Ms, Vs, Ps, Fs = _op_args(ngt, grid, cams)
    out_dhw = (depth,) + hw
    r = Renderer(filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="NONE", blend_mode="ADDITIVE")
    it = dev(imgs).requires_grad_()
    with ops.semantics(semantics):
        vol = r.raymarch_camera(it, grid, cams, inverse=True, squeeze_batch=False)
        exp = RM.raymarch_grid_transform_grad(np.zeros((1,) + dims + (2,)), imgs, imgs, Ms, Vs, Ps, Fs, out_dhw,
                                              LINEAR, BORDER, B.ADDITIVE, semantics=semantics)
        assert tuple(vol.shape) == exp.shape
        assert rel(host(vol), exp) < GRAD_TOL
        G = f32(rng.standard_normal(exp.shape))
        vol.backward(dev(G))
    eg = RM.raymarch_grid_transform(G, Ms, Vs, Ps, Fs, out_dhw, LINEAR, BORDER, B.ADDITIVE)
    assert rel(host(it.grad), eg) < FWD_TOL


def test_unproject_min_of_inverse_samples(ngt):
    """Renderer.unproject: minimum over the cameras of each target extruded
    along depth and sampled back into the grid (TRANSFORM_LINDEPTH_REVERSE)."""
    from ngt.renderer import Renderer
    rng = np.random.default_rng(52)
    dims, hw, depth = (10, 9, 11), (8, 12), 6
    grid, cams = _scene(ngt, dims, hw, depth, views=2)
    targets = f32((rng.uniform(0.0, 1.0, (2,) + hw + (1,)) > 0.3).astype(np.float64))
    r = Renderer(filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="NONE")
    got = host(r.unproject(grid, dev(targets), cams))
    Ms, Vs, Ps, Fs = _op_args(ngt, grid, cams)
    exp = np.ones((1,) + dims + (1,))
    for v in range(2):
        tar = np.broadcast_to(targets[v][None, None], (1, depth) + hw + (1,))
        s = R.sample_grid_transform(tar, Ms, Vs[v:v + 1], Ps[v:v + 1], Fs[v:v + 1], dims, LINEAR, BORDER,
                                    R.TRANSFORM_LINDEPTH_REVERSE)[:, 0]
        exp = np.minimum(exp, s)
    assert rel(got, exp) < FWD_TOL
