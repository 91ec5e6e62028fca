"""GPU parity of every op of libngt_transport.so against the FP64 oracle on
the same seeded inputs.

Tolerances (written per assertion): forward 1e-5 normwise relative
(max |gpu - oracle| / max |oracle|); gradients 1e-4 normwise relative (fp32
atomics accumulate in arbitrary order).  Masks (BORDER zero regions, MC_SMOOTH
revert decisions) are compared exactly where they are observable.
"""
import numpy as np
import pytest
import torch

from oracle import BORDER, CLAMP, WRAP, LINEAR, NEAREST, MIN, MAX
from oracle import geometry as G
from oracle import resample as R
from oracle import blend as B
from oracle import raymarch as RM
from oracle import advect as A

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
GRAD_TOL = 1e-4

BNAME = {BORDER: "BORDER", CLAMP: "CLAMP", WRAP: "WRAP"}
FNAME = {NEAREST: "NEAREST", LINEAR: "LINEAR", MIN: "MIN", MAX: "MAX"}
MNAME = {B.BEER_LAMBERT: "BEER_LAMBERT", B.ALPHA: "ALPHA", B.ALPHA_ADDITIVE: "ALPHA_ADDITIVE", B.ADDITIVE: "ADDITIVE"}


@pytest.fixture(scope="module")
def ngt_ops():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from ngt import ops, _lib
    _lib.load()
    return ops


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float32))).to("cuda:0")


def host(t):
    return t.detach().cpu().double().numpy()


def rel(got, exp):
    scale = max(float(np.max(np.abs(exp))), 1e-30)
    return float(np.max(np.abs(got - exp))) / scale


def f32(a):
    """Round to fp32 (the GPU sees fp32 inputs) and return float64."""
    return np.asarray(a, dtype=np.float32).astype(np.float64)


# --- SampleGridLut -------------------------------------------------------------------

@pytest.mark.parametrize("C", [1, 2, 4])
@pytest.mark.parametrize("filt", [LINEAR, NEAREST, MIN, MAX])
@pytest.mark.parametrize("bound", [BORDER, CLAMP, WRAP])
def test_lut_forward(ngt_ops, C, filt, bound):
    rng = np.random.default_rng(100 + C * 10 + filt * 3 + bound)
    inp = f32(rng.standard_normal((2, 5, 6, 7, C)))
    lut = f32(rng.uniform(-2.5, 2.5, (3, 4, 6, 9, 4)))
    lut[..., 3] = 0
    exp = R.sample_grid_lut(inp, lut, filt, bound, 0.0, True, True, False)
    got = host(ngt_ops.sample_grid_lut(dev(inp), dev(lut), interpolation=FNAME[filt], boundary=BNAME[bound]))
    assert got.shape == exp.shape
    assert rel(got, exp) < FWD_TOL
    if bound == BORDER:
        np.testing.assert_array_equal(got == 0, exp == 0)


@pytest.mark.parametrize("separate,relative,normalized", [(False, True, False), (True, False, True), (False, False, False)])
def test_lut_forward_modes(ngt_ops, separate, relative, normalized):
    rng = np.random.default_rng(7)
    inp = f32(rng.standard_normal((2, 6, 5, 4, 2)))
    lut = f32(rng.uniform(0.0, 1.0, (2, 3, 4, 5, 4))) if normalized else f32(rng.uniform(-1, 6, (2, 3, 4, 5, 4)))
    cco = 0.5 if normalized else 0.25
    exp = R.sample_grid_lut(inp, lut, LINEAR, CLAMP, cco, separate, relative, normalized)
    got = host(ngt_ops.sample_grid_lut(dev(inp), dev(lut), boundary="CLAMP", cell_center_offset=cco,
                                       separate_camera_batch=separate, relative_coords=relative,
                                       normalized_coords=normalized))
    assert rel(got, exp) < FWD_TOL


@pytest.mark.parametrize("semantics", ["exact", "reference"])
@pytest.mark.parametrize("filt", [LINEAR, NEAREST, MIN])
@pytest.mark.parametrize("bound", [BORDER, CLAMP, WRAP])
def test_lut_backward(ngt_ops, semantics, filt, bound):
    rng = np.random.default_rng(200 + filt * 3 + bound)
    inp = f32(rng.standard_normal((2, 5, 6, 7, 2)))
    lut = f32(rng.uniform(-1.7, 1.7, (2, 4, 5, 6, 4)))
    lut[..., 3] = 0
    g = f32(rng.standard_normal((2, 2, 4, 5, 6, 2)))
    ei, el = R.sample_grid_lut_grad(inp, lut, g, filt, bound, 0.0, True, True, False, semantics=semantics)
    with ngt_ops.semantics(semantics):
        gi, gl = ngt_ops.sample_grid_lut_grad(dev(inp), dev(g), dev(lut), interpolation=FNAME[filt],
                                              boundary=BNAME[bound])
    assert rel(host(gi), ei) < GRAD_TOL
    if filt == LINEAR:
        assert rel(host(gl), el) < GRAD_TOL
    else:
        assert np.all(host(gl) == 0)


def test_lut_autograd_matches_grad_op(ngt_ops):
    rng = np.random.default_rng(3)
    inp = dev(rng.standard_normal((1, 6, 6, 6, 1))).requires_grad_()
    lut = dev(rng.uniform(-1, 1, (1, 6, 6, 6, 4))).requires_grad_()
    out = ngt_ops.sample_grid_lut(inp, lut, separate_camera_batch=False)
    g = torch.randn_like(out)
    out.backward(g)
    gi, gl = ngt_ops.sample_grid_lut_grad(inp.detach(), g, lut.detach(), separate_camera_batch=False)
    assert torch.allclose(inp.grad, gi, atol=1e-6) and torch.allclose(lut.grad, gl, atol=1e-6)


# --- SampleGridTransform ----------------------------------------------------------

def ring_args(dims, out_dhw, views=3, dist=3.0, generic=False):
    """Ring cameras around a centred unit grid.  generic=True shifts the grid
    by a fraction of a cell so no sample lands exactly on a cell centre (the
    count-normalised reference gradient is discontinuous there: fp32 and fp64
    positions can fall on different sides of an integer)."""
    t = (0.0123, -0.0217, 0.0311) if generic else (0, 0, 0)
    M = G.grid_transform(dims, translation_=t, center=True, scale_=[1.0 / dims[2]] * 3)
    cams = G.look_at_ring_cameras(views, dist, 40.0, out_dhw[2] / out_dhw[1], dist - 1.0, dist + 1.0)
    Ms = [G.op_matrix(M).astype(np.float32)]
    Vs = [G.op_matrix(c[0]).astype(np.float32) for c in cams]
    Ps = [G.op_matrix(c[1]).astype(np.float32) for c in cams]
    Fs = [c[2].astype(np.float32) for c in cams]
    return Ms, Vs, Ps, Fs


@pytest.mark.parametrize("mode", ["TRANSFORM_LINDEPTH", "TRANSFORM", "TRANSFORM_LINDEPTH_REVERSE", "TRANSFORM_REVERSE"])
@pytest.mark.parametrize("C", [1, 4])
@pytest.mark.parametrize("bound", [BORDER, CLAMP])
def test_transform_forward(ngt_ops, mode, C, bound):
    rng = np.random.default_rng(300 + C + bound)
    dims = (10, 9, 11)
    Ms, Vs, Ps, Fs = ring_args(dims, (8, 7, 9))
    if mode.endswith("REVERSE"):
        # sample a frustum grid back into the volume
        inp = f32(rng.standard_normal((1, 8, 7, 9, C)))
        out_dhw = dims
    else:
        inp = f32(rng.standard_normal((1,) + dims + (C,)))
        out_dhw = (8, 7, 9)
    cm = R.COORDINATE_MODE[mode]
    exp = R.sample_grid_transform(inp, Ms, Vs, Ps, Fs, out_dhw, LINEAR, bound, cm)
    got = host(ngt_ops.sample_grid_transform(dev(inp), Ms, Vs, Ps, Fs, out_dhw, boundary=BNAME[bound],
                                             coordinate_mode=mode))
    assert rel(got, exp) < FWD_TOL


@pytest.mark.parametrize("semantics", ["exact", "reference"])
def test_transform_backward(ngt_ops, semantics):
    rng = np.random.default_rng(4)
    dims = (9, 9, 9)
    Ms, Vs, Ps, Fs = ring_args(dims, (7, 6, 8), generic=True)
    inp = f32(rng.standard_normal((1,) + dims + (2,)))
    g = f32(rng.standard_normal((1, 3, 7, 6, 8, 2)))
    exp = R.sample_grid_transform_grad(inp, g, Ms, Vs, Ps, Fs, semantics=semantics)
    with ngt_ops.semantics(semantics):
        got = host(ngt_ops.sample_grid_transform_grad(dev(inp), dev(g), Ms, Vs, Ps, Fs))
    assert rel(got, exp) < GRAD_TOL


# --- ReduceGridBlend ---------------------------------------------------------------

@pytest.mark.parametrize("mode", [B.BEER_LAMBERT, B.ALPHA, B.ALPHA_ADDITIVE, B.ADDITIVE])
@pytest.mark.parametrize("C", [1, 2, 4])
@pytest.mark.parametrize("keep", [False, True])
def test_blend(ngt_ops, mode, C, keep):
    rng = np.random.default_rng(400 + mode * 10 + C + keep)
    x = f32(rng.uniform(0.0, 0.3, (2, 11, 5, 70, C)))
    exp = B.reduce_grid_blend(x, mode, keep)
    xt = dev(x).requires_grad_()
    out = ngt_ops.reduce_grid_blend(xt, MNAME[mode], keep)
    assert rel(host(out), exp) < FWD_TOL
    g = f32(rng.standard_normal(exp.shape))
    out.backward(dev(g))
    eg = B.reduce_grid_blend_grad(g, exp, x, mode, keep, "exact")
    assert rel(host(xt.grad), eg) < GRAD_TOL
    if mode != B.ALPHA_ADDITIVE:
        with ngt_ops.semantics("reference"):
            rg = host(ngt_ops.reduce_grid_blend_grad(dev(g), out.detach(), dev(x), MNAME[mode], keep))
        assert rel(rg, B.reduce_grid_blend_grad(g, exp, x, mode, keep, "reference")) < GRAD_TOL


# --- RaymarchGridTransform ---------------------------------------------------------

@pytest.mark.parametrize("mode", [B.BEER_LAMBERT, B.ALPHA, B.ALPHA_ADDITIVE, B.ADDITIVE])
@pytest.mark.parametrize("C", [1, 4])
@pytest.mark.parametrize("bound", [BORDER, CLAMP, WRAP])
def test_raymarch(ngt_ops, mode, C, bound):
    rng = np.random.default_rng(500 + mode * 10 + C + bound)
    dims = (12, 11, 13)
    out_dhw = (14, 9, 12)
    Ms, Vs, Ps, Fs = ring_args(dims, out_dhw)
    x = f32(rng.uniform(0.0, 0.25, (1,) + dims + (C,)))
    exp = RM.raymarch_grid_transform(x, Ms, Vs, Ps, Fs, out_dhw, LINEAR, bound, mode)
    xt = dev(x).requires_grad_()
    out = ngt_ops.raymarch_grid_transform(xt, Ms, Vs, Ps, Fs, out_dhw, boundary=BNAME[bound],
                                          blending_mode=MNAME[mode])
    assert rel(host(out), exp) < FWD_TOL
    g = f32(rng.standard_normal(exp.shape))
    out.backward(dev(g))
    eg = RM.raymarch_grid_transform_grad(x, exp, g, Ms, Vs, Ps, Fs, out_dhw, LINEAR, bound, mode)
    assert rel(host(xt.grad), eg) < GRAD_TOL


@pytest.mark.parametrize("mode", [B.BEER_LAMBERT, B.ADDITIVE])
@pytest.mark.parametrize("bound", [BORDER, CLAMP])
def test_raymarch_reference_gradient(ngt_ops, mode, bound):
    rng = np.random.default_rng(600 + mode + bound)
    dims = (10, 10, 10)
    out_dhw = (12, 8, 9)
    Ms, Vs, Ps, Fs = ring_args(dims, out_dhw, views=2, generic=True)
    x = f32(rng.uniform(0.0, 0.25, (1,) + dims + (4,)))
    exp = RM.raymarch_grid_transform(x, Ms, Vs, Ps, Fs, out_dhw, LINEAR, bound, mode)
    g = f32(rng.standard_normal(exp.shape))
    eg = RM.raymarch_grid_transform_grad(x, exp, g, Ms, Vs, Ps, Fs, out_dhw, LINEAR, bound, mode, semantics="reference")
    with ngt_ops.semantics("reference"):
        got = host(ngt_ops.raymarch_grid_transform_grad(dev(x), dev(exp), dev(g), Ms, Vs, Ps, Fs, out_dhw,
                                                        boundary=BNAME[bound], blending_mode=MNAME[mode]))
    assert rel(got, eg) < GRAD_TOL


def test_render_density_fused_matches_concatenated_raymarch(ngt_ops):
    rng = np.random.default_rng(7)
    dims = (12, 12, 12)
    out_dhw = (16, 10, 12)
    Ms, Vs, Ps, Fs = ring_args(dims, out_dhw)
    rho = f32(rng.uniform(0.0, 0.5, (1,) + dims + (1,)))
    light = np.array([0.9, 0.7, 0.5])
    ds = 0.125
    cat = np.concatenate([rho * light, rho], axis=-1) * ds
    exp = RM.raymarch_grid_transform(cat, Ms, Vs, Ps, Fs, out_dhw, LINEAR, BORDER, B.BEER_LAMBERT)
    rt = dev(rho).requires_grad_()
    out = ngt_ops.render_density_fused(rt, light, ds, Ms, Vs, Ps, Fs, out_dhw)
    assert rel(host(out), exp) < FWD_TOL
    g = f32(rng.standard_normal(exp.shape))
    out.backward(dev(g))
    eg4 = RM.raymarch_grid_transform_grad(cat, exp, g, Ms, Vs, Ps, Fs, out_dhw, LINEAR, BORDER, B.BEER_LAMBERT)
    eg = ds * (eg4[..., :3] @ light + eg4[..., 3])[..., None]
    assert rel(host(rt.grad), eg) < GRAD_TOL


# --- advection -----------------------------------------------------------------------

def smooth_field(rng, shape, amp):
    from scipy.ndimage import gaussian_filter
    f = gaussian_filter(rng.standard_normal(shape), 1.5)
    return amp * f / np.max(np.abs(f))


@pytest.mark.parametrize("order,clamp", [(1, "NONE"), (2, "NONE"), (2, "MC_SMOOTH"), (1, "NEGATIVE"), (2, "NEGATIVE")])
@pytest.mark.parametrize("bound", [BORDER, CLAMP, WRAP])
@pytest.mark.parametrize("staggered", [False, True])
def test_advect(ngt_ops, order, clamp, bound, staggered):
    rng = np.random.default_rng(700 + order + bound + 3 * staggered)
    D, H, W = 10, 11, 12
    rho = f32(np.clip(smooth_field(rng, (D, H, W), 1.0), -0.2, 1.0))[..., None]
    if staggered:
        vx = f32(smooth_field(rng, (D, H, W + 1), 1.8))[..., None]
        vy = f32(smooth_field(rng, (D, H + 1, W), 1.8))[..., None]
        vz = f32(smooth_field(rng, (D + 1, H, W), 1.8))[..., None]
        vc = A.centered_from_staggered(vx, vy, vz)
        vel = (dev(vx[None]).requires_grad_(), dev(vy[None]).requires_grad_(), dev(vz[None]).requires_grad_())
    else:
        vc = f32(np.stack([smooth_field(rng, (D, H, W), 1.8) for _ in range(3)], -1))
        vel = dev(vc[None]).requires_grad_()
    exp = A.warp_forward(rho, vc, 1.0, order, clamp, bound)
    rt = dev(rho[None]).requires_grad_()
    out = ngt_ops.advect_grid_semi_langrange(rt, vel, timestep=-1.0, order=order, clamp_extrema=clamp,
                                             boundary=BNAME[bound], separate_velocity_batch=False)
    got = host(out)[0, 0]
    assert rel(got, exp) < FWD_TOL
    g = f32(rng.standard_normal(exp.shape))
    out.backward(dev(g[None, None]))
    eg_rho, eg_v = A.warp_backward(rho, vc, g, 1.0, order, clamp, bound)
    assert rel(host(rt.grad)[0], eg_rho) < GRAD_TOL
    if staggered:
        egx, egy, egz = A.centered_from_staggered_adjoint(eg_v)
        for t, e in zip(vel, (egx, egy, egz)):
            assert rel(host(t.grad)[0], e) < GRAD_TOL
    else:
        assert rel(host(vel.grad)[0], eg_v) < GRAD_TOL


def test_advect_reference_gradient(ngt_ops):
    rng = np.random.default_rng(9)
    D, H, W = 9, 10, 11
    rho = f32(np.clip(smooth_field(rng, (D, H, W), 1.0), 0, 1))[..., None]
    vc = f32(np.stack([smooth_field(rng, (D, H, W), 1.5) for _ in range(3)], -1))
    g = f32(rng.standard_normal((D, H, W, 1)))
    eg_rho, eg_v = A.warp_backward(rho, vc, g, 1.0, 2, "MC_SMOOTH", BORDER, semantics="reference")
    rt = dev(rho[None]).requires_grad_()
    vt = dev(vc[None]).requires_grad_()
    with ngt_ops.semantics("reference"):
        out = ngt_ops.advect_grid_semi_langrange(rt, vt, -1.0, 2, "MC_SMOOTH", "BORDER", False)
        out.backward(dev(g[None, None]))
    assert rel(host(rt.grad)[0], eg_rho) < GRAD_TOL
    assert rel(host(vt.grad)[0], eg_v) < GRAD_TOL


def test_error_messages(ngt_ops):
    x = torch.zeros((1, 4, 4, 4, 3), device="cuda:0")
    lut = torch.zeros((1, 4, 4, 4, 4), device="cuda:0")
    with pytest.raises(NotImplementedError, match="Only 1,2 and 4 Channel supported."):
        ngt_ops.sample_grid_lut(x, lut)
    with pytest.raises(ValueError, match="camera batch must match data batch"):
        ngt_ops.sample_grid_lut(torch.zeros((2, 4, 4, 4, 1), device="cuda:0"), lut, separate_camera_batch=False)


@pytest.mark.parametrize("vmax,C,separate", [(5.0, 1, False), (1.5, 2, True), (4.0, 2, True), (1.5, 4, False)])
def test_advect_overflow_channels_batches(ngt_ops, vmax, C, separate):
    """Displacements beyond the backward's LDS tile halo (vmax 4-5 cells),
    2 / 4 channels, and separate velocity batches where several data entries
    share a velocity (N=2 x V=2 pairs)."""
    rng = np.random.default_rng(int(vmax * 10) + C)
    D, H, W = 9, 17, 37
    N, V = (2, 2) if separate else (1, 1)
    rho = f32(np.stack([np.stack([np.clip(smooth_field(rng, (D, H, W), 1.0), 0, 1) for _ in range(C)], -1)
                        for _ in range(N)]))
    vx = f32(np.stack([smooth_field(rng, (D, H, W + 1), vmax) for _ in range(V)]))[..., None]
    vy = f32(np.stack([smooth_field(rng, (D, H + 1, W), vmax) for _ in range(V)]))[..., None]
    vz = f32(np.stack([smooth_field(rng, (D + 1, H, W), vmax) for _ in range(V)]))[..., None]
    vel = tuple(dev(c).requires_grad_() for c in (vx, vy, vz))
    rt = dev(rho).requires_grad_()
    out = ngt_ops.advect_grid_semi_langrange(rt, vel, timestep=-1.0, order=2, clamp_extrema="MC_SMOOTH",
                                             boundary="BORDER", separate_velocity_batch=separate)
    pairs = [(n, v, v) for n in range(N) for v in range(V)] if separate else [(n, n, 0) for n in range(N)]
    g = f32(rng.standard_normal(out.shape))
    out.backward(dev(g))
    eg_rho = np.zeros_like(rho)
    eg_v = [np.zeros_like(c) for c in (vx, vy, vz)]
    for n, v, slot in pairs:
        vc = A.centered_from_staggered(vx[v], vy[v], vz[v])
        exp = A.warp_forward(rho[n], vc, 1.0, 2, "MC_SMOOTH", BORDER)
        got = host(out)[n, slot]
        assert rel(got, exp) < FWD_TOL
        gr, gv = A.warp_backward(rho[n], vc, g[n, slot], 1.0, 2, "MC_SMOOTH", BORDER)
        eg_rho[n] += gr
        for k, e in enumerate(A.centered_from_staggered_adjoint(gv)):
            eg_v[k][v] += e
    assert rel(host(rt.grad), eg_rho) < GRAD_TOL
    for t, e in zip(vel, eg_v):
        assert rel(host(t.grad), e) < GRAD_TOL


@pytest.mark.parametrize("C,bound,mode", [(1, BORDER, B.BEER_LAMBERT), (2, CLAMP, B.ALPHA_ADDITIVE),
                                          (4, BORDER, B.ALPHA), (4, CLAMP, B.ADDITIVE), (1, WRAP, B.BEER_LAMBERT),
                                          (2, WRAP, B.ADDITIVE)])
def test_raymarch_bricked_layout_is_bit_identical(ngt_ops, C, bound, mode):
    """The forward over the haloed 4x4x4-brick copy (dims not multiples of 4;
    halo cells hold the boundary mode's values) equals the direct NDHWC read
    bit for bit, for the generic raymarch and the fused density renderer."""
    rng = np.random.default_rng(800 + C)
    dims = (13, 10, 11)
    out_dhw = (17, 9, 12)
    Ms, Vs, Ps, Fs = ring_args(dims, out_dhw, views=3, generic=True)
    x = dev(rng.uniform(0.0, 0.3, (1,) + dims + (C,)))
    old = ngt_ops.use_bricked_layout
    try:
        outs = []
        for bricked in (True, False):
            ngt_ops.use_bricked_layout = bricked
            outs.append(ngt_ops.raymarch_grid_transform(x, Ms, Vs, Ps, Fs, out_dhw, boundary=BNAME[bound],
                                                        blending_mode=MNAME[mode]))
        assert torch.equal(outs[0], outs[1])
        rho = dev(rng.uniform(0.0, 0.5, (2,) + dims + (1,)))
        Ms2 = Ms * 2
        imgs = []
        for bricked in (True, False):
            ngt_ops.use_bricked_layout = bricked
            imgs.append(ngt_ops.render_density_fused(rho, (0.9, 0.8, 0.7), 0.1, Ms2, Vs, Ps, Fs, out_dhw,
                                                     boundary=BNAME[bound]))
        assert torch.equal(imgs[0], imgs[1])
    finally:
        ngt_ops.use_bricked_layout = old


@pytest.mark.parametrize("clamp", ["NONE", "MC_SMOOTH"])
def test_advect_three_channels_self_advection(ngt_ops, clamp):
    """3-channel input (the centered velocity) advected by itself: the input
    and velocity gradients of the same tensor add up."""
    rng = np.random.default_rng(900 + len(clamp))
    D, H, W = 9, 11, 10
    vc = f32(np.stack([smooth_field(rng, (D, H, W), 1.6) for _ in range(3)], -1))
    vt = dev(vc[None]).requires_grad_()
    out = ngt_ops.advect_grid_semi_langrange(vt, vt, timestep=-1.0, order=2, clamp_extrema=clamp,
                                             boundary="BORDER", separate_velocity_batch=False)
    exp = A.warp_forward(vc, vc, 1.0, 2, clamp, BORDER)
    assert rel(host(out)[0, 0], exp) < FWD_TOL
    g = f32(rng.standard_normal(exp.shape))
    out.backward(dev(g[None, None]))
    eg_in, eg_v = A.warp_backward(vc, vc, g, 1.0, 2, clamp, BORDER)
    assert rel(host(vt.grad)[0], eg_in + eg_v) < GRAD_TOL


@pytest.mark.parametrize("shape,C,order,clamp,staggered", [((5, 6, 7), 1, 2, "MC_SMOOTH", True),
                                                            ((4, 9, 13), 2, 2, "MC_SMOOTH", False),
                                                            ((6, 5, 20), 1, 1, "NONE", False),
                                                            ((3, 7, 33), 2, 2, "NEGATIVE", True)])
def test_advect_wrap_grids_narrower_than_a_tile(ngt_ops, shape, C, order, clamp, staggered):
    """WRAP on grids narrower than the backward's LDS tile: wrapped corners
    land inside the tile at slot steps other than 0 / 1, and neighbouring
    work-items' corners coincide in wrapped and unwrapped arrangements."""
    rng = np.random.default_rng(sum(shape) + 10 * C + order)
    D, H, W = shape
    rho = f32(np.stack([np.clip(smooth_field(rng, shape, 1.0), -0.2, 1.0) for _ in range(C)], -1))
    if staggered:
        vx = f32(rng.uniform(-1.9, 1.9, (D, H, W + 1)))[..., None]
        vy = f32(rng.uniform(-1.9, 1.9, (D, H + 1, W)))[..., None]
        vz = f32(rng.uniform(-1.9, 1.9, (D + 1, H, W)))[..., None]
        vc = A.centered_from_staggered(vx, vy, vz)
        vel = (dev(vx[None]).requires_grad_(), dev(vy[None]).requires_grad_(), dev(vz[None]).requires_grad_())
    else:
        vc = f32(rng.uniform(-1.9, 1.9, shape + (3,)))
        vel = dev(vc[None]).requires_grad_()
    exp = A.warp_forward(rho, vc, 1.0, order, clamp, WRAP)
    rt = dev(rho[None]).requires_grad_()
    out = ngt_ops.advect_grid_semi_langrange(rt, vel, timestep=-1.0, order=order, clamp_extrema=clamp,
                                             boundary="WRAP", separate_velocity_batch=False)
    assert rel(host(out)[0, 0], exp) < FWD_TOL
    g = f32(rng.standard_normal(exp.shape))
    out.backward(dev(g[None, None]))
    eg_rho, eg_v = A.warp_backward(rho, vc, g, 1.0, order, clamp, WRAP)
    assert rel(host(rt.grad)[0], eg_rho) < GRAD_TOL
    if staggered:
        for t, e in zip(vel, A.centered_from_staggered_adjoint(eg_v)):
            assert rel(host(t.grad)[0], e) < GRAD_TOL
    else:
        assert rel(host(vel.grad)[0], eg_v) < GRAD_TOL
