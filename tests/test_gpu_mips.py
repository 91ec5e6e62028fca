"""GPU parity of the mip-mapped sampling paths (SampleGridLut /
SampleGridTransform with mipmapping NEAREST / LINEAR, their exact adjoints,
LodTransform, the Renderer's lookup-table and default mip-mapped paths)
against the FP64 oracle on the same seeded inputs.

Tolerances: forward 1e-5 and gradients 1e-4, normwise relative
(max |gpu - oracle| / max |oracle|), as for the other ops.  LoD values of
LodTransform: 1e-4 absolute.  With NEAREST mip selection the level is a step
function of the LoD; output cells whose oracle LoD lies within 2e-3 of a
rounding point (fp32 and fp64 LoDs could pick different levels there) get a
zero output gradient and are left out of the forward comparison.
"""
import numpy as np
import pytest
import torch

from oracle import BORDER, CLAMP, WRAP, LINEAR, NEAREST, MIN, MAX
from oracle import geometry as G
from oracle import mips as MP
from oracle import resample as R
from oracle import blend as B

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
GRAD_TOL = 1e-4

BNAME = {BORDER: "BORDER", CLAMP: "CLAMP", WRAP: "WRAP"}
FNAME = {NEAREST: "NEAREST", LINEAR: "LINEAR", MIN: "MIN", MAX: "MAX"}
MIPNAME = {MP.MIP_NEAREST: "NEAREST", MP.MIP_LINEAR: "LINEAR"}


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
    return np.asarray(a, dtype=np.float32).astype(np.float64)


# --- SampleGridLut with mips ---------------------------------------------------------

def _lut_case(rng, C, dims=(9, 10, 11), L=3):
    inp = f32(rng.standard_normal((2,) + dims + (C,)))
    pos = rng.uniform(-1.0, 1.0, (3, 4, 6, 9, 3)) * (np.array(dims[::-1]) + 2.0) * 0.5 + np.array(dims[::-1]) * 0.5
    lod = rng.uniform(-0.6, L + 0.4, (3, 4, 6, 9, 1))
    return inp, f32(np.concatenate([pos, lod], -1))


@pytest.mark.parametrize("C", [1, 4])
@pytest.mark.parametrize("filt", [LINEAR, NEAREST, MIN, MAX])
@pytest.mark.parametrize("bound", [BORDER, CLAMP, WRAP])
@pytest.mark.parametrize("mip", [MP.MIP_NEAREST, MP.MIP_LINEAR])
def test_lut_mip_forward(ngt_ops, C, filt, bound, mip):
    rng = np.random.default_rng(500 + C * 17 + filt * 5 + bound * 3 + mip)
    inp, lut = _lut_case(rng, C)
    exp = R.sample_grid_lut(inp, lut, filt, bound, 0.5, True, False, False, mip, 3, 0.25)
    got = host(ngt_ops.sample_grid_lut(dev(inp), dev(lut), interpolation=FNAME[filt], boundary=BNAME[bound],
                                       mipmapping=MIPNAME[mip], num_mipmaps=3, mip_bias=0.25, cell_center_offset=0.5,
                                       relative_coords=False))
    assert got.shape == exp.shape
    assert rel(got, exp) < FWD_TOL


def test_lut_mip_forward_relative_normalized(ngt_ops):
    rng = np.random.default_rng(77)
    inp = f32(rng.standard_normal((1, 12, 10, 8, 2)))
    lut = f32(np.concatenate([rng.uniform(-0.3, 0.3, (2, 5, 4, 6, 3)), rng.uniform(0, 2.2, (2, 5, 4, 6, 1))], -1))
    exp = R.sample_grid_lut(inp, lut, LINEAR, CLAMP, 0.5, True, True, True, MP.MIP_LINEAR, 2, -0.1)
    got = host(ngt_ops.sample_grid_lut(dev(inp), dev(lut), boundary="CLAMP", mipmapping="LINEAR", num_mipmaps=2,
                                       mip_bias=-0.1, cell_center_offset=0.5, relative_coords=True,
                                       normalized_coords=True))
    assert rel(got, exp) < FWD_TOL


@pytest.mark.parametrize("filt", [LINEAR, NEAREST, MIN])
@pytest.mark.parametrize("bound", [BORDER, CLAMP, WRAP])
@pytest.mark.parametrize("mip", [MP.MIP_NEAREST, MP.MIP_LINEAR])
def test_lut_mip_backward_exact(ngt_ops, filt, bound, mip):
    rng = np.random.default_rng(600 + filt * 5 + bound * 3 + mip)
    inp, lut = _lut_case(rng, 2)
    g = f32(rng.standard_normal((2, 3, 4, 6, 9, 2)))
    ei, el = R.sample_grid_lut_grad(inp, lut, g, filt, bound, 0.5, True, False, False, "exact", mip, 3, 0.25)
    gi, gl = ngt_ops.sample_grid_lut_grad(dev(inp), dev(g), dev(lut), interpolation=FNAME[filt],
                                          boundary=BNAME[bound], mipmapping=MIPNAME[mip], num_mipmaps=3,
                                          mip_bias=0.25, cell_center_offset=0.5, relative_coords=False)
    assert rel(host(gi), ei) < GRAD_TOL
    gl = host(gl)
    if filt == LINEAR:
        assert rel(gl, el) < GRAD_TOL
    else:
        assert np.all(gl[..., :3] == 0)
    if mip == MP.MIP_LINEAR:
        assert rel(gl[..., 3], el[..., 3]) < GRAD_TOL
    else:
        assert np.all(gl[..., 3] == 0)


def test_lut_mip_autograd_and_reference_rejects(ngt_ops):
    rng = np.random.default_rng(9)
    inp, lut = _lut_case(rng, 1)
    it, lt = dev(inp).requires_grad_(), dev(lut).requires_grad_()
    kw = dict(mipmapping="LINEAR", num_mipmaps=3, mip_bias=0.25, cell_center_offset=0.5, relative_coords=False)
    out = ngt_ops.sample_grid_lut(it, lt, **kw)
    g = torch.randn_like(out)
    out.backward(g)
    gi, gl = ngt_ops.sample_grid_lut_grad(it.detach(), g, lt.detach(), **kw)
    assert torch.allclose(it.grad, gi, atol=1e-6) and torch.allclose(lt.grad, gl, atol=1e-6)
    with ngt_ops.semantics("reference"):
        with pytest.raises(NotImplementedError, match="mipmapping 'NONE' only"):
            ngt_ops.sample_grid_lut_grad(it.detach(), g, lt.detach(), **kw)


# --- SampleGridTransform with mips ---------------------------------------------------

def ring_args(dims, out_dhw, views=3, dist=3.0, fov=40.0):
    M = G.grid_transform(dims, translation_=(0.0123, -0.0217, 0.0311), center=True, scale_=[1.0 / dims[2]] * 3)
    cams = G.look_at_ring_cameras(views, dist, fov, out_dhw[2] / out_dhw[1], dist - 1.0, dist + 1.0)
    Ms = [G.op_matrix(M).astype(np.float32)]
    Vs = [G.op_matrix(c[0]).astype(np.float32) for c in cams]
    Ps = [G.op_matrix(c[1]).astype(np.float32) for c in cams]
    Fs = [c[2].astype(np.float32) for c in cams]
    return Ms, Vs, Ps, Fs


def _xform_case(mode):
    """(input dims, output dims, camera args, mip bias): forward modes sample a
    16x14x18 volume into a coarse 5x4x6 frustum grid, reverse modes a 16x12x16
    frustum grid into a 5x5x6 volume; biased LoDs lie between 0.3 and 2.3."""
    if mode.endswith("REVERSE"):
        in_dhw, out_dhw = (16, 12, 16), (5, 5, 6)
        return in_dhw, out_dhw, ring_args(out_dhw, in_dhw), 0.8
    in_dhw, out_dhw = (16, 14, 18), (5, 4, 6)
    return in_dhw, out_dhw, ring_args(in_dhw, out_dhw, fov=25.0), -1.6


def _nearest_safe(cams, in_dhw, out_dhw, cm, bias, L):
    """[N, V, D, H, W] mask of cells whose LoD is not within 2e-3 of a NEAREST
    rounding point."""
    Ms, Vs, Ps, Fs = cams
    safe = np.ones((1, len(Vs)) + tuple(out_dhw), dtype=bool)
    for v in range(len(Vs)):
        cx = G.CameraXform.from_op_args(Ms[0], Vs[v], Ps[v], Fs[v])
        lod = np.clip(R.transform_lod(cx, out_dhw, in_dhw, cm) + bias, 0.0, L)
        safe[0, v] = np.abs((lod - np.floor(lod)) - 0.5) > 2e-3
    return safe


@pytest.mark.parametrize("mode", ["TRANSFORM_LINDEPTH", "TRANSFORM", "TRANSFORM_LINDEPTH_REVERSE", "TRANSFORM_REVERSE"])
@pytest.mark.parametrize("C", [1, 4])
@pytest.mark.parametrize("bound", [BORDER, CLAMP])
@pytest.mark.parametrize("mip", [MP.MIP_NEAREST, MP.MIP_LINEAR])
def test_transform_mip_forward(ngt_ops, mode, C, bound, mip):
    rng = np.random.default_rng(700 + C + bound * 3 + mip * 7)
    in_dhw, out_dhw, cams, bias = _xform_case(mode)
    Ms, Vs, Ps, Fs = cams
    inp = f32(rng.standard_normal((1,) + in_dhw + (C,)))
    cm = R.COORDINATE_MODE[mode]
    exp = R.sample_grid_transform(inp, Ms, Vs, Ps, Fs, out_dhw, LINEAR, bound, cm, True, mip, 3, bias)
    got = host(ngt_ops.sample_grid_transform(dev(inp), Ms, Vs, Ps, Fs, out_dhw, boundary=BNAME[bound],
                                             coordinate_mode=mode, mipmapping=MIPNAME[mip], num_mipmaps=3,
                                             mip_bias=bias))
    if mip == MP.MIP_NEAREST:
        safe = _nearest_safe(cams, in_dhw, out_dhw, cm, bias, 3)[..., None]
        got, exp = np.where(safe, got, 0.0), np.where(safe, exp, 0.0)
    assert rel(got, exp) < FWD_TOL


@pytest.mark.parametrize("mode", ["TRANSFORM_LINDEPTH", "TRANSFORM_LINDEPTH_REVERSE"])
@pytest.mark.parametrize("mip", [MP.MIP_NEAREST, MP.MIP_LINEAR])
@pytest.mark.parametrize("bound", [BORDER, CLAMP])
def test_transform_mip_backward_exact(ngt_ops, mode, mip, bound):
    rng = np.random.default_rng(800 + mip + bound * 3)
    in_dhw, out_dhw, cams, bias = _xform_case(mode)
    Ms, Vs, Ps, Fs = cams
    cm = R.COORDINATE_MODE[mode]
    inp = f32(rng.standard_normal((1,) + in_dhw + (2,)))
    g = f32(rng.standard_normal((1, 3) + out_dhw + (2,)))
    if mip == MP.MIP_NEAREST:
        g = g * _nearest_safe(cams, in_dhw, out_dhw, cm, bias, 3)[..., None]
    exp = R.sample_grid_transform_grad(inp, g, Ms, Vs, Ps, Fs, LINEAR, bound, cm, True, "exact", mip, 3, bias)
    it = dev(inp).requires_grad_()
    out = ngt_ops.sample_grid_transform(it, Ms, Vs, Ps, Fs, out_dhw, boundary=BNAME[bound], coordinate_mode=mode,
                                        mipmapping=MIPNAME[mip], num_mipmaps=3, mip_bias=bias)
    out.backward(dev(g))
    assert rel(host(it.grad), exp) < GRAD_TOL
    with ngt_ops.semantics("reference"):
        with pytest.raises(NotImplementedError, match="mipmapping 'NONE' only"):
            ngt_ops.sample_grid_transform_grad(it.detach(), dev(g), Ms, Vs, Ps, Fs, boundary=BNAME[bound],
                                               coordinate_mode=mode, mipmapping=MIPNAME[mip], num_mipmaps=3,
                                               mip_bias=bias)


def test_mip_pyramid_of_large_grid(ngt_ops):
    """A 128^3 4-channel volume through 4 levels: a field linear in the
    position is reproduced at every level (closed form, no oracle pass over
    the whole grid)."""
    D = 128
    z, y, x = torch.meshgrid(*(torch.arange(D, dtype=torch.float32, device="cuda:0") + 0.5 for _ in range(3)),
                             indexing="ij")
    a = (0.03, -0.02, 0.01)
    f = (a[0] * x + a[1] * y + a[2] * z + 1.0)[..., None]
    inp = torch.cat([f, 2 * f, -f, f + 3], -1)[None].contiguous()
    rng = np.random.default_rng(5)
    pos = rng.uniform(12.0, 116.0, (1, 8, 16, 32, 3))
    lod = rng.uniform(0.0, 4.0, (1, 8, 16, 32, 1))
    lut = dev(np.concatenate([pos, lod], -1))
    got = host(ngt_ops.sample_grid_lut(inp, lut, mipmapping="LINEAR", num_mipmaps=4, cell_center_offset=0.5,
                                       relative_coords=False))[0, 0]
    p = f32(pos[0])
    e = p[..., 0] * a[0] + p[..., 1] * a[1] + p[..., 2] * a[2] + 1.0
    exp = np.stack([e, 2 * e, -e, e + 3], -1)
    assert rel(got, exp) < FWD_TOL


# --- LodTransform --------------------------------------------------------------------

@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("lin", [True, False])
@pytest.mark.parametrize("relative", [False, True])
def test_lod_transform(ngt_ops, inverse, lin, relative):
    vol, fr = (12, 10, 14), ((36, 32, 44) if inverse else (9, 8, 11))
    M = G.grid_transform(vol, center=True, scale_=[1.0 / vol[2]] * 3)
    cams = G.look_at_ring_cameras(2, 3.0, 40.0, fr[2] / fr[1], 2.0, 4.0)
    MV = [G.op_matrix(c[0] @ M).astype(np.float32) for c in cams]
    P = [G.op_matrix(c[1]).astype(np.float32) for c in cams]
    F = [c[2].astype(np.float32) for c in cams]
    in_dhw, out_dhw = (fr, vol) if inverse else (vol, fr)
    exp = R.lod_transform(in_dhw, MV, P, F, out_dhw, inverse, lin, relative)
    got = host(ngt_ops.lod_transform(in_dhw, MV, P, F, out_dhw, inverse_transform=inverse, linearize_depth=lin,
                                     relative_coords=relative))
    assert got.shape == exp.shape
    assert rel(got[..., :3], exp[..., :3]) < FWD_TOL
    assert np.max(np.abs(got[..., 3] - exp[..., 3])) < 1e-4
    assert np.max(exp[..., 3]) > 0.3  # the LoD is exercised


# --- Renderer -------------------------------------------------------------------------

def _renderer_scene(static):
    from ngt import synth
    grid, cams = synth.ring_setup((16, 16, 16), (6, 8), 5, 2)
    grid.set_translation([0.0123, -0.0217, 0.0311])
    if static:
        for c in cams:
            c.static = grid
    return grid, cams


def _cam_mats(grid, cams):
    from ngt.renderer import Renderer
    return [np.asarray(a, np.float32) for a in Renderer()._transform_params([grid], cams)]


def test_renderer_static_camera_lut_path(ngt_ops):
    """Static cameras sample through LodTransform lookup tables (cell-centre
    offset 0, the Renderer's default mipmapping LINEAR with 2 levels)."""
    from ngt.renderer import Renderer
    rng = np.random.default_rng(11)
    grid, cams = _renderer_scene(True)
    x = f32(rng.standard_normal((1, 16, 16, 16, 2)))
    r = Renderer()
    xt = dev(x).requires_grad_()
    y = r.sample_camera(xt, grid, cams, squeeze_batch=False)
    assert all(c.LuT is not None for c in cams)
    M, V, P, F = _cam_mats(grid, cams)
    MV = [G.op_matrix(G.math_matrix(v) @ G.math_matrix(M[0])) for v in V]
    luts = R.lod_transform((16, 16, 16), MV, P, F, cams[0].transform.grid_size)
    exp = R.sample_grid_lut(x, luts, LINEAR, BORDER, 0.0, True, False, False, MP.MIP_LINEAR, 2, 0.0)
    assert rel(host(y), exp) < FWD_TOL
    g = f32(rng.standard_normal(exp.shape))
    y.backward(dev(g))
    eg, _ = R.sample_grid_lut_grad(x, luts, g, LINEAR, BORDER, 0.0, True, False, False, "exact", MP.MIP_LINEAR, 2,
                                   0.0)
    assert rel(host(xt.grad), eg) < GRAD_TOL


def test_renderer_default_mips_frustum_render(ngt_ops):
    """render_density through the frustum grid (not fused) with the
    Renderer's default mipmapping (LINEAR, 2 levels)."""
    from ngt.renderer import Renderer
    from ngt.lighting import Light
    rng = np.random.default_rng(12)
    grid, cams = _renderer_scene(False)
    rho = f32(rng.uniform(0.0, 0.6, (1, 16, 16, 16, 1)))
    light = Light(color=(0.9, 0.8, 0.7), intensity=1.2)
    r = Renderer(fused=False)
    rt = dev(rho).requires_grad_()
    grid.data = rt
    imgs = r.render_density(grid, [light], cams, cut_alpha=False)
    got = np.stack([host(i) for i in imgs], 1)
    Ms, Vs, Ps, Fs = _cam_mats(grid, cams)
    ds = float(cams[0].depth_step)
    cat = np.concatenate([rho * light.rgb(), rho], -1) * ds
    out_dhw = tuple(cams[0].transform.grid_size)
    fr = R.sample_grid_transform(cat, Ms, Vs, Ps, Fs, out_dhw, LINEAR, BORDER, R.TRANSFORM_LINDEPTH, True,
                                 MP.MIP_LINEAR, 2, 0.0)
    exp = np.stack([B.reduce_grid_blend(fr[:, v], B.BEER_LAMBERT) for v in range(len(cams))], 1)
    assert rel(got, exp) < FWD_TOL
    g = f32(rng.standard_normal(got.shape))
    sum((imgs[v] * dev(g[:, v])).sum() for v in range(len(cams))).backward()
    g_cat = np.zeros_like(cat)
    for v in range(len(cams)):
        gf = B.reduce_grid_blend_grad(g[:, v], exp[:, v], fr[:, v], B.BEER_LAMBERT)
        g_cat += R.sample_grid_transform_grad(cat, gf[:, None], Ms, Vs[v:v + 1], Ps[v:v + 1], Fs[v:v + 1], LINEAR,
                                              BORDER, R.TRANSFORM_LINDEPTH, True, "exact", MP.MIP_LINEAR, 2, 0.0)
    eg = ds * (g_cat[..., :3] @ light.rgb() + g_cat[..., 3])[..., None]
    assert rel(host(rt.grad), eg) < GRAD_TOL
