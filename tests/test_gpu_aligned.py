"""GPU parity of the aligned resampling between resolutions
(Renderer._sample_transform / resample_grid3D_aligned, DensityGrid.scaled,
VelocityGrid.scaled / resample_velocity / _centered_to_staggered /
_scalar_centered_to_staggered / _sampled_to_shape) against the FP64 oracle.

Tolerances: forward 1e-5, gradients 1e-4, normwise relative, as for the ops.
"""
import numpy as np
import pytest
import torch

from oracle import BORDER, CLAMP, LINEAR
from oracle import geometry as G
from oracle import resample as R

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


def _renderer(boundary="BORDER", sample_gradients=False):
    from ngt.renderer import Renderer
    return Renderer(filter_mode="LINEAR", boundary_mode=boundary, mipmapping="NONE", blend_mode="ADDITIVE",
                    sample_gradients=sample_gradients)


ALIGNS = [("BORDER", "BORDER", "BORDER"), ("CENTER", "BORDER", "CENTER"), ("STAGGER_OUTPUT", "CENTER", "BORDER"),
          ("BORDER", "STAGGER_INPUT", "STAGGER_OUTPUT")]


@pytest.mark.parametrize("align", ALIGNS)
@pytest.mark.parametrize("C", [1, 2, 3, 4])
@pytest.mark.parametrize("bound", [BORDER, CLAMP])
def test_resample_aligned(ngt, align, C, bound):
    rng = np.random.default_rng(900 + C * 7 + bound + ALIGNS.index(align) * 3)
    x = f32(rng.standard_normal((2, 7, 9, 8, C)))
    out_dhw = (11, 6, 13)
    r = _renderer("BORDER" if bound == BORDER else "CLAMP")
    xt = dev(x).requires_grad_()
    y = r.resample_grid3D_aligned(xt, out_dhw, *align, allow_split_channels=(C == 3))
    exp = R.resample_grid3D_aligned(x, out_dhw, align, LINEAR, bound)
    assert tuple(y.shape) == exp.shape
    assert rel(host(y), exp) < FWD_TOL
    g = f32(rng.standard_normal(exp.shape))
    y.backward(dev(g))
    eg = R.resample_grid3D_aligned_grad(x.shape, g, align, LINEAR, bound)
    assert rel(host(xt.grad), eg) < GRAD_TOL


def test_resample_aligned_rejects_three_channels(ngt):
    with pytest.raises(ValueError):
        _renderer().resample_grid3D_aligned(torch.zeros((1, 4, 4, 4, 3), device="cuda:0"), (5, 5, 5))


def test_sample_transform_gather_gradient(ngt):
    """sample_gradients=True: the backward of _sample_transform is the
    TRANSFORM_REVERSE read of dy (renderer.py:239-268)."""
    rng = np.random.default_rng(5)
    x = f32(rng.standard_normal((1, 7, 9, 8, 2)))
    out_dhw = (11, 6, 13)
    r = _renderer(sample_gradients=True)
    xt = dev(x).requires_grad_()
    y = r.resample_grid3D_aligned(xt, out_dhw)
    assert rel(host(y), R.resample_grid3D_aligned(x, out_dhw)) < FWD_TOL
    g = f32(rng.standard_normal(tuple(y.shape)))
    y.backward(dev(g))
    M, V, P, F = R.aligned_args(x.shape[1:4], out_dhw, ("BORDER",) * 3)
    eg = R.sample_grid_transform(g, [M], [V], [P], [F], x.shape[1:4], LINEAR, BORDER, R.TRANSFORM_REVERSE)[:, 0]
    assert rel(host(xt.grad), eg) < GRAD_TOL


def _mac(rng, cs):
    D, H, W = cs
    return [f32(rng.standard_normal((1,) + s + (1,))) for s in ((D, H, W + 1), (D, H + 1, W), (D + 1, H, W))]


def test_velocity_scaled_and_resample(ngt):
    from ngt.data_structures import VelocityGrid
    rng = np.random.default_rng(6)
    cs, new = [6, 7, 8], [9, 5, 12]
    comps = _mac(rng, cs)
    vg = VelocityGrid(cs, x=dev(comps[0]), y=dev(comps[1]), z=dev(comps[2]))
    got = vg.scaled(new)
    xs, ys, zs = VelocityGrid.component_shapes(new)
    s = [o / i for i, o in zip(cs, new)]
    exp = [R.resample_grid3D_aligned(comps[0], xs, ("CENTER", "BORDER", "BORDER")) * s[2],
           R.resample_grid3D_aligned(comps[1], ys, ("BORDER", "CENTER", "BORDER")) * s[1],
           R.resample_grid3D_aligned(comps[2], zs, ("BORDER", "BORDER", "CENTER")) * s[0]]
    for gt, e in zip(got, exp):
        assert rel(host(gt), e) < FWD_TOL
    # resample_velocity by shape: magnitudes scaled by the mean ratio
    rv = VelocityGrid.resample_velocity(vg.scale_renderer, (vg.x, vg.y, vg.z), shape=new)
    m = float(np.mean(s))
    for gt, e, si in zip(rv, exp, (s[2], s[1], s[0])):
        assert rel(host(gt), e / si * m) < FWD_TOL
    # centred velocity [N,D,H,W,3] by scale: xy and z resampled as 2 + 1 channels
    c = f32(rng.standard_normal((1, 6, 7, 8, 3)))
    rc = VelocityGrid.resample_velocity(vg.scale_renderer, dev(c), scale=1.5, is_staggered=False)
    tgt = [9, 10, 12]
    ec = np.concatenate([R.resample_grid3D_aligned(c[..., :2], tgt), R.resample_grid3D_aligned(c[..., 2:], tgt)], -1)
    assert rel(host(rc), ec * 1.5) < FWD_TOL


def test_centered_to_staggered_and_scalar(ngt):
    """CLAMP boundary: the staggered targets put sample positions exactly on
    the BORDER validity edge (q = -0.5), where fp32 and fp64 may disagree."""
    from ngt.data_structures import VelocityGrid
    rng = np.random.default_rng(7)
    cs = [6, 7, 8]
    vg = VelocityGrid(cs, x=dev(np.zeros((1, 6, 7, 9, 1))), y=dev(np.zeros((1, 6, 8, 8, 1))),
                      z=dev(np.zeros((1, 7, 7, 8, 1))), scale_renderer=_renderer("CLAMP"))
    c = f32(rng.standard_normal((1, 6, 7, 8, 3)))
    out_cs = [8, 7, 10]
    got = vg._centered_to_staggered(dev(c), out_cs)
    xs, ys, zs = VelocityGrid.component_shapes(out_cs)
    f = [o / i for o, i in zip(out_cs, cs)]
    exp = [R.resample_grid3D_aligned(c[..., 0:1], xs, ("CENTER", "BORDER", "BORDER"), LINEAR, CLAMP) * f[2],
           R.resample_grid3D_aligned(c[..., 1:2], ys, ("BORDER", "CENTER", "BORDER"), LINEAR, CLAMP) * f[1],
           R.resample_grid3D_aligned(c[..., 2:3], zs, ("BORDER", "BORDER", "CENTER"), LINEAR, CLAMP) * f[0]]
    for gt, e in zip(got, exp):
        assert rel(host(gt), e) < FWD_TOL
    sc = f32(rng.standard_normal((1, 6, 7, 8, 2)))
    st = vg._scalar_centered_to_staggered(dev(sc))
    est = R.resample_grid3D_aligned(sc, (7, 8, 9), ("STAGGER_OUTPUT",) * 3, LINEAR, CLAMP)
    assert rel(host(st), est) < FWD_TOL
    # components onto one common grid with uniform scaling
    comps = _mac(rng, cs)
    vg2 = VelocityGrid(cs, x=dev(comps[0]), y=dev(comps[1]), z=dev(comps[2]), scale_renderer=_renderer("CLAMP"))
    shape = [10, 9, 11]
    outs = vg2._sampled_to_shape(shape)
    scale = [2.0 / v for v in shape[::-1]]
    for comp, o in zip(comps, outs):
        M = G.op_matrix(G.grid_transform(comp.shape[1:4], scale_=scale, center=True))
        V = G.op_matrix(np.eye(4))
        e = R.sample_grid_transform(comp, [M], [V], [V], [np.array([-1.0, 1, -1, 1, 1, -1])], shape, LINEAR, CLAMP,
                                    R.TRANSFORM)[:, 0]
        assert rel(host(o), e) < FWD_TOL


def test_density_scaled(ngt):
    from ngt.data_structures import DensityGrid
    rng = np.random.default_rng(8)
    d = f32(rng.uniform(0, 1, (1, 8, 9, 10, 1)))
    dg = DensityGrid([8, 9, 10], d=dev(d))
    got = dg.scaled([12, 6, 15])
    assert rel(host(got), R.resample_grid3D_aligned(d, (12, 6, 15))) < FWD_TOL
    assert dg.scaled([8, 9, 10]) is dg.d
    cp = dg.copy_scaled([4, 5, 5])
    assert cp.shape == [4, 5, 5] and rel(host(cp.d), R.resample_grid3D_aligned(d, (4, 5, 5))) < FWD_TOL
