"""GPU parity of QUADRATIC (quadratic B-spline) transform sampling
(SampleGridTransform / SampleGridTransformGrad with interpolation QUADRATIC,
sampling.hpp:247-365, resample_grid.cu.cc:904-1185) against the FP64
oracle: forward, exact and reference gradients, mips, and the aligned
resampling of a renderer whose filter mode is QUADRATIC.

Tolerances: forward 1e-5, gradients 1e-4, normwise relative.
"""
import numpy as np
import pytest
import torch

from oracle import BORDER, CLAMP, QUADRATIC
from oracle import geometry as G
from oracle import mips as MP
from oracle import resample as R

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
GRAD_TOL = 1e-4
BNAME = {BORDER: "BORDER", CLAMP: "CLAMP"}


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


def ring_args(dims, out_dhw, views=3, dist=3.0):
    M = G.grid_transform(dims, translation_=(0.0123, -0.0217, 0.0311), center=True, scale_=[1.0 / dims[2]] * 3)
    cams = G.look_at_ring_cameras(views, dist, 40.0, out_dhw[2] / out_dhw[1], dist - 1.0, dist + 1.0)
    return ([G.op_matrix(M).astype(np.float32)], [G.op_matrix(c[0]).astype(np.float32) for c in cams],
            [G.op_matrix(c[1]).astype(np.float32) for c in cams], [c[2].astype(np.float32) for c in cams])


@pytest.mark.parametrize("mode", ["TRANSFORM_LINDEPTH", "TRANSFORM_REVERSE"])
@pytest.mark.parametrize("C", [1, 2, 4])
def test_quadratic_forward_clamp(ops, mode, C):
    rng = np.random.default_rng(1000 + C)
    dims = (10, 9, 11)
    Ms, Vs, Ps, Fs = ring_args(dims, (8, 7, 9))
    if mode.endswith("REVERSE"):
        inp, out_dhw = f32(rng.standard_normal((1, 8, 7, 9, C))), dims
    else:
        inp, out_dhw = f32(rng.standard_normal((1,) + dims + (C,))), (8, 7, 9)
    exp = R.sample_grid_transform(inp, Ms, Vs, Ps, Fs, out_dhw, QUADRATIC, CLAMP, R.COORDINATE_MODE[mode])
    got = host(ops.sample_grid_transform(dev(inp), Ms, Vs, Ps, Fs, out_dhw, interpolation="QUADRATIC",
                                         boundary="CLAMP", coordinate_mode=mode))
    assert rel(got, exp) < FWD_TOL


def test_quadratic_border_reads_zero(ops):
    """The reference's QUADRATIC read returns 0 under BORDER (sampling.hpp:284-287)."""
    rng = np.random.default_rng(1001)
    dims = (8, 8, 8)
    Ms, Vs, Ps, Fs = ring_args(dims, (6, 5, 7))
    x = dev(rng.standard_normal((1,) + dims + (2,))).requires_grad_()
    y = ops.sample_grid_transform(x, Ms, Vs, Ps, Fs, (6, 5, 7), interpolation="QUADRATIC", boundary="BORDER")
    assert torch.count_nonzero(y) == 0
    y.backward(torch.ones_like(y))
    assert torch.count_nonzero(x.grad) == 0


@pytest.mark.parametrize("semantics", ["exact", "reference"])
@pytest.mark.parametrize("bound", [BORDER, CLAMP])
def test_quadratic_backward(ops, semantics, bound):
    rng = np.random.default_rng(1002 + bound)
    dims = (9, 9, 9)
    Ms, Vs, Ps, Fs = ring_args(dims, (7, 6, 8))
    inp = f32(rng.standard_normal((1,) + dims + (2,)))
    g = f32(rng.standard_normal((1, 3, 7, 6, 8, 2)))
    exp = R.sample_grid_transform_grad(inp, g, Ms, Vs, Ps, Fs, QUADRATIC, bound, semantics=semantics)
    with ops.semantics(semantics):
        got = host(ops.sample_grid_transform_grad(dev(inp), dev(g), Ms, Vs, Ps, Fs, interpolation="QUADRATIC",
                                                  boundary=BNAME[bound]))
    if semantics == "exact" and bound == BORDER:
        assert np.all(exp == 0.0) and np.all(got == 0.0)
    else:
        assert rel(got, exp) < GRAD_TOL


@pytest.mark.parametrize("mip", [MP.MIP_NEAREST, MP.MIP_LINEAR])
def test_quadratic_mips(ops, mip):
    rng = np.random.default_rng(1004 + mip)
    dims = (16, 16, 16)
    Ms, Vs, Ps, Fs = ring_args(dims, (6, 5, 7), dist=6.0)
    inp = f32(rng.standard_normal((1,) + dims + (1,)))
    mname = {MP.MIP_NEAREST: "NEAREST", MP.MIP_LINEAR: "LINEAR"}[mip]
    exp = R.sample_grid_transform(inp, Ms, Vs, Ps, Fs, (6, 5, 7), QUADRATIC, CLAMP, R.TRANSFORM_LINDEPTH,
                                  mipmapping=mip, num_mipmaps=2, mip_bias=0.3)
    xt = dev(inp).requires_grad_()
    y = ops.sample_grid_transform(xt, Ms, Vs, Ps, Fs, (6, 5, 7), interpolation="QUADRATIC", boundary="CLAMP",
                                  mipmapping=mname, num_mipmaps=2, mip_bias=0.3)
    assert rel(host(y), exp) < FWD_TOL
    g = f32(rng.standard_normal(exp.shape))
    y.backward(dev(g))
    eg = R.sample_grid_transform_grad(inp, g, Ms, Vs, Ps, Fs, QUADRATIC, CLAMP, R.TRANSFORM_LINDEPTH,
                                      mipmapping=mip, num_mipmaps=2, mip_bias=0.3)
    assert rel(host(xt.grad), eg) < GRAD_TOL


def test_quadratic_aligned_resampling(ops):
    """A renderer with filter mode QUADRATIC (the reference's scale renderer
    setting, neural_data_structures.py:1984) resamples between resolutions
    through the QUADRATIC transform op."""
    from ngt.renderer import Renderer
    rng = np.random.default_rng(1006)
    x = f32(rng.standard_normal((1, 7, 9, 8, 2)))
    r = Renderer(filter_mode="QUADRATIC", boundary_mode="CLAMP", mipmapping="NONE", blend_mode="ADDITIVE")
    xt = dev(x).requires_grad_()
    y = r.resample_grid3D_aligned(xt, (11, 6, 13), "CENTER", "BORDER", "CENTER")
    exp = R.resample_grid3D_aligned(x, (11, 6, 13), ("CENTER", "BORDER", "CENTER"), QUADRATIC, CLAMP)
    assert rel(host(y), exp) < FWD_TOL
    g = f32(rng.standard_normal(exp.shape))
    y.backward(dev(g))
    eg = R.resample_grid3D_aligned_grad(x.shape, g, ("CENTER", "BORDER", "CENTER"), QUADRATIC, CLAMP)
    assert rel(host(xt.grad), eg) < GRAD_TOL
