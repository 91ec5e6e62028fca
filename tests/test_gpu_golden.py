"""The HIP library against the committed golden fixtures (tests/golden/):
forward 1e-5, gradients 1e-4 normwise relative, no oracle run needed."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FWD_TOL, GRAD_TOL = 1e-5, 1e-4


@pytest.fixture(scope="module")
def ngt_ops():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from ngt import ops, _lib
    _lib.load()
    return ops


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def rel(got, exp):
    got = got.detach().cpu().double().numpy()
    return float(np.max(np.abs(got - exp))) / max(float(np.max(np.abs(exp))), 1e-30)


@pytest.mark.parametrize("name,fm,bm", [("lut_linear_border", "LINEAR", "BORDER"), ("lut_min_clamp", "MIN", "CLAMP")])
def test_lut(ngt_ops, name, fm, bm):
    d = load(name)
    x, lut = dev(d["input"]).requires_grad_(), dev(d["lut"]).requires_grad_()
    y = ngt_ops.sample_grid_lut(x, lut, interpolation=fm, boundary=bm, relative_coords=True)
    assert rel(y, d["output"]) < FWD_TOL
    y.backward(dev(d["out_grad"]))
    assert rel(x.grad, d["input_grad"]) < GRAD_TOL
    assert rel(lut.grad, d["lut_grad"]) < GRAD_TOL


def test_transform(ngt_ops):
    d = load("transform_lindepth")
    x = dev(d["input"]).requires_grad_()
    y = ngt_ops.sample_grid_transform(x, d["M"], d["V"], d["P"], d["F"], d["output"].shape[2:5])
    assert rel(y, d["output"]) < FWD_TOL
    y.backward(dev(d["out_grad"]))
    assert rel(x.grad, d["input_grad"]) < GRAD_TOL


def test_blend(ngt_ops):
    d = load("blend_beer_lambert")
    x = dev(d["input"]).requires_grad_()
    y = ngt_ops.reduce_grid_blend(x, "BEER_LAMBERT")
    assert rel(y, d["output"]) < FWD_TOL
    y.backward(dev(d["out_grad"]))
    assert rel(x.grad, d["input_grad"]) < GRAD_TOL


def test_raymarch(ngt_ops):
    d = load("raymarch_beer_lambert")
    x = dev(d["input"]).requires_grad_()
    y = ngt_ops.raymarch_grid_transform(x, d["M"], d["V"], d["P"], d["F"], (9,) + d["output"].shape[2:4])
    assert rel(y, d["output"]) < FWD_TOL
    y.backward(dev(d["out_grad"]))
    assert rel(x.grad, d["input_grad"]) < GRAD_TOL


def test_advect(ngt_ops):
    d = load("advect_maccormack")
    x = dev(d["input"][None]).requires_grad_()
    v = dev(d["velocity"][None]).requires_grad_()
    y = ngt_ops.advect_grid_semi_langrange(x, v, -1.0, 2, "MC_SMOOTH", "BORDER", False)
    assert rel(y[0, 0], d["output"]) < FWD_TOL
    y.backward(dev(d["out_grad"][None, None]))
    assert rel(x.grad[0], d["input_grad"]) < GRAD_TOL
    assert rel(v.grad[0], d["velocity_grad"]) < GRAD_TOL
