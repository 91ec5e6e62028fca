"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py) are reproduced by the oracle: a change to the
oracle that moves any output shows up here."""
import os

import numpy as np
import pytest

from oracle import BORDER, CLAMP, LINEAR, MIN
from oracle import resample as R
from oracle import blend as B
from oracle import raymarch as RM
from oracle import advect as A

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.mark.parametrize("name,fm,bm", [("lut_linear_border", LINEAR, BORDER), ("lut_min_clamp", MIN, CLAMP)])
def test_lut(name, fm, bm):
    d = load(name)
    np.testing.assert_allclose(R.sample_grid_lut(d["input"], d["lut"], fm, bm, relative=True), d["output"],
                               rtol=0, atol=1e-12)
    gi, gl = R.sample_grid_lut_grad(d["input"], d["lut"], d["out_grad"], fm, bm, relative=True)
    np.testing.assert_allclose(gi, d["input_grad"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(gl, d["lut_grad"], rtol=0, atol=1e-12)


def test_transform():
    d = load("transform_lindepth")
    y = R.sample_grid_transform(d["input"], d["M"], d["V"], d["P"], d["F"], d["output"].shape[2:5], LINEAR, BORDER)
    np.testing.assert_allclose(y, d["output"], rtol=0, atol=1e-12)
    gi = R.sample_grid_transform_grad(d["input"], d["out_grad"], d["M"], d["V"], d["P"], d["F"], LINEAR, BORDER)
    np.testing.assert_allclose(gi, d["input_grad"], rtol=0, atol=1e-12)


def test_blend():
    d = load("blend_beer_lambert")
    y = B.reduce_grid_blend(d["input"], B.BEER_LAMBERT)
    np.testing.assert_allclose(y, d["output"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(B.reduce_grid_blend_grad(d["out_grad"], y, d["input"], B.BEER_LAMBERT), d["input_grad"],
                               rtol=0, atol=1e-12)


def test_raymarch():
    d = load("raymarch_beer_lambert")
    out_dhw = (9,) + d["output"].shape[2:4]
    y = RM.raymarch_grid_transform(d["input"], d["M"], d["V"], d["P"], d["F"], out_dhw, LINEAR, BORDER, B.BEER_LAMBERT)
    np.testing.assert_allclose(y, d["output"], rtol=0, atol=1e-12)
    gi = RM.raymarch_grid_transform_grad(d["input"], y, d["out_grad"], d["M"], d["V"], d["P"], d["F"], out_dhw,
                                         LINEAR, BORDER, B.BEER_LAMBERT)
    np.testing.assert_allclose(gi, d["input_grad"], rtol=0, atol=1e-12)


def test_advect():
    d = load("advect_maccormack")
    np.testing.assert_allclose(A.warp_forward(d["input"], d["velocity"], 1.0, 2, "MC_SMOOTH", BORDER), d["output"],
                               rtol=0, atol=1e-12)
    gr, gv = A.warp_backward(d["input"], d["velocity"], d["out_grad"], 1.0, 2, "MC_SMOOTH", BORDER)
    np.testing.assert_allclose(gr, d["input_grad"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(gv, d["velocity_grad"], rtol=0, atol=1e-12)
