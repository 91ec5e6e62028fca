"""Pins the oracle's aligned resampling (resample_grid3D_aligned,
renderer.py:843-905) by closed forms independent of its camera chain: per
alignment, where each output cell centre lands in the input, checked on
fields linear in the input index (reproduced exactly by trilinear reads away
from the BORDER band)."""
import numpy as np
import pytest

from oracle import BORDER, CLAMP
from oracle import resample as R


def _ramp(shape_dhw, axis, fn):
    """Field of value fn(index along axis) ([1, D, H, W, 1])."""
    idx = np.arange(shape_dhw[axis], dtype=np.float64)
    v = fn(idx)
    sh = [1, 1, 1]
    sh[axis] = shape_dhw[axis]
    return np.broadcast_to(v.reshape(sh), shape_dhw)[None, ..., None].copy()


AX = {0: 2, 1: 1, 2: 0}  # x, y, z -> DHW axis


@pytest.mark.parametrize("ax", [0, 1, 2])
def test_border_alignment_maps_normalised_centres(ax):
    ind, outd = (6, 7, 8), (9, 5, 12)
    a = AX[ax]
    n_in, n_out = ind[a], outd[a]
    data = _ramp(ind, a, lambda j: (j + 0.5) / n_in)
    out = R.resample_grid3D_aligned(data, outd, boundary=CLAMP)
    exp = _ramp(outd, a, lambda i: np.clip((i + 0.5) / n_out, 0.5 / n_in, 1 - 0.5 / n_in))
    np.testing.assert_allclose(out, exp, atol=1e-12)


@pytest.mark.parametrize("ax", [0, 1, 2])
def test_center_alignment_maps_corner_centres(ax):
    ind, outd = (6, 7, 8), (11, 4, 13)
    a = AX[ax]
    n_in, n_out = ind[a], outd[a]
    align = ["BORDER"] * 3
    align[ax] = "CENTER"
    data = _ramp(ind, a, lambda j: j / (n_in - 1.0))
    out = R.resample_grid3D_aligned(data, outd, tuple(align), boundary=CLAMP)
    exp = _ramp(outd, a, lambda i: i / (n_out - 1.0))
    np.testing.assert_allclose(out, exp, atol=1e-12)
    # the aligned axis never leaves [0, n_in - 1]: BORDER attenuates only through the other axes' half-cell bands
    out_b = R.resample_grid3D_aligned(data, outd, tuple(align), boundary=BORDER)
    inner = [slice(None)] * 5
    for b in range(1, 4):
        if b - 1 != a:
            inner[b] = slice(2, -2)
    np.testing.assert_allclose(out_b[tuple(inner)], exp[tuple(inner)], atol=1e-12)


@pytest.mark.parametrize("ax", [0, 1, 2])
def test_stagger_alignments(ax):
    a = AX[ax]
    # centred -> staggered (STAGGER_OUTPUT): output centre i sits on input border i - 1/2
    ind = (6, 7, 8)
    outd = list(ind)
    outd[a] += 1
    n_in = ind[a]
    align = ["BORDER"] * 3
    align[ax] = "STAGGER_OUTPUT"
    data = _ramp(ind, a, lambda j: j + 3.0)
    out = R.resample_grid3D_aligned(data, tuple(outd), tuple(align), boundary=CLAMP)
    exp = _ramp(tuple(outd), a, lambda i: np.clip(i - 0.5, 0.0, n_in - 1.0) + 3.0)
    np.testing.assert_allclose(out, exp, atol=1e-12)
    # staggered -> centred (STAGGER_INPUT): input centres on output borders
    ind2 = tuple(outd)
    outd2 = ind
    align[ax] = "STAGGER_INPUT"
    data2 = _ramp(ind2, a, lambda j: 2.0 * j)
    out2 = R.resample_grid3D_aligned(data2, outd2, tuple(align), boundary=CLAMP)
    exp2 = _ramp(outd2, a, lambda i: 2.0 * (i + 0.5))
    np.testing.assert_allclose(out2, exp2, atol=1e-12)


def test_aligned_grad_is_adjoint():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 5, 6, 7, 2))
    for align in (("BORDER",) * 3, ("CENTER", "BORDER", "STAGGER_OUTPUT")):
        y = R.resample_grid3D_aligned(x, (8, 4, 9), align)
        g = rng.standard_normal(y.shape)
        gx = R.resample_grid3D_aligned_grad(x.shape, g, align)
        assert abs(float((y * g).sum()) - float((x * gx).sum())) < 1e-10
