"""Pins oracle/advect.py: identity and pure-shift closed forms, the revert
clamp invariant with an exact mask check, finite differences of the exact
adjoint, and the staggered<->centered adjoint identity."""
import numpy as np
import pytest

from oracle import BORDER, CLAMP, WRAP
from oracle import advect as A
from oracle import sampling as S


def _blob(shape, rng):
    D, H, W = shape
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    c = rng.uniform(0.3, 0.7, 3) * np.array([W, H, D])
    r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
    return np.exp(-r2 / (2 * 2.0 ** 2))[..., None]


@pytest.mark.parametrize("order,clamp", [(1, "NONE"), (2, "NONE"), (2, "MC_SMOOTH"), (1, "NEGATIVE")])
def test_dt_zero_identity(order, clamp):
    rng = np.random.default_rng(0)
    rho = rng.standard_normal((5, 6, 7, 1))
    v = rng.standard_normal((5, 6, 7, 3))
    out = A.warp_forward(rho, v, dt=0.0, order=order, clamp=clamp)
    exp = np.maximum(rho, 0) if clamp == "NEGATIVE" else rho
    np.testing.assert_array_equal(out, exp)


@pytest.mark.parametrize("order", [1, 2])
def test_integer_velocity_is_shift(order):
    rng = np.random.default_rng(1)
    rho = rng.standard_normal((6, 6, 6, 1))
    v = np.zeros((6, 6, 6, 3))
    v[..., 0] = 1.0   # x velocity +1 cell/step: out(x) = rho(x-1)
    v[..., 2] = -2.0  # z velocity -2: out(z) = rho(z+2)
    out = A.warp_forward(rho, v, dt=1.0, order=order, clamp="MC_SMOOTH" if order == 2 else "NONE", boundary=WRAP)
    np.testing.assert_allclose(out, np.roll(np.roll(rho, 1, axis=2), -2, axis=0), atol=1e-12)


def test_subcell_velocity_on_affine_field_interior():
    z, y, x = np.meshgrid(range(10), range(10), range(10), indexing="ij")
    rho = (0.5 * x + 0.2 * y - 0.3 * z)[..., None].astype(float)
    v = np.zeros((10, 10, 10, 3))
    v[...] = [0.3, -0.45, 0.7]
    out1 = A.warp_forward(rho, v, 1.0, 1)
    out2 = A.warp_forward(rho, v, 1.0, 2)
    exp = rho - (0.5 * 0.3 + 0.2 * -0.45 - 0.3 * 0.7)
    sl = (slice(2, -2),) * 3
    np.testing.assert_allclose(out1[sl], exp[sl], atol=1e-12)
    np.testing.assert_allclose(out2[sl], exp[sl], atol=1e-12)


@pytest.mark.parametrize("boundary", [BORDER, CLAMP])
def test_mc_smooth_revert_invariant_and_mask(boundary):
    rng = np.random.default_rng(2)
    rho = np.clip(_blob((12, 12, 12), rng) + 0.3 * rng.standard_normal((12, 12, 12, 1)), 0, 1)
    v = rng.uniform(-1.5, 1.5, (12, 12, 12, 3))
    out, parts = A.warp_forward(rho, v, 1.0, 2, "MC_SMOOTH", boundary, return_parts=True)
    # explicit recomputation of the clamp mask from independently computed reads
    i = A._idx(rho.shape[:3])
    p = i - v
    fwd = S.sample_linear(rho, p, boundary)
    back = S.sample_linear(fwd, i + v, boundary)
    corr = fwd + 0.5 * (rho - back)
    mn = S.sample_minmax(rho, p, boundary, False)
    mx = S.sample_minmax(rho, p, boundary, True)
    mask = (corr < mn) | (corr > mx)
    np.testing.assert_array_equal(parts["oob"], mask)
    assert mask.any() and (~mask).any()
    np.testing.assert_array_equal(out[mask], fwd[mask])
    assert np.all((out[~mask] >= mn[~mask]) & (out[~mask] <= mx[~mask]))


@pytest.mark.parametrize("order,clamp,boundary", [(1, "NONE", BORDER), (2, "NONE", BORDER), (2, "MC_SMOOTH", BORDER),
                                                   (2, "MC_SMOOTH", CLAMP), (1, "NEGATIVE", CLAMP),
                                                   (1, "NONE", WRAP), (2, "MC_SMOOTH", WRAP), (2, "NEGATIVE", WRAP)])
def test_warp_backward_fd(order, clamp, boundary):
    rng = np.random.default_rng(3)
    shape = (7, 6, 8)
    rho = _blob(shape, rng) + 0.1 * rng.standard_normal(shape + (1,))
    v = rng.uniform(-1.2, 1.2, shape + (3,))
    g = rng.standard_normal(shape + (1,))
    g_rho, g_v = A.warp_backward(rho, v, g, 1.0, order, clamp, boundary)
    h = 1e-7

    def L(r, vv):
        return np.sum(A.warp_forward(r, vv, 1.0, order, clamp, boundary) * g)

    checked = 0
    for (z, y, x) in [(3, 3, 4), (2, 1, 5), (5, 4, 2), (1, 2, 1), (4, 4, 6), (3, 2, 2)]:
        rp = rho.copy(); rp[z, y, x, 0] += h
        rm = rho.copy(); rm[z, y, x, 0] -= h
        fd = (L(rp, v) - L(rm, v)) / (2 * h)
        assert g_rho[z, y, x, 0] == pytest.approx(fd, rel=1e-5, abs=1e-7)
        for a in range(3):
            vp = v.copy(); vp[z, y, x, a] += h
            vm = v.copy(); vm[z, y, x, a] -= h
            fd = (L(rho, vp) - L(rho, vm)) / (2 * h)
            assert g_v[z, y, x, a] == pytest.approx(fd, rel=1e-5, abs=1e-7)
            checked += 1
    assert checked == 18


def test_staggered_centering_adjoint():
    rng = np.random.default_rng(4)
    D, H, W = 4, 5, 6
    vx = rng.standard_normal((D, H, W + 1, 1))
    vy = rng.standard_normal((D, H + 1, W, 1))
    vz = rng.standard_normal((D + 1, H, W, 1))
    c = A.centered_from_staggered(vx, vy, vz)
    g = rng.standard_normal(c.shape)
    gx, gy, gz = A.centered_from_staggered_adjoint(g)
    lhs = np.sum(c * g)
    rhs = np.sum(vx * gx) + np.sum(vy * gy) + np.sum(vz * gz)
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_velocity_self_advect_uniform_is_shift_and_fd():
    rng = np.random.default_rng(5)
    D, H, W = 5, 6, 7
    # uniform velocity field: self-advection leaves it unchanged (WRAP)
    vx = np.full((D, H, W + 1, 1), 0.4)
    vy = np.full((D, H + 1, W, 1), -0.3)
    vz = np.full((D + 1, H, W, 1), 0.2)
    out = A.velocity_self_advect_forward(vx, vy, vz, 1.0, 2, "MC_SMOOTH", WRAP)
    for o, ref in zip(out, (vx, vy, vz)):
        np.testing.assert_allclose(o, ref, atol=1e-12)
    # FD of the exact adjoint on a random field
    vx = rng.uniform(-0.8, 0.8, (D, H, W + 1, 1))
    vy = rng.uniform(-0.8, 0.8, (D, H + 1, W, 1))
    vz = rng.uniform(-0.8, 0.8, (D + 1, H, W, 1))
    gs = [rng.standard_normal(c.shape) for c in (vx, vy, vz)]
    grads = A.velocity_self_advect_backward(vx, vy, vz, gs, 1.0, 2, "NONE", BORDER)

    def L(c):
        return sum(np.sum(o * g) for o, g in zip(A.velocity_self_advect_forward(*c, 1.0, 2, "NONE", BORDER), gs))

    h = 1e-7
    comps = [vx, vy, vz]
    for ci, idx in [(0, (2, 3, 4, 0)), (1, (1, 2, 3, 0)), (2, (3, 3, 3, 0)), (0, (4, 1, 1, 0))]:
        cp = [c.copy() for c in comps]; cp[ci][idx] += h
        cm = [c.copy() for c in comps]; cm[ci][idx] -= h
        fd = (L(cp) - L(cm)) / (2 * h)
        assert grads[ci][idx] == pytest.approx(fd, rel=1e-5, abs=1e-7)
