"""Pins oracle/sampling.py against closed forms, an independent brute-force
corner-weight formulation, finite differences and adjoint identities."""
import itertools

import numpy as np
import pytest

from oracle import BORDER, CLAMP, WRAP
from oracle import sampling as S

BOUNDS = [BORDER, CLAMP, WRAP]


def brute_linear(grid, q, boundary):
    """Sum over the 8 corners of prod(weights) * value, one position at a time,
    written from the boundary rules (not from the lerp chain)."""
    D, H, W, C = grid.shape
    dims = (W, H, D)
    out = np.zeros((len(q), C))
    for n, p in enumerate(q):
        if boundary == BORDER and not all(0 <= p[a] + 0.5 < dims[a] for a in range(3)):
            continue
        fl = [int(np.floor(p[a])) for a in range(3)]
        ce = [int(np.ceil(p[a])) for a in range(3)]
        fr = [p[a] - np.floor(p[a]) for a in range(3)]
        acc = np.zeros(C)
        for bits in itertools.product((0, 1), repeat=3):
            idx = [ce[a] if bits[a] else fl[a] for a in range(3)]
            wgt = np.prod([fr[a] if bits[a] else 1.0 - fr[a] for a in range(3)])
            if boundary == BORDER:
                if not all(0 <= idx[a] < dims[a] for a in range(3)):
                    continue
            elif boundary == CLAMP:
                idx = [min(max(idx[a], 0), dims[a] - 1) for a in range(3)]
            else:
                idx = [idx[a] % dims[a] for a in range(3)]
            acc += wgt * grid[idx[2], idx[1], idx[0]]
        out[n] = acc
    return out


def rand_positions(rng, dims_xyz, n, lo=-2.0, hi_pad=2.0):
    return np.stack([rng.uniform(lo, dims_xyz[a] - 1 + hi_pad, n) for a in range(3)], axis=-1)


@pytest.mark.parametrize("boundary", BOUNDS)
def test_linear_matches_bruteforce(boundary):
    rng = np.random.default_rng(0)
    grid = rng.standard_normal((4, 5, 6, 2))
    q = rand_positions(rng, (6, 5, 4), 300)
    np.testing.assert_allclose(S.sample_linear(grid, q, boundary), brute_linear(grid, q, boundary), rtol=0, atol=1e-12)


@pytest.mark.parametrize("boundary", BOUNDS)
def test_integer_positions_copy(boundary):
    rng = np.random.default_rng(1)
    grid = rng.standard_normal((3, 4, 5, 1))
    z, y, x = np.meshgrid(range(3), range(4), range(5), indexing="ij")
    q = np.stack([x, y, z], -1).astype(float)
    np.testing.assert_array_equal(S.sample_linear(grid, q, boundary), grid)


def test_affine_field_exact_interior():
    z, y, x = np.meshgrid(range(6), range(7), range(8), indexing="ij")
    grid = (0.3 * x - 1.7 * y + 0.25 * z + 2.0)[..., None].astype(float)
    rng = np.random.default_rng(2)
    q = np.stack([rng.uniform(0, 7, 200), rng.uniform(0, 6, 200), rng.uniform(0, 5, 200)], -1)
    exp = 0.3 * q[:, 0] - 1.7 * q[:, 1] + 0.25 * q[:, 2] + 2.0
    for b in BOUNDS:
        np.testing.assert_allclose(S.sample_linear(grid, q, b)[:, 0], exp, atol=1e-12)


def test_border_outside_is_zero_and_band_is_half_weighted():
    grid = np.ones((4, 4, 4, 1))
    q = np.array([[-0.5000001, 1, 1], [3.5, 1, 1], [1, -0.6, 1], [1, 1, 3.51]])
    assert np.all(S.sample_linear(grid, q, BORDER) == 0.0)
    # q=-0.5: half of the ghost cell (0) and half of cell 0 (1)
    v = S.sample_linear(grid, np.array([[-0.5, 1.0, 1.0]]), BORDER)
    assert v[0, 0] == pytest.approx(0.5)
    # CLAMP keeps the constant everywhere
    assert np.allclose(S.sample_linear(grid, q, CLAMP), 1.0)


def test_wrap_periodic():
    rng = np.random.default_rng(3)
    grid = rng.standard_normal((4, 4, 4, 1))
    q = rng.uniform(0, 3, (50, 3))
    for shift in (4, -4, 8):
        np.testing.assert_allclose(S.sample_linear(grid, q + shift, WRAP), S.sample_linear(grid, q, WRAP), atol=1e-12)


def test_nearest_truncation_quirk():
    grid = np.arange(4, dtype=float).reshape(1, 1, 4, 1)
    grid = np.broadcast_to(grid, (1, 1, 4, 1)).copy()
    # q=-0.9: int(-0.4) = 0 -> cell 0 is read under BORDER (truncation toward 0)
    q = np.array([[-0.9, 0.0, 0.0], [1.49, 0, 0], [1.5, 0, 0], [3.6, 0, 0]])
    v = S.sample_nearest(grid, q, BORDER)[:, 0]
    np.testing.assert_array_equal(v, [0.0, 1.0, 2.0, 0.0])
    np.testing.assert_array_equal(S.sample_nearest(grid, q, CLAMP)[:, 0], [0.0, 1.0, 2.0, 3.0])


@pytest.mark.parametrize("boundary", BOUNDS)
@pytest.mark.parametrize("is_max", [False, True])
def test_minmax_bruteforce(boundary, is_max):
    rng = np.random.default_rng(4)
    grid = rng.standard_normal((4, 4, 5, 2))
    q = rand_positions(rng, (5, 4, 4), 200, lo=-1.5, hi_pad=1.5)
    got = S.sample_minmax(grid, q, boundary, is_max)
    dims = (5, 4, 4)
    for n, p in enumerate(q):
        if boundary == BORDER and not all(0 <= p[a] + 0.5 < dims[a] for a in range(3)):
            assert np.all(got[n] == 0)
            continue
        vals = []
        any_out = False
        for bits in itertools.product((0, 1), repeat=3):
            idx = [int(np.ceil(p[a])) if bits[a] else int(np.floor(p[a])) for a in range(3)]
            if boundary == BORDER:
                if not all(0 <= idx[a] < dims[a] for a in range(3)):
                    any_out = True
                    continue
            elif boundary == CLAMP:
                idx = [min(max(idx[a], 0), dims[a] - 1) for a in range(3)]
            else:
                idx = [idx[a] % dims[a] for a in range(3)]
            vals.append(grid[idx[2], idx[1], idx[0]])
        if any_out:
            vals.append(np.zeros(2))
        exp = np.max(vals, 0) if is_max else np.min(vals, 0)
        np.testing.assert_allclose(got[n], exp)


@pytest.mark.parametrize("boundary", BOUNDS)
def test_data_gradient_fd(boundary):
    rng = np.random.default_rng(5)
    grid = rng.standard_normal((5, 5, 5, 2))
    q = rand_positions(rng, (5, 5, 5), 200, lo=-0.4, hi_pad=0.3)
    # keep away from integer kinks
    q = np.floor(q) + np.clip(q - np.floor(q), 0.05, 0.95)
    dx, dy, dz = S.data_gradient(grid, q, boundary)
    h = 1e-6
    for a, d in enumerate((dx, dy, dz)):
        e = np.zeros(3)
        e[a] = h
        fd = (S.sample_linear(grid, q + e, boundary) - S.sample_linear(grid, q - e, boundary)) / (2 * h)
        np.testing.assert_allclose(d, fd, atol=1e-6)


def test_data_gradient_zero_at_integer_coordinate():
    rng = np.random.default_rng(6)
    grid = rng.standard_normal((4, 4, 4, 1))
    q = np.array([[1.0, 1.3, 2.2]])
    dx, dy, dz = S.data_gradient(grid, q, CLAMP)
    assert dx[0, 0] == 0.0 and dy[0, 0] != 0.0


@pytest.mark.parametrize("boundary", BOUNDS)
def test_scatter_is_adjoint(boundary):
    rng = np.random.default_rng(7)
    grid = rng.standard_normal((4, 5, 6, 2))
    q = rand_positions(rng, (6, 5, 4), 400)
    g = rng.standard_normal((400, 2))
    lhs = np.sum(S.sample_linear(grid, q, boundary) * g)
    rhs = np.sum(grid * S.scatter_linear(grid.shape, q, g, boundary, mask="forward"))
    assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))
    lhs = np.sum(S.sample_nearest(grid, q, boundary) * g)
    rhs = np.sum(grid * S.scatter_nearest(grid.shape, q, g, boundary))
    assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))
    for is_max in (False, True):
        lhs = np.sum(S.sample_minmax(grid, q, boundary, is_max) * g)
        rhs = np.sum(grid * S.scatter_minmax(grid, q, g, boundary, is_max))
        assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))


def test_reference_mask_band_and_counts():
    # position in (-1, -0.5): forward reads 0, reference scatter still deposits
    q = np.array([[-0.75, 0.0, 0.0]])
    g = np.ones((1, 1))
    fwd = S.scatter_linear((2, 2, 2, 1), q, g, BORDER, mask="forward")
    cnt = np.zeros((2, 2, 2), dtype=np.int64)
    ref = S.scatter_linear((2, 2, 2, 1), q, g, BORDER, mask="reference", count=cnt)
    assert np.all(fwd == 0)
    assert ref[0, 0, 0, 0] == pytest.approx(0.25)  # ceil corner x=0 with weight fract(-0.75)=0.25
    # integer y,z: ceil==floor, so each in-grid corner is hit 4 times (duplicates counted)
    assert cnt[0, 0, 0] == 4 and cnt.sum() == 4
