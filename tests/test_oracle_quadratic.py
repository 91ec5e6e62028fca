"""QUADRATIC (quadratic B-spline) sampling of the oracle pinned by closed
forms: the per-axis weights against the B-spline pieces written in the
offset t from the nearest cell, partition of unity and reproduction of linear
fields, the BORDER quirk (always 0), the adjoint identity of the scatter, and
the reference scatter's masks (sampling.hpp:247-365,
resample_grid.cu.cc:904-1080)."""
import numpy as np
import pytest

from oracle import BORDER, CLAMP, WRAP, QUADRATIC
from oracle import sampling as S
from oracle import resample as R


def _rand_q(rng, n, dims, margin):
    return np.stack([rng.uniform(-margin, d - 1 + margin, n) for d in dims], -1)


def test_weights_are_the_bspline_pieces():
    rng = np.random.default_rng(0)
    q = rng.uniform(-3.0, 9.0, (500, 3))
    idx, w = S.quadratic_stencil(q, np.array([8, 8, 8]))
    # t = offset from the nearest cell (ties round up), in [-1/2, 1/2)
    t = q - np.floor(q + 0.5)
    np.testing.assert_array_equal(idx[1], np.floor(q + 0.5).astype(np.int64))
    np.testing.assert_allclose(w[0], 0.5 * (0.5 - t) ** 2, atol=1e-15)
    np.testing.assert_allclose(w[1], 0.75 - t * t, atol=1e-15)
    np.testing.assert_allclose(w[2], 0.5 * (0.5 + t) ** 2, atol=1e-15)
    np.testing.assert_allclose(w[0] + w[1] + w[2], 1.0, atol=1e-15)
    # first moment: the weighted cell index is the position itself
    np.testing.assert_allclose(w[0] * idx[0] + w[1] * idx[1] + w[2] * idx[2], q, atol=1e-12)


@pytest.mark.parametrize("bound", [CLAMP, WRAP])
def test_linear_field_reproduced_inside(bound):
    rng = np.random.default_rng(1)
    D, H, W = 7, 8, 9
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    coef = rng.standard_normal(4)
    grid = (coef[0] + coef[1] * x + coef[2] * y + coef[3] * z)[..., None] * np.array([1.0, -2.0])
    q = np.stack([rng.uniform(1.0, d - 2.0, 300) for d in (W, H, D)], -1)  # stencils inside the grid
    got = S.sample_quadratic(grid, q, bound)
    exp = (coef[0] + coef[1] * q[:, 0] + coef[2] * q[:, 1] + coef[3] * q[:, 2])[:, None] * np.array([1.0, -2.0])
    np.testing.assert_allclose(got, exp, atol=1e-12)


def test_brute_force_and_border_quirk():
    rng = np.random.default_rng(2)
    grid = rng.standard_normal((5, 6, 7, 2))
    q = _rand_q(rng, 50, (7, 6, 5), 2.0)
    got = S.sample_quadratic(grid, q, CLAMP)
    for k in range(len(q)):
        acc = np.zeros(2)
        c = np.floor(q[k] + 0.5).astype(int)
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    i = c + np.array([dx, dy, dz])
                    w = np.prod(S.bspline2_weight(np.abs(q[k] - i)))
                    ii = np.clip(i, 0, np.array([6, 5, 4]))
                    acc += w * grid[ii[2], ii[1], ii[0]]
        np.testing.assert_allclose(got[k], acc, atol=1e-12)
    assert np.all(S.sample_quadratic(grid, q, BORDER) == 0.0)


@pytest.mark.parametrize("bound", [CLAMP, WRAP])
def test_scatter_is_the_adjoint(bound):
    rng = np.random.default_rng(3)
    shape = (5, 6, 7, 2)
    x = rng.standard_normal(shape)
    q = _rand_q(rng, 80, (7, 6, 5), 3.0)
    y = rng.standard_normal((80, 2))
    lhs = np.sum(S.sample_quadratic(x, q, bound) * y)
    rhs = np.sum(x * S.scatter_quadratic(shape, q, y, bound))
    assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))
    assert np.all(S.scatter_quadratic(shape, q, y, BORDER) == 0.0)


def test_reference_border_scatter_masks():
    """Inside the grid the reference's BORDER scatter is the exact (CLAMP)
    scatter without the centre cell; it counts one hit per scattered cell."""
    rng = np.random.default_rng(4)
    shape = (6, 6, 6, 1)
    q = rng.uniform(1.0, 4.0, (40, 3))
    g = rng.standard_normal((40, 1))
    count = np.zeros(shape[:3], dtype=np.int64)
    ref = S.scatter_quadratic(shape, q, g, BORDER, mask="reference", count=count)
    full = S.scatter_quadratic(shape, q, g, CLAMP)
    idx, w = S.quadratic_stencil(q, np.array([6, 6, 6]))
    centre = np.zeros(shape)
    np.add.at(centre, (idx[1][:, 2], idx[1][:, 1], idx[1][:, 0]), g * (w[1][:, 0] * w[1][:, 1] * w[1][:, 2])[:, None])
    np.testing.assert_allclose(ref + centre, full, atol=1e-12)
    assert count.sum() == 26 * len(q)
    # outside -1.5 < q < dims nothing is scattered
    far = np.array([[-1.6, 2.0, 2.0], [2.0, 6.0, 2.0]])
    assert np.all(S.scatter_quadratic(shape, far, np.ones((2, 1)), BORDER, mask="reference") == 0.0)


def test_transform_grad_quadratic_adjoint():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((1, 6, 7, 8, 2))
    out_dhw = (9, 5, 11)
    M, V, P, F = R.aligned_args(x.shape[1:4], out_dhw, ("BORDER",) * 3)
    y = R.sample_grid_transform(x, [M], [V], [P], [F], out_dhw, QUADRATIC, CLAMP, R.TRANSFORM)
    g = rng.standard_normal(y.shape)
    gx = R.sample_grid_transform_grad(x, g, [M], [V], [P], [F], QUADRATIC, CLAMP, R.TRANSFORM)
    assert abs(np.sum(y * g) - np.sum(x * gx)) < 1e-10 * abs(np.sum(y * g))
