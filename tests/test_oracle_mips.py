"""Pins the mip-mapped sampling of the oracle (oracle/mips.py and its use in
oracle/resample.py) with closed forms, brute force and adjoint identities:

* mip generation = block means over 2^m cubes (floor dims), computed by a
  reshape instead of the pairwise averaging;
* level selection: levels holding the constant m return the clamped LoD
  (LINEAR) or its rounding (NEAREST);
* level positions: a field linear in the cell-centre position is reproduced
  exactly at every level with a 0.5 cell-centre offset, and shifted by
  2^(m-1) with offset 0;
* the LoD: per-axis (forward modes, central differences) vs per-difference
  (reverse modes) lengths on a linear map, and log2 of the cell size of an
  orthographic camera;
* gradients: adjoint identities of the level scatter and of the collapse,
  finite differences of the LUT position / LoD gradient.
"""
import numpy as np
import pytest

from oracle import BORDER, CLAMP, LINEAR, NEAREST, MIN, MAX
from oracle import geometry as G
from oracle import mips as MP
from oracle import resample as R


def _block_mean(g, m):
    k = 1 << m
    d, h, w = (s // k for s in g.shape[:3])
    c = g[:d * k, :h * k, :w * k]
    return c.reshape(d, k, h, k, w, k, -1).mean(axis=(1, 3, 5))


def test_generate_mips_is_block_mean():
    g = np.random.default_rng(1).standard_normal((9, 10, 13, 2))
    levels = MP.generate_mips(g, 3)
    assert [l.shape[:3] for l in levels] == [(9, 10, 13), (4, 5, 6), (2, 2, 3), (1, 1, 1)]
    for m in range(4):
        np.testing.assert_allclose(levels[m], _block_mean(g, m), rtol=1e-13, atol=1e-13)


def test_collapse_is_adjoint_of_generation():
    rng = np.random.default_rng(2)
    g = rng.standard_normal((8, 7, 6, 3))
    levels = MP.generate_mips(g, 2)
    ys = [rng.standard_normal(l.shape) for l in levels]
    lhs = sum(float((l * y).sum()) for l, y in zip(levels, ys))
    rhs = float((g * MP.collapse_adjoint(ys)).sum())
    assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))


@pytest.mark.parametrize("bias", [0.0, 0.7, -0.4])
def test_level_selection_closed_form(bias):
    L = 3
    levels = [np.full((16 >> m, 16 >> m, 16 >> m, 1), float(m)) for m in range(L + 1)]
    lod = np.linspace(-1.0, 4.5, 37)
    pos = np.tile(np.array([[8.0, 8.0, 8.0]]), (lod.size, 1))
    lin = MP.sample_levels(levels, pos, lod, MP.MIP_LINEAR, bias, 0.5, LINEAR, CLAMP)[:, 0]
    np.testing.assert_allclose(lin, np.clip(lod + bias, 0.0, L), atol=1e-12)
    near = MP.sample_levels(levels, pos, lod, MP.MIP_NEAREST, bias, 0.5, LINEAR, CLAMP)[:, 0]
    np.testing.assert_array_equal(near, np.minimum(np.floor(np.clip(lod + bias, 0.0, L) + 0.5), L))


def _linear_field(shape, a, b):
    z, y, x = np.meshgrid(*(np.arange(s) + 0.5 for s in shape), indexing="ij")
    return (a[0] * x + a[1] * y + a[2] * z + b)[..., None]


@pytest.mark.parametrize("mode", [MP.MIP_LINEAR, MP.MIP_NEAREST])
def test_linear_field_reproduced_at_every_level(mode):
    a, b = np.array([0.3, -1.1, 0.6]), 2.0
    L = 2
    levels = MP.generate_mips(_linear_field((16, 16, 16), a, b), L)
    rng = np.random.default_rng(3)
    pos = rng.uniform(2.0, 14.0, size=(200, 3))  # inside the last level's interpolation range
    lod = rng.uniform(-0.5, 2.5, size=200)
    out = MP.sample_levels(levels, pos, lod, mode, 0.0, 0.5, LINEAR, BORDER)[:, 0]
    np.testing.assert_allclose(out, pos @ a + b, atol=1e-12)


def test_zero_cell_centre_offset_shifts_each_level():
    a, b = np.array([1.0, 2.0, -0.5]), 0.25
    levels = MP.generate_mips(_linear_field((16, 16, 16), a, b), 2)
    pos = np.array([[6.3, 7.1, 8.4], [5.0, 9.5, 7.25]])
    for m in range(3):
        out = MP.sample_levels(levels, pos, np.full(2, float(m)), MP.MIP_NEAREST, 0.0, 0.0, LINEAR, CLAMP)[:, 0]
        # level m holds f at 2^m (i + 0.5); read at q = pos / 2^m it gives f(pos + 2^(m-1))
        np.testing.assert_allclose(out, (pos + 0.5 * (1 << m)) @ a + b, atol=1e-12)


@pytest.mark.parametrize("fm", [LINEAR, NEAREST, MIN, MAX])
@pytest.mark.parametrize("mode", [MP.MIP_LINEAR, MP.MIP_NEAREST])
def test_level_scatter_is_adjoint(fm, mode):
    rng = np.random.default_rng(4)
    levels = MP.generate_mips(rng.standard_normal((12, 10, 9, 2)), 2)
    pos = rng.uniform(-1.0, 13.0, size=(300, 3))
    lod = rng.uniform(-0.3, 2.6, size=300)
    g = rng.standard_normal((300, 2))
    out = MP.sample_levels(levels, pos, lod, mode, 0.1, 0.5, fm, BORDER)
    gl = MP.scatter_levels([l.shape for l in levels], pos, g, lod, mode, 0.1, 0.5, fm, BORDER, levels)
    lhs = float((out * g).sum())
    rhs = sum(float((l * x).sum()) for l, x in zip(levels, gl))
    assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))


def test_linear_map_lod_forward_vs_reverse():
    A = np.array([[1.0, 1.0, 1.0], [0.0, 0.5, 0.0], [0.0, 0.0, 0.25]])
    fn = lambda idx: idx @ A.T
    rows = np.linalg.norm(A, axis=1).max()  # per object-space axis
    cols = np.linalg.norm(A, axis=0).max()  # per difference vector
    np.testing.assert_allclose(MP.transform_lod(fn, (2, 3, 4), True), np.log2(rows), atol=1e-12)
    np.testing.assert_allclose(MP.transform_lod(fn, (2, 3, 4), False), max(0.0, np.log2(cols)), atol=1e-12)
    np.testing.assert_allclose(MP.central_lod(fn, (2, 3, 4)), np.log2(rows), atol=1e-12)


def _ortho_camera():
    # 16 x 8 x 4 view volume over a 4 x 4 x 4 frustum grid: cells of 4 x 2 x 1 -> LoD 2
    trbl = [4.0, 8.0, -4.0, -8.0]
    near, far = 1.0, 5.0
    P = G.projection(near, far, trbl, perspective=False)
    return G.op_matrix(np.eye(4)), G.op_matrix(P), G.frustum_params(near, far, trbl)


@pytest.mark.parametrize("cm", [R.TRANSFORM, R.TRANSFORM_LINDEPTH])
def test_orthographic_lod_closed_form(cm):
    V, P, F = _ortho_camera()
    cx = G.CameraXform.from_op_args(np.eye(4).reshape(-1), V, P, F)
    np.testing.assert_allclose(R.transform_lod(cx, (4, 4, 4), (8, 8, 8), cm), 2.0, atol=1e-9)
    lut = R.lod_transform((8, 8, 8), [V], [P], [F], (4, 4, 4), inverse_transform=False,
                          linearize_depth=cm == R.TRANSFORM_LINDEPTH)
    np.testing.assert_allclose(lut[0, ..., 3], 2.0, atol=1e-9)
    pos = R.transform_positions(cx, (4, 4, 4), (8, 8, 8), cm)
    np.testing.assert_allclose(lut[0, ..., :3], pos, atol=1e-12)
    rel = R.lod_transform((8, 8, 8), [V], [P], [F], (4, 4, 4), linearize_depth=cm == R.TRANSFORM_LINDEPTH,
                          relative_coords=True)
    np.testing.assert_allclose(rel[0, ..., :3], pos - R._index_grid(4, 4, 4), atol=1e-12)
    # reverse: one object-space cell moves 1/4, 1/2, 1 frustum cells -> LoD 0
    np.testing.assert_allclose(R.lod_transform((4, 4, 4), [V], [P], [F], (8, 8, 8), inverse_transform=True,
                                               linearize_depth=cm == R.TRANSFORM_LINDEPTH)[0, ..., 3], 0.0,
                               atol=1e-9)


def test_transform_mips_constant_field_and_linear_mode():
    V, P, F = _ortho_camera()
    M = np.eye(4).reshape(1, 16)
    inp = np.full((1, 8, 8, 8, 2), 3.5)
    for mode in (MP.MIP_LINEAR, MP.MIP_NEAREST):
        out = R.sample_grid_transform(inp, M, [V], [P], [F], (4, 4, 4), LINEAR, CLAMP, R.TRANSFORM, True, mode, 2,
                                      0.0)
        np.testing.assert_allclose(out, 3.5, atol=1e-12)


def test_lut_mips_adjoint_and_fd():
    rng = np.random.default_rng(5)
    inp = rng.standard_normal((2, 8, 9, 10, 2))
    lut = np.concatenate([rng.uniform(0.5, 8.0, size=(3, 3, 4, 5, 3)), rng.uniform(0.1, 1.9, size=(3, 3, 4, 5, 1))],
                         -1)
    g = rng.standard_normal((2, 3, 3, 4, 5, 2))
    kw = dict(filter_mode=LINEAR, boundary=BORDER, cell_center_offset=0.5, relative=False, normalized=False,
              mipmapping=MP.MIP_LINEAR, num_mipmaps=2, mip_bias=0.05)
    out = R.sample_grid_lut(inp, lut, **kw)
    ig, lg = R.sample_grid_lut_grad(inp, lut, g, **kw)
    assert abs(float((out * g).sum()) - float((inp * ig).sum())) < 1e-9
    # LUT gradient (positions and LoD) by central differences
    eps = 1e-6
    for _ in range(12):
        v, z, y, x, c = (rng.integers(0, s) for s in lut.shape)
        lp, lm = lut.copy(), lut.copy()
        lp[v, z, y, x, c] += eps
        lm[v, z, y, x, c] -= eps
        fd = (float((R.sample_grid_lut(inp, lp, **kw) * g).sum()) - float((R.sample_grid_lut(inp, lm, **kw) * g).sum())
              ) / (2 * eps)
        assert abs(fd - lg[v, z, y, x, c]) < 1e-5 * max(1.0, abs(fd))


def test_transform_mips_exact_grad_is_adjoint():
    rng = np.random.default_rng(6)
    cams = G.look_at_ring_cameras(2, 2.5, 40.0, 1.0, 0.5, 5.0)
    gt = G.grid_transform((8, 8, 8), center=True, normalize="MAX")
    M = np.stack([G.op_matrix(gt)])
    V = np.stack([G.op_matrix(c[0]) for c in cams])
    P = np.stack([G.op_matrix(c[1]) for c in cams])
    F = np.stack([c[2] for c in cams])
    inp = rng.standard_normal((1, 8, 8, 8, 1))
    for mode in (MP.MIP_LINEAR, MP.MIP_NEAREST):
        out = R.sample_grid_transform(inp, M, V, P, F, (6, 5, 4), LINEAR, BORDER, R.TRANSFORM_LINDEPTH, True, mode, 2,
                                      -0.5)
        g = rng.standard_normal(out.shape)
        ig = R.sample_grid_transform_grad(inp, g, M, V, P, F, LINEAR, BORDER, R.TRANSFORM_LINDEPTH, True, "exact",
                                          mode, 2, -0.5)
        assert abs(float((out * g).sum()) - float((inp * ig).sum())) < 1e-9
    with pytest.raises(ValueError):
        R.sample_grid_transform_grad(inp, g, M, V, P, F, LINEAR, BORDER, R.TRANSFORM_LINDEPTH, True, "reference",
                                     MP.MIP_LINEAR, 2, 0.0)
