"""Pins oracle/geometry.py, resample.py, blend.py and raymarch.py against closed
forms, round trips, finite differences and adjoint identities."""
import numpy as np
import pytest

from oracle import BORDER, CLAMP, WRAP, LINEAR, NEAREST
from oracle import geometry as G
from oracle import resample as R
from oracle import blend as B
from oracle import raymarch as RM


# --- geometry -----------------------------------------------------------------

def test_projection_closed_form():
    n, f = 1.0, 10.0
    trbl = G.set_fov(90.0, 2.0, n)
    assert trbl == pytest.approx([0.5, 1.0, -0.5, -1.0])
    P = G.projection(n, f, trbl)
    # a point on the near plane at the right edge maps to ndc (1, 0, -1)
    p = P @ np.array([1.0, 0.0, -1.0, 1.0])
    np.testing.assert_allclose(p[:3] / p[3], [1.0, 0.0, -1.0], atol=1e-12)
    p = P @ np.array([0.0, 5.0, -10.0, 1.0])
    np.testing.assert_allclose(p[:3] / p[3], [0.0, 1.0, 1.0], atol=1e-12)


def test_op_matrix_roundtrip_is_transpose():
    A = np.arange(16.0).reshape(4, 4)
    flat = G.op_matrix(A)
    # the ops read 16 floats column-major: element (r,c) at index 4c+r
    assert flat[4 * 3 + 1] == A[1, 3]
    np.testing.assert_array_equal(G.math_matrix(flat), A)


def _ring_cam(dims=(16, 16, 16), out=(12, 10, 14), idx=0, views=5, dist=3.0):
    M = G.grid_transform(dims, center=True, scale_=[1.0 / dims[2]] * 3)
    cams = G.look_at_ring_cameras(views, dist, 40.0, out[2] / out[1], dist - 1.0, dist + 1.0)
    V, P, F = cams[idx]
    return M, V, P, F


def test_ring_camera_looks_at_grid_center():
    M, V, P, F = _ring_cam()
    cx = G.CameraXform(M, V, P, F)
    center_os = np.array([8.0, 8.0, 8.0])
    ndc = G.os_to_ndc(cx, center_os[None])[0]
    np.testing.assert_allclose(ndc[:2], [0.0, 0.0], atol=1e-12)


@pytest.mark.parametrize("lin", [False, True])
def test_idx_os_roundtrip(lin):
    M, V, P, F = _ring_cam(idx=2)
    cx = G.CameraXform(M, V, P, F)
    out_dims = np.array([14.0, 10.0, 12.0])
    rng = np.random.default_rng(0)
    idx = rng.uniform(0, 1, (100, 3)) * out_dims
    if lin:
        os_ = G.idx_to_os_lindepth(cx, idx, out_dims)
        back = G.os_to_idx_lindepth(cx, os_, out_dims)
    else:
        os_ = G.idx_to_os(cx, idx, out_dims)
        back = G.os_to_idx(cx, os_, out_dims)
    np.testing.assert_allclose(back, idx, atol=1e-9)


def test_lindepth_slices_uniform_in_view_depth():
    M, V, P, F = _ring_cam(idx=1)
    cx = G.CameraXform(M, V, P, F)
    D = 12
    out_dims = np.array([14.0, 10.0, float(D)])
    idx = np.stack([np.full(D, 7.0), np.full(D, 5.0), np.arange(D) + 0.5], -1)
    os_ = G.idx_to_os_lindepth(cx, idx, out_dims)
    vs = (cx.MV @ np.concatenate([os_, np.ones((D, 1))], -1).T).T
    t = (np.arange(D) + 0.5) / D
    np.testing.assert_allclose(-vs[:, 2], F[0] + t * (F[1] - F[0]), atol=1e-9)


# --- resample ---------------------------------------------------------------

def test_lut_zero_relative_is_identity():
    rng = np.random.default_rng(1)
    inp = rng.standard_normal((2, 5, 6, 7, 2))
    lut = np.zeros((2, 5, 6, 7, 4))
    out = R.sample_grid_lut(inp, lut, separate_camera_batch=False)
    np.testing.assert_array_equal(out[:, 0], inp)
    out = R.sample_grid_lut(inp, lut, separate_camera_batch=True)
    assert out.shape == (2, 2, 5, 6, 7, 2)
    np.testing.assert_array_equal(out[1, 0], inp[1])


def test_lut_normalized_scales_by_output_dims():
    rng = np.random.default_rng(2)
    inp = rng.standard_normal((1, 6, 6, 6, 1))
    lut = rng.uniform(0, 1, (1, 3, 4, 5, 4))
    a = R.sample_grid_lut(inp, lut, relative=False, normalized=True, cell_center_offset=0.5)
    lut2 = lut.copy()
    lut2[..., :3] *= np.array([5.0, 4.0, 3.0])
    b = R.sample_grid_lut(inp, lut2, relative=False, normalized=False, cell_center_offset=0.5)
    np.testing.assert_allclose(a, b)


@pytest.mark.parametrize("boundary", [BORDER, CLAMP, WRAP])
def test_lut_grad_exact_adjoint_and_fd(boundary):
    rng = np.random.default_rng(3)
    inp = rng.standard_normal((2, 4, 5, 6, 2))
    lut = rng.uniform(-1.3, 1.3, (2, 4, 5, 6, 4))
    lut[..., 3] = 0
    g = rng.standard_normal((2, 2, 4, 5, 6, 2))
    out = R.sample_grid_lut(inp, lut, boundary=boundary)
    gi, gl = R.sample_grid_lut_grad(inp, lut, g, boundary=boundary)
    assert abs(np.sum(out * g) - np.sum(inp * gi)) < 1e-10 * np.abs(out * g).sum()
    # FD w.r.t. a few LUT entries
    h = 1e-6
    for (v, z, y, x, a) in [(0, 1, 2, 3, 0), (1, 2, 1, 4, 1), (0, 3, 4, 1, 2)]:
        lp = lut.copy(); lp[v, z, y, x, a] += h
        lm = lut.copy(); lm[v, z, y, x, a] -= h
        fd = (np.sum(R.sample_grid_lut(inp, lp, boundary=boundary) * g) - np.sum(R.sample_grid_lut(inp, lm, boundary=boundary) * g)) / (2 * h)
        assert gl[v, z, y, x, a] == pytest.approx(fd, abs=1e-5)


def test_lut_grad_reference_count_normalisation():
    inp = np.zeros((1, 3, 3, 3, 1))
    lut = np.zeros((1, 3, 3, 3, 4))
    lut[..., 0] = 0.25  # every sample 0.25 cells right of its own centre
    g = np.ones((1, 1, 3, 3, 3, 1))
    gi, _ = R.sample_grid_lut_grad(inp, lut, g, boundary=CLAMP, separate_camera_batch=False, semantics="reference")
    gx, _ = R.sample_grid_lut_grad(inp, lut, g, boundary=CLAMP, separate_camera_batch=False, semantics="exact")
    # exact: cell x=0 gets 0.75, x=1 gets 0.25+0.75, x=2 gets 0.25+0.75+0.25(clamped ceil of x=2)
    np.testing.assert_allclose(gx[0, 1, 1, :, 0], [0.75, 1.0, 1.25])
    # reference: integer y,z -> each corner hit 4x; counts x=0:4, x=1:8, x=2:12 (clamped duplicates)
    np.testing.assert_allclose(gi[0, 1, 1, :, 0], [0.75 / 4, 1.0 / 8, 1.25 / 12])


def _ortho_identity_setup(d, h, w):
    """Orthographic camera whose frustum grid coincides with the volume grid:
    camera at z>0 looking down -z, frustum [-w/2,w/2]x[-h/2,h/2], depth
    [1, 1+d]; model centres the grid at (0,0,-(1+d/2))."""
    M = G.translation([0, 0, -(1.0 + d / 2.0)]) @ G.translation([-w / 2.0, -h / 2.0, -d / 2.0])
    V = np.eye(4)
    trbl = [h / 2.0, w / 2.0, -h / 2.0, -w / 2.0]
    P = G.projection(1.0, 1.0 + d, trbl, perspective=False)
    F = G.frustum_params(1.0, 1.0 + d, trbl)
    return M, V, P, F


def test_transform_ortho_is_flip_copy():
    """TRANSFORM with an orthographic frustum matching the grid reproduces the
    grid with z flipped (NDC -1 is the near plane, at the grid's max z)."""
    d, h, w = 4, 5, 6
    M, V, P, F = _ortho_identity_setup(d, h, w)
    rng = np.random.default_rng(4)
    inp = rng.standard_normal((1, d, h, w, 1))
    out = R.sample_grid_transform(inp, [G.op_matrix(M)], [G.op_matrix(V)], [G.op_matrix(P)], [F], (d, h, w),
                                  coordinate_mode=R.TRANSFORM)
    np.testing.assert_allclose(out[0, 0], inp[0, ::-1], atol=1e-9)
    # (TRANSFORM_LINDEPTH assumes a perspective matrix: ndc.z = -P22 - P23/z_view)


def test_transform_grad_adjoint():
    rng = np.random.default_rng(5)
    dims = (8, 8, 8)
    M = G.grid_transform(dims, center=True, scale_=[1.0 / 8] * 3)
    cams = G.look_at_ring_cameras(3, 3.0, 40.0, 1.2, 2.0, 4.0)
    Ms = [G.op_matrix(M)]
    Vs = [G.op_matrix(c[0]) for c in cams]
    Ps = [G.op_matrix(c[1]) for c in cams]
    Fs = [c[2] for c in cams]
    inp = rng.standard_normal((1,) + dims + (2,))
    out = R.sample_grid_transform(inp, Ms, Vs, Ps, Fs, (6, 5, 6))
    g = rng.standard_normal(out.shape)
    gi = R.sample_grid_transform_grad(inp, g, Ms, Vs, Ps, Fs)
    assert abs(np.sum(out * g) - np.sum(inp * gi)) < 1e-10 * np.abs(out * g).sum()


# --- blend ---------------------------------------------------------------------

def test_blend_closed_forms():
    D, s, l = 7, 0.3, np.array([0.2, 0.5, 0.9])
    col = np.zeros((1, D, 1, 1, 4))
    col[..., :3] = l
    col[..., 3] = s
    bl = B.reduce_grid_blend(col, B.BEER_LAMBERT)[0, 0, 0]
    k = np.arange(1, D + 1)
    np.testing.assert_allclose(bl[:3], l * np.sum(np.exp(-k * s)))
    assert bl[3] == pytest.approx(D * s)
    np.testing.assert_allclose(B.reduce_grid_blend(col, B.ADDITIVE)[0, 0, 0], D * col[0, 0, 0, 0])
    a = B.reduce_grid_blend(col, B.ALPHA)[0, 0, 0]
    assert a[3] == pytest.approx(1 - (1 - s) ** D)
    np.testing.assert_allclose(a[:3], l * np.sum((1 - s) ** np.arange(D)))
    aa = B.reduce_grid_blend(col, B.ALPHA_ADDITIVE)[0, 0, 0]
    np.testing.assert_allclose(aa, np.concatenate([D * (1 - s) * l, [D * s]]))
    # single channel: plain sum for every mode
    c1 = col[..., 3:]
    for m in (B.BEER_LAMBERT, B.ALPHA, B.ALPHA_ADDITIVE, B.ADDITIVE):
        assert B.reduce_grid_blend(c1, m)[0, 0, 0, 0] == pytest.approx(D * s)


@pytest.mark.parametrize("mode", [B.BEER_LAMBERT, B.ALPHA, B.ALPHA_ADDITIVE, B.ADDITIVE])
@pytest.mark.parametrize("C", [1, 2, 4])
@pytest.mark.parametrize("keep", [False, True])
def test_blend_grad_fd(mode, C, keep):
    rng = np.random.default_rng(6)
    x = rng.uniform(0.05, 0.4, (1, 6, 2, 3, C))
    out = B.reduce_grid_blend(x, mode, keep)
    g = rng.standard_normal(out.shape)
    gx = B.reduce_grid_blend_grad(g, out, x, mode, keep)
    h = 1e-6
    for idx in [(0, 0, 0, 0, 0), (0, 3, 1, 2, C - 1), (0, 5, 1, 1, 0)]:
        xp = x.copy(); xp[idx] += h
        xm = x.copy(); xm[idx] -= h
        fd = (np.sum(B.reduce_grid_blend(xp, mode, keep) * g) - np.sum(B.reduce_grid_blend(xm, mode, keep) * g)) / (2 * h)
        assert gx[idx] == pytest.approx(fd, rel=1e-6, abs=1e-8)


@pytest.mark.parametrize("mode", [B.BEER_LAMBERT, B.ALPHA_ADDITIVE, B.ADDITIVE])
def test_blend_reference_grad_equals_exact(mode):
    rng = np.random.default_rng(7)
    x = rng.uniform(0.05, 0.4, (2, 9, 2, 3, 4))
    for keep in (False, True):
        out = B.reduce_grid_blend(x, mode, keep)
        g = rng.standard_normal(out.shape)
        np.testing.assert_allclose(B.reduce_grid_blend_grad(g, out, x, mode, keep, "reference"),
                                   B.reduce_grid_blend_grad(g, out, x, mode, keep, "exact"), atol=1e-12)


def test_blend_reference_alpha_quirk():
    """The reference ALPHA backward weights the cell terms with the
    reconstructed accumulated alpha instead of the transmissivity 1-a."""
    x = np.zeros((1, 1, 1, 1, 2))
    x[..., 0] = 0.5
    x[..., 1] = 0.25
    out = B.reduce_grid_blend(x, B.ALPHA)
    g = np.ones_like(out)
    ref = B.reduce_grid_blend_grad(g, out, x, B.ALPHA, semantics="reference")
    exact = B.reduce_grid_blend_grad(g, out, x, B.ALPHA)
    np.testing.assert_allclose(exact[0, 0, 0, 0], [1.0, 1.0])  # t = 1 - a_prev = 1
    np.testing.assert_allclose(ref[0, 0, 0, 0], [0.0, 0.0])    # t = a_prev = 0


# --- raymarch ------------------------------------------------------------------

def test_raymarch_ortho_additive_is_column_sum():
    d, h, w = 5, 4, 6
    M, V, P, F = _ortho_identity_setup(d, h, w)
    rng = np.random.default_rng(8)
    inp = rng.standard_normal((1, d, h, w, 2))
    args = ([G.op_matrix(M)], [G.op_matrix(V)], [G.op_matrix(P)], [F], (d, h, w))
    out = RM.raymarch_grid_transform(inp, *args, boundary=CLAMP, blend_mode=B.ADDITIVE)
    np.testing.assert_allclose(out[0, 0], inp[0].sum(0), atol=1e-9)
    # BORDER clips to the AABB +-1 cell: the extra samples read the zero border
    out = RM.raymarch_grid_transform(inp, *args, boundary=BORDER, blend_mode=B.ADDITIVE)
    np.testing.assert_allclose(out[0, 0], inp[0].sum(0), atol=1e-9)


def _ring_args(dims, out_dhw, views=3):
    M = G.grid_transform(dims, center=True, scale_=[1.0 / dims[2]] * 3)
    cams = G.look_at_ring_cameras(views, 3.0, 40.0, out_dhw[2] / out_dhw[1], 2.0, 4.0)
    return ([G.op_matrix(M)], [G.op_matrix(c[0]) for c in cams], [G.op_matrix(c[1]) for c in cams],
            [c[2] for c in cams], out_dhw)


@pytest.mark.parametrize("boundary", [BORDER, CLAMP])
def test_raymarch_additive_adjoint(boundary):
    rng = np.random.default_rng(9)
    dims = (8, 8, 8)
    args = _ring_args(dims, (10, 6, 7))
    inp = rng.standard_normal((1,) + dims + (1,))
    out = RM.raymarch_grid_transform(inp, *args, boundary=boundary, blend_mode=B.ADDITIVE)
    g = rng.standard_normal(out.shape)
    gi = RM.raymarch_grid_transform_grad(inp, out, g, *args, boundary=boundary, blend_mode=B.ADDITIVE)
    assert abs(np.sum(out * g) - np.sum(inp * gi)) < 1e-10 * np.abs(out * g).sum()


@pytest.mark.parametrize("mode", [B.BEER_LAMBERT, B.ALPHA, B.ALPHA_ADDITIVE])
def test_raymarch_grad_fd(mode):
    rng = np.random.default_rng(10)
    dims = (6, 6, 6)
    args = _ring_args(dims, (8, 5, 6), views=2)
    inp = rng.uniform(0.0, 0.3, (1,) + dims + (4,))
    out = RM.raymarch_grid_transform(inp, *args, blend_mode=mode)
    g = rng.standard_normal(out.shape)
    gi = RM.raymarch_grid_transform_grad(inp, out, g, *args, blend_mode=mode)
    h = 1e-6
    for idx in [(0, 3, 3, 3, 0), (0, 2, 3, 4, 3), (0, 3, 2, 2, 3)]:
        xp = inp.copy(); xp[idx] += h
        xm = inp.copy(); xm[idx] -= h
        fd = (np.sum(RM.raymarch_grid_transform(xp, *args, blend_mode=mode) * g)
              - np.sum(RM.raymarch_grid_transform(xm, *args, blend_mode=mode) * g)) / (2 * h)
        assert gi[idx] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_raymarch_reference_grad_normalisation_ortho():
    """Orthographic column through cell centres with CLAMP (same lattice both
    ways): every sample sits on a cell centre with weight 1, so the reference
    gradient of ADDITIVE is g / (number of rays through the cell) = g."""
    d, h, w = 4, 3, 3
    M, V, P, F = _ortho_identity_setup(d, h, w)
    args = ([G.op_matrix(M)], [G.op_matrix(V)], [G.op_matrix(P)], [F], (d, h, w))
    inp = np.ones((1, d, h, w, 1))
    out = RM.raymarch_grid_transform(inp, *args, boundary=CLAMP, blend_mode=B.ADDITIVE)
    g = np.arange(h * w, dtype=float).reshape(1, 1, h, w, 1)
    ref = RM.raymarch_grid_transform_grad(inp, out, g, *args, boundary=CLAMP, blend_mode=B.ADDITIVE, semantics="reference")
    exact = RM.raymarch_grid_transform_grad(inp, out, g, *args, boundary=CLAMP, blend_mode=B.ADDITIVE)
    for z in range(d):
        np.testing.assert_allclose(ref[0, z], g[0, 0], atol=1e-9)
        np.testing.assert_allclose(exact[0, z], g[0, 0], atol=1e-9)


# --- BORDER ray clip, pinned by an independent box intersection -----------------

def _box_hits(origin, direction, lo, hi):
    """Entry / exit ray parameters of the box [lo, hi] found by intersecting
    the ray with each of the six face planes and keeping the hits that lie
    on the face (not the slab formula); None when the ray misses."""
    ts = []
    for a in range(3):
        if direction[a] == 0.0:
            continue
        for b in (lo[a], hi[a]):
            t = (b - origin[a]) / direction[a]
            p = origin + t * direction
            others = [o for o in range(3) if o != a]
            if all(lo[o] - 1e-9 <= p[o] <= hi[o] + 1e-9 for o in others):
                ts.append(t)
    if not ts:
        return None
    return min(ts), max(ts)


def _border_band_weight(q, dims):
    """BORDER trilinear read of a constant 1 field at q (per-axis in-grid
    corner weights; 0 outside [-0.5, dims - 0.5))."""
    w = 1.0
    for a in range(3):
        if not (-0.5 <= q[a] < dims[a] - 0.5):
            return 0.0
        w *= min(1.0, q[a] + 1.0, dims[a] - q[a])
    return w


@pytest.mark.parametrize("views,dist,near,dims,out_dhw", [(5, 3.0, 2.0, (12, 10, 14), (20, 9, 16)),
                                                          (3, 1.2, 0.9, (9, 13, 11), (15, 12, 10))])
def test_raymarch_border_clip_independent(views, dist, near, dims, out_dhw):
    """Every pixel of perspective ring cameras (one set far outside the grid,
    one whose near plane cuts the grid, so that rays start inside it): the first sample
    is 1 unit along the normalised ray before the entry into the cell-centre
    box [0.5, dims - 0.5] (or the unclipped start when that entry is less
    than 1 unit ahead), the last sample lies past exit + 1 by at most one
    step, misses are exactly the rays with no face hit ahead, and a constant
    field's ADDITIVE render is the band-weighted count of those
    independently generated positions (raymarch_grid.cu.cc:109-200)."""
    M = G.grid_transform(dims, translation_=(0.013, -0.021, 0.008), center=True, scale_=[1.0 / max(dims)] * 3)
    cams = G.look_at_ring_cameras(views, dist, 50.0, out_dhw[2] / out_dhw[1], near, dist + 1.0,
                                  height=0.05, start_deg=7.0)
    D, H, W = out_dhw
    in_dims = np.array([dims[2], dims[1], dims[0]], dtype=np.float64)
    lo, hi = np.full(3, 0.5), in_dims - 0.5
    zmax = float(max(D - 1, 1))
    const = 0.7
    n_inside = n_miss = 0
    for V, P, F in cams:
        args = ([G.op_matrix(M)], [G.op_matrix(V)], [G.op_matrix(P)], [F], out_dhw)
        cx = G.CameraXform(M, V, P, F)
        start, end, step, steps, hit = RM.ray_setup(cx, out_dhw, dims, BORDER)
        img = RM.raymarch_grid_transform(np.full((1,) + tuple(dims) + (1,), const), *args, boundary=BORDER,
                                         blend_mode=B.ADDITIVE)[0, 0, ..., 0]
        for y in range(H):
            for x in range(W):
                s0 = G.idx_to_os(cx, np.array([[x + 0.5, y + 0.5, 0.5]]), np.array([W, H, D], float))[0]
                e0 = G.idx_to_os(cx, np.array([[x + 0.5, y + 0.5, zmax + 0.5]]), np.array([W, H, D], float))[0]
                st = (e0 - s0) / zmax
                d = (e0 - s0) / np.linalg.norm(e0 - s0)
                box = _box_hits(s0, d, lo, hi)
                if box is None or box[1] < 0.0:
                    n_miss += 1
                    assert not hit[y, x] and img[y, x] == 0.0
                    continue
                assert hit[y, x]
                t_in, t_out = box
                n_inside += t_in < 1.0
                first = s0 + d * max(t_in - 1.0, 0.0)
                np.testing.assert_allclose(start[y, x], first, rtol=0, atol=1e-9)
                np.testing.assert_allclose(step[y, x], st, rtol=0, atol=1e-12)
                last = start[y, x] + (steps[y, x] - 1) * step[y, x]
                exit1 = s0 + d * (t_out + 1.0)
                past = np.dot(last - exit1, d)
                assert 0.0 < past <= np.linalg.norm(st) + 1e-9, (past, np.linalg.norm(st))
                np.testing.assert_allclose(np.cross(last - s0, d), 0.0, atol=1e-9)
                k = np.arange(steps[y, x])
                pos = first[None] + k[:, None] * st[None]
                expect = const * sum(_border_band_weight(p - 0.5, in_dims) for p in pos)
                assert img[y, x] == pytest.approx(expect, rel=1e-12, abs=1e-12)
    assert n_miss > 0
    if dist < 2.0:
        assert n_inside > 0
