"""Per-position grid reads, data gradients and scatter adjoints (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Conventions (all follow the reference):
  * a grid is an array ``[D, H, W, C]``; x is the fastest axis
    (vector_io.inl:21-24 ``flat = X*(Y*z + y) + x``);
  * a position ``q`` is ``[..., 3]`` in (x, y, z) order and in
    *integer-centred* cell units: cell ``i`` has its centre at ``q = i``.
    Callers subtract the cell-centre offset before calling (sampling.hpp:
    802-826 ``position - settings.cellCenterOffset``);
  * corners are ``floor(q)`` and ``ceil(q)`` (so both corners coincide when a
    coordinate is an integer) and the weights are ``fract(q)``
    (sampling.hpp:375-379);
  * interpolation order is x, then y, then z with ``lerp(a,b,t) = a + t(b-a)``
    (sampling.hpp:389-430, vectormath.hpp:255-259).
"""
import numpy as np

from . import BORDER, CLAMP, WRAP

FLOAT_MAX = float(np.finfo(np.float32).max)


def dims_xyz(grid_shape):
    """(W, H, D) of a ``[D, H, W, ...]`` grid shape."""
    return np.array([grid_shape[2], grid_shape[1], grid_shape[0]], dtype=np.int64)


def wrap_index(i, m):
    """Periodic index.  The reference computes ``(i % m) + m*(i < 0)`` with
    C's truncating ``%`` (vectormath.hpp:262-268); that agrees with this
    floor-mod except for ``i = -k*m`` (k >= 1), where the reference produces the
    out-of-range index ``m`` (an out-of-bounds read).  Only reachable for shifts
    of at least a full grid extent."""
    return np.mod(i, m)


def _fix_index(idx, dims, boundary):
    if boundary == CLAMP:
        return np.clip(idx, 0, dims - 1)
    if boundary == WRAP:
        return wrap_index(idx, dims)
    return idx


def _gather(grid, ix, iy, iz, valid):
    """grid[iz, iy, ix, :] where ``valid`` else 0."""
    zi = np.where(valid, iz, 0)
    yi = np.where(valid, iy, 0)
    xi = np.where(valid, ix, 0)
    out = grid[zi, yi, xi]
    return np.where(valid[..., None], out, 0.0)


def border_position_valid(q, dims):
    """Forward BORDER validity: ``0 <= q+0.5 < dims`` on every axis
    (sampling.hpp:381-384)."""
    p = q + 0.5
    return np.all((p >= 0.0) & (p < dims), axis=-1)


def _corners(q, dims, boundary):
    f = np.floor(q).astype(np.int64)
    c = np.ceil(q).astype(np.int64)
    w = q - np.floor(q)
    if boundary == BORDER:
        fv = (f > -1) & (f < dims)
        cv = (c > -1) & (c < dims)
    else:
        fv = np.ones(f.shape, dtype=bool)
        cv = np.ones(c.shape, dtype=bool)
        f = _fix_index(f, dims, boundary)
        c = _fix_index(c, dims, boundary)
    return f, c, w, fv, cv


def _corner_values(grid, f, c, fv, cv):
    """The 8 corner values, keyed by (x-corner, y-corner, z-corner) with
    0 = floor, 1 = ceil."""
    vals = {}
    for bz in (0, 1):
        for by in (0, 1):
            for bx in (0, 1):
                ix = (c if bx else f)[..., 0]
                iy = (c if by else f)[..., 1]
                iz = (c if bz else f)[..., 2]
                v = (cv if bx else fv)[..., 0] & (cv if by else fv)[..., 1] & (cv if bz else fv)[..., 2]
                vals[(bx, by, bz)] = (_gather(grid, ix, iy, iz, v), v, (ix, iy, iz))
    return vals


def _lerp(a, b, t):
    return a + t * (b - a)


def sample_linear(grid, q, boundary):
    """Trilinear read (sampling.hpp:371-465)."""
    dims = dims_xyz(grid.shape)
    f, c, w, fv, cv = _corners(q, dims, boundary)
    vals = _corner_values(grid, f, c, fv, cv)
    wx, wy, wz = (w[..., i:i + 1] for i in range(3))
    v0 = _lerp(_lerp(vals[(0, 0, 0)][0], vals[(1, 0, 0)][0], wx),
               _lerp(vals[(0, 1, 0)][0], vals[(1, 1, 0)][0], wx), wy)
    v1 = _lerp(_lerp(vals[(0, 0, 1)][0], vals[(1, 0, 1)][0], wx),
               _lerp(vals[(0, 1, 1)][0], vals[(1, 1, 1)][0], wx), wy)
    out = _lerp(v0, v1, wz)
    if boundary == BORDER:
        out = np.where(border_position_valid(q, dims)[..., None], out, 0.0)
    return out


def nearest_index(q, dims, boundary):
    """Nearest cell: ``int(q + 0.5)`` truncates toward zero, so BORDER keeps
    ``q in (-1.5, -0.5)`` as cell 0 (sampling.hpp:570-584)."""
    idx = np.trunc(q + 0.5).astype(np.int64)
    if boundary == BORDER:
        valid = np.all((idx >= 0) & (idx < dims), axis=-1)
    else:
        valid = np.ones(idx.shape[:-1], dtype=bool)
        idx = _fix_index(idx, dims, boundary)
    return idx, valid


def sample_nearest(grid, q, boundary):
    dims = dims_xyz(grid.shape)
    idx, valid = nearest_index(q, dims, boundary)
    return _gather(grid, idx[..., 0], idx[..., 1], idx[..., 2], valid)


def sample_minmax(grid, q, boundary, is_max):
    """Min / max over the 8 corners (sampling.hpp:585-704).  BORDER: 0 outside
    the valid range; if any corner is out of the grid the running value starts
    at 0 instead of +-FLOAT_MAX and only in-grid corners are compared."""
    dims = dims_xyz(grid.shape)
    f, c, w, fv, cv = _corners(q, dims, boundary)
    vals = _corner_values(grid, f, c, fv, cv)
    C = grid.shape[-1]
    init = -FLOAT_MAX if is_max else FLOAT_MAX
    acc = np.full(q.shape[:-1] + (C,), init)
    if boundary == BORDER:
        all_valid = np.all(fv & cv, axis=-1)
        acc = np.where(all_valid[..., None], acc, 0.0)
    op = np.maximum if is_max else np.minimum
    for key in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)]:
        v, valid, _ = vals[key]
        acc = np.where(valid[..., None], op(acc, v), acc)
    if boundary == BORDER:
        acc = np.where(border_position_valid(q, dims)[..., None], acc, 0.0)
    return acc


def bspline2_weight(dist):
    """Quadratic B-spline weight of a cell at distance ``dist`` from the
    sample (sampling.hpp:247-256): 3/4 - d^2 below 1/2, (3/2 - d)^2 / 2 below
    3/2, else 0."""
    d = np.asarray(dist, dtype=np.float64)
    return np.where(d < 0.5, 0.75 - d * d, np.where(d < 1.5, 0.5 * (1.5 - d) ** 2, 0.0))


def quadratic_stencil(q, dims):
    """The 3x3x3 stencil of ``read3DQuadratic`` (sampling.hpp:268-283):
    per axis the (unfixed) indices centre-1, centre, centre+1 with
    centre = floor(q + 0.5), and their weights from the distances |q - i|."""
    centre = np.floor(q + 0.5).astype(np.int64)
    idx = [centre - 1, centre, centre + 1]
    w = [bspline2_weight(np.abs(q - i)) for i in idx]
    return idx, w


def sample_quadratic(grid, q, boundary):
    """read3DQuadratic (sampling.hpp:268-365): BORDER returns 0 everywhere
    (the reference's early return, :284-287); CLAMP / WRAP fix the indices
    and interpolate along x, then y, then z."""
    dims = dims_xyz(grid.shape)
    C = grid.shape[-1]
    if boundary == BORDER:
        return np.zeros(q.shape[:-1] + (C,))
    idx, w = quadratic_stencil(q, dims)
    idx = [_fix_index(i, dims, boundary) for i in idx]
    ones = np.ones(q.shape[:-1], dtype=bool)
    out = 0.0
    for cz in range(3):
        vz = 0.0
        for cy in range(3):
            vy = 0.0
            for cx in range(3):
                val = _gather(grid, idx[cx][..., 0], idx[cy][..., 1], idx[cz][..., 2], ones)
                vy = vy + w[cx][..., 0:1] * val
            vz = vz + w[cy][..., 1:2] * vy
        out = out + w[cz][..., 2:3] * vz
    return out


def scatter_quadratic(grid_shape, q, g, boundary, mask="forward", count=None, out=None):
    """Adjoint of ``sample_quadratic`` w.r.t. the grid values.

    mask="forward": the exact adjoint (nothing for BORDER, whose forward is 0).
    mask="reference": scatterGradQuadratic (resample_grid.cu.cc:904-1080):
      BORDER positions with -1.5 < q < dims scatter to the stencil cells whose
      'prev' indices are >= 0 and 'next' indices < dims, skipping the
      centre-centre-centre cell; its 'centre' indices are not checked there,
      which writes out of range for q+0.5 < 0 or q+0.5 >= dims, so here
      those cells are skipped.  ``count`` receives one hit per scattered cell.
    """
    D, H, W = grid_shape[:3]
    C = g.shape[-1]
    if out is None:
        out = np.zeros((D, H, W, C))
    dims = dims_xyz(grid_shape)
    if boundary == BORDER and mask == "forward":
        return out
    idx, w = quadratic_stencil(q, dims)
    if boundary == BORDER:
        pos_ok = np.all((q > -1.5) & (q < dims), axis=-1)
    else:
        pos_ok = np.ones(q.shape[:-1], dtype=bool)
        idx = [_fix_index(i, dims, boundary) for i in idx]
    for cz in range(3):
        for cy in range(3):
            for cx in range(3):
                ix, iy, iz = idx[cx][..., 0], idx[cy][..., 1], idx[cz][..., 2]
                v = pos_ok.copy()
                if boundary == BORDER:
                    if (cx, cy, cz) == (1, 1, 1):
                        continue
                    for c_, i_, m in ((cx, ix, dims[0]), (cy, iy, dims[1]), (cz, iz, dims[2])):
                        if c_ == 0:
                            v &= i_ >= 0
                        elif c_ == 2:
                            v &= i_ < m
                        else:
                            v &= (i_ >= 0) & (i_ < m)
                wt = w[cx][..., 0] * w[cy][..., 1] * w[cz][..., 2]
                _accumulate(out, count, ix, iy, iz, v, g * wt[..., None])
    return out


def sample(grid, q, filter_mode, boundary):
    from . import NEAREST, LINEAR, MIN, MAX, QUADRATIC
    if filter_mode == LINEAR:
        return sample_linear(grid, q, boundary)
    if filter_mode == QUADRATIC:
        return sample_quadratic(grid, q, boundary)
    if filter_mode == NEAREST:
        return sample_nearest(grid, q, boundary)
    if filter_mode == MIN:
        return sample_minmax(grid, q, boundary, False)
    if filter_mode == MAX:
        return sample_minmax(grid, q, boundary, True)
    raise ValueError(filter_mode)


def data_gradient(grid, q, boundary):
    """d sample_linear / d q per channel, as (dx, dy, dz) each ``[..., C]``
    (sampling.hpp:470-565).  Note the differences of coinciding floor/ceil
    corners vanish, so the derivative along an axis is 0 at integer
    coordinates, exactly as in the reference."""
    dims = dims_xyz(grid.shape)
    f, c, w, fv, cv = _corners(q, dims, boundary)
    vals = _corner_values(grid, f, c, fv, cv)
    g = {k: v[0] for k, v in vals.items()}
    wx, wy, wz = (w[..., i:i + 1] for i in range(3))
    dx = _lerp(_lerp(g[1, 0, 0] - g[0, 0, 0], g[1, 1, 0] - g[0, 1, 0], wy),
               _lerp(g[1, 0, 1] - g[0, 0, 1], g[1, 1, 1] - g[0, 1, 1], wy), wz)
    dy = _lerp(_lerp(g[0, 1, 0] - g[0, 0, 0], g[1, 1, 0] - g[1, 0, 0], wx),
               _lerp(g[0, 1, 1] - g[0, 0, 1], g[1, 1, 1] - g[1, 0, 1], wx), wz)
    dz = _lerp(_lerp(g[0, 0, 1] - g[0, 0, 0], g[1, 0, 1] - g[1, 0, 0], wx),
               _lerp(g[0, 1, 1] - g[0, 1, 0], g[1, 1, 1] - g[1, 1, 0], wx), wy)
    if boundary == BORDER:
        m = border_position_valid(q, dims)[..., None]
        dx, dy, dz = (np.where(m, d, 0.0) for d in (dx, dy, dz))
    return dx, dy, dz


def _accumulate(out, count, ix, iy, iz, valid, contrib):
    zi, yi, xi = iz[valid], iy[valid], ix[valid]
    np.add.at(out, (zi, yi, xi), contrib[valid])
    if count is not None:
        np.add.at(count, (zi, yi, xi), 1)


def scatter_linear(grid_shape, q, g, boundary, mask="forward", count=None, out=None):
    """Adjoint of ``sample_linear`` w.r.t. the grid values: scatter ``g`` with
    the trilinear weights to the 8 corners.

    mask="forward": the exact adjoint, BORDER positions outside
      ``[-0.5, dims-0.5)`` contribute nothing (the forward returned 0).
    mask="reference": the reference's wider acceptance ``-1 < q < dims``
      (resample_grid.cu.cc:1082, sampling_v2.hpp:475).
    ``count`` (int array [D,H,W]) receives one hit per in-grid corner,
    duplicates included (resample_grid.cu.cc:1097-1129).
    """
    D, H, W = grid_shape[:3]
    C = g.shape[-1]
    if out is None:
        out = np.zeros((D, H, W, C))
    dims = dims_xyz(grid_shape)
    f, c, w, fv, cv = _corners(q, dims, boundary)
    if boundary == BORDER:
        if mask == "forward":
            pos_ok = border_position_valid(q, dims)
        else:
            pos_ok = np.all((q > -1.0) & (q < dims), axis=-1)
    else:
        pos_ok = np.ones(q.shape[:-1], dtype=bool)
    cw = w
    fw = 1.0 - w
    for bz in (0, 1):
        for by in (0, 1):
            for bx in (0, 1):
                ix = (c if bx else f)[..., 0]
                iy = (c if by else f)[..., 1]
                iz = (c if bz else f)[..., 2]
                v = (cv if bx else fv)[..., 0] & (cv if by else fv)[..., 1] & (cv if bz else fv)[..., 2]
                v = v & pos_ok
                wt = (cw if bx else fw)[..., 0] * (cw if by else fw)[..., 1] * (cw if bz else fw)[..., 2]
                _accumulate(out, count, ix, iy, iz, v, g * wt[..., None])
    return out


def scatter_nearest(grid_shape, q, g, boundary, out=None):
    """Exact adjoint of ``sample_nearest``."""
    D, H, W = grid_shape[:3]
    if out is None:
        out = np.zeros((D, H, W, g.shape[-1]))
    idx, valid = nearest_index(q, dims_xyz(grid_shape), boundary)
    _accumulate(out, None, idx[..., 0], idx[..., 1], idx[..., 2], valid, g)
    return out


def scatter_minmax(grid, q, g, boundary, is_max, out=None):
    """Exact adjoint of ``sample_minmax``: per channel the gradient goes to the
    first corner (evaluation order x, y, z fastest-first) that attains the
    selected value; nothing if the BORDER start value 0 won."""
    dims = dims_xyz(grid.shape)
    if out is None:
        out = np.zeros(grid.shape)
    f, c, w, fv, cv = _corners(q, dims, boundary)
    vals = _corner_values(grid, f, c, fv, cv)
    C = grid.shape[-1]
    init = -FLOAT_MAX if is_max else FLOAT_MAX
    acc = np.full(q.shape[:-1] + (C,), init)
    if boundary == BORDER:
        all_valid = np.all(fv & cv, axis=-1)
        acc = np.where(all_valid[..., None], acc, 0.0)
    sel = np.full(q.shape[:-1] + (C,), -1, dtype=np.int64)
    order = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)]
    for k, key in enumerate(order):
        v, valid, _ = vals[key]
        better = (v > acc) if is_max else (v < acc)
        upd = valid[..., None] & better
        acc = np.where(upd, v, acc)
        sel = np.where(upd, k, sel)
    posok = border_position_valid(q, dims) if boundary == BORDER else np.ones(q.shape[:-1], dtype=bool)
    for k, key in enumerate(order):
        _, valid, (ix, iy, iz) = vals[key]
        for ch in range(C):
            m = (sel[..., ch] == k) & posok
            zi, yi, xi = iz[m], iy[m], ix[m]
            np.add.at(out[..., ch], (zi, yi, xi), g[..., ch][m])
    return out
