"""Mip-mapped sampling (FP64): mip generation, level of detail, and the
sampler that reads one or two mip levels, with its exact adjoint.

Test infrastructure only (see ``oracle/__init__.py``).

* ``generate_mips``: level m+1 has floor(dims_m / 2) cells, each the mean of
  its 2x2x2 children, averaged pairwise along z, then y, then x
  (kGenerateMip3D, resample_grid.cu.cc:58-87).
* ``transform_lod``: level of detail of a camera-transform sample
  (samplePos3D<CM, true>, sampling.hpp:151-228): forward differences of the
  sample position over one output cell along x, y and z; the transform modes
  take, per object-space axis, the length of its three differences, the
  reverse modes the length of each difference vector; lod = max(0, log2 of
  the largest).
* ``central_lod``: LodTransform's level of detail (calcLoD<CM>,
  sampling.hpp:72-139): central differences and per-axis lengths in every
  mode.
* ``sample_levels``: Sampler2::sample3D with mips (sampling.hpp:778-862,
  991-1010): lod = clamp(raw + bias, 0, L); NEAREST reads level
  min(int(lod + .5), L); LINEAR reads floor level min(int(lod), L) and, when
  min(ceil(lod), L) is larger, lerps toward it by fract(lod); level m samples
  at position / 2^m minus the cell-centre offset on a grid of dims >> m.
"""
import numpy as np

from . import LINEAR, NEAREST, MIN, MAX, QUADRATIC
from . import sampling as S

MIP_NONE, MIP_NEAREST, MIP_LINEAR = 0, 1, 2
MIPMODE = {"NONE": MIP_NONE, "NEAREST": MIP_NEAREST, "LINEAR": MIP_LINEAR}


def generate_mips(grid, levels):
    """[level 0 = grid, level 1, ..., level L] for grid [D, H, W, C]."""
    out = [np.asarray(grid, dtype=np.float64)]
    for _ in range(levels):
        g = out[-1]
        d, h, w = (s // 2 for s in g.shape[:3])
        g = g[:2 * d, :2 * h, :2 * w]
        a = (g[0::2] + g[1::2]) * 0.5
        b = (a[:, 0::2] + a[:, 1::2]) * 0.5
        out.append((b[:, :, 0::2] + b[:, :, 1::2]) * 0.5)
    return out


def collapse_adjoint(level_grads):
    """Adjoint of ``generate_mips``: every coarse gradient adds 1/8 to each of
    its 8 children, from the coarsest level down; returns level 0."""
    g = [np.array(x, dtype=np.float64) for x in level_grads]
    for m in range(len(g) - 1, 0, -1):
        d, h, w = g[m].shape[:3]
        for a in (0, 1):
            for b in (0, 1):
                for c in (0, 1):
                    g[m - 1][a:2 * d:2, b:2 * h:2, c:2 * w:2] += g[m] * 0.125
    return g[0]


def transform_lod(positions_fn, out_dhw, forward_mode):
    """Raw LoD per output cell.  ``positions_fn(idx)`` maps cell coordinates
    (index + 0.5, xyz) to sample positions; ``forward_mode`` selects the
    per-axis (TRANSFORM*) or per-difference (…_REVERSE) lengths."""
    d, h, w = out_dhw
    z, y, x = np.meshgrid(np.arange(d), np.arange(h), np.arange(w), indexing="ij")
    idx = np.stack([x, y, z], -1).astype(np.float64) + 0.5
    p = positions_fn(idx)
    diffs = [p - positions_fn(idx + e) for e in np.eye(3)]
    if forward_mode:
        lens = [np.sqrt(diffs[0][..., a] ** 2 + diffs[1][..., a] ** 2 + diffs[2][..., a] ** 2) for a in range(3)]
    else:
        lens = [np.sqrt(np.sum(dd * dd, -1)) for dd in diffs]
    dmax = np.maximum(np.maximum(lens[0], lens[1]), lens[2])
    with np.errstate(divide="ignore"):
        return np.maximum(0.0, np.log2(dmax))


def level_weights(lod_raw, mip_mode, levels, bias):
    """[(level, weight array)] per cell, and dweight/dlod of the ceil level
    (0 where the lod is clamped or a single level is read)."""
    lod = np.clip(lod_raw + bias, 0.0, float(levels))
    if mip_mode == MIP_NEAREST:
        lf = np.minimum((lod + 0.5).astype(np.int64), levels)
        return lf, None, np.ones_like(lod), np.zeros_like(lod), np.zeros_like(lod)
    lf = np.minimum(lod.astype(np.int64), levels)
    lc = np.minimum(np.ceil(lod).astype(np.int64), levels)
    frac = lod - np.floor(lod)
    two = lc > lf
    wc = np.where(two, frac, 0.0)
    wf = 1.0 - wc
    inside = (lod_raw + bias > 0.0) & (lod_raw + bias < float(levels))
    dw = np.where(two & inside, 1.0, 0.0)
    return lf, lc, wf, wc, dw


def _sample_at(levels_data, m, pos, cco, filter_mode, boundary, sampler):
    return sampler(levels_data[m], pos / float(1 << m) - cco, filter_mode, boundary)


def sample_levels(levels_data, pos, lod_raw, mip_mode, bias, cco, filter_mode, boundary, sampler=None):
    """Mip-mapped read of ``pos`` (sample positions before the cell-centre
    offset, [..., 3]) with per-sample raw LoD."""
    sampler = sampler or S.sample
    L = len(levels_data) - 1
    if mip_mode == MIP_NONE:
        return sampler(levels_data[0], pos - cco, filter_mode, boundary)
    lf, lc, wf, wc, _ = level_weights(lod_raw, mip_mode, L, bias)
    C = levels_data[0].shape[-1]
    out = np.zeros(pos.shape[:-1] + (C,))
    for m in range(L + 1):
        sel_f = lf == m
        sel_c = (lc == m) & (wc > 0) if lc is not None else np.zeros_like(sel_f)
        if not (np.any(sel_f) or np.any(sel_c)):
            continue
        val = _sample_at(levels_data, m, pos, cco, filter_mode, boundary, sampler)
        out += np.where(sel_f[..., None], wf[..., None] * val, 0.0)
        out += np.where(sel_c[..., None], wc[..., None] * val, 0.0)
    return out


def scatter_levels(level_shapes, pos, g, lod_raw, mip_mode, bias, cco, filter_mode, boundary, levels_data=None):
    """Exact adjoint of ``sample_levels`` w.r.t. the level data: per-level
    gradients (collapse them with ``collapse_adjoint``)."""
    L = len(level_shapes) - 1
    grads = [np.zeros(s) for s in level_shapes]

    def scat(m, wgt):
        gm = g * wgt[..., None]
        q = pos / float(1 << m) - cco
        if filter_mode == LINEAR:
            S.scatter_linear(level_shapes[m], q, gm, boundary, mask="forward", out=grads[m])
        elif filter_mode == QUADRATIC:
            S.scatter_quadratic(level_shapes[m], q, gm, boundary, mask="forward", out=grads[m])
        elif filter_mode == NEAREST:
            S.scatter_nearest(level_shapes[m], q, gm, boundary, out=grads[m])
        else:
            S.scatter_minmax(levels_data[m], q, gm, boundary, filter_mode == MAX, out=grads[m])

    if mip_mode == MIP_NONE:
        scat(0, np.ones(pos.shape[:-1]))
        return grads
    lf, lc, wf, wc, _ = level_weights(lod_raw, mip_mode, L, bias)
    for m in range(L + 1):
        wsum = np.where(lf == m, wf, 0.0)
        if lc is not None:
            wsum = wsum + np.where((lc == m) & (wc > 0), wc, 0.0)
        if np.any(wsum != 0):
            scat(m, wsum)
    return grads


def position_gradient(levels_data, pos, g, lod_raw, mip_mode, bias, cco, filter_mode, boundary):
    """d <g, sample_levels> / d pos and / d lod_raw.  The position part is
    nonzero for the LINEAR filter only (the other filters are piecewise
    constant in the position); the LoD part is the derivative of the level
    lerp, f(ceil) - f(floor) where two levels are read and the LoD is not
    clamped, for every filter."""
    L = len(levels_data) - 1

    def dgrad(m):
        q = pos / float(1 << m) - cco
        dx, dy, dz = S.data_gradient(levels_data[m], q, boundary)
        s = 1.0 / float(1 << m)
        return np.stack([np.sum(g * dx, -1) * s, np.sum(g * dy, -1) * s, np.sum(g * dz, -1) * s], -1)

    gp = np.zeros(pos.shape)
    glod = np.zeros(pos.shape[:-1])
    if mip_mode == MIP_NONE:
        if filter_mode == LINEAR:
            gp = dgrad(0)
        return gp, glod
    lf, lc, wf, wc, dw = level_weights(lod_raw, mip_mode, L, bias)
    for m in range(L + 1):
        wsum = np.where(lf == m, wf, 0.0)
        if lc is not None:
            wsum = wsum + np.where((lc == m) & (wc > 0), wc, 0.0)
        if filter_mode == LINEAR and np.any(wsum != 0):
            gp += dgrad(m) * wsum[..., None]
    if mip_mode == MIP_LINEAR:
        for m in range(L + 1):
            val = np.sum(g * _sample_at(levels_data, m, pos, cco, filter_mode, boundary, S.sample), -1)
            glod += np.where((lc == m) & (dw > 0), val, 0.0) - np.where((lf == m) & (dw > 0), val, 0.0)
    return gp, glod


def central_lod(positions_fn, out_dhw):
    """LodTransform's level of detail (calcLoD<CM>, sampling.hpp:72-139):
    central differences (p(idx - e) - p(idx + e)) / 2 over one output cell,
    per object-space axis the length of its three differences (every
    coordinate mode), lod = max(0, log2 of the largest)."""
    d, h, w = out_dhw
    z, y, x = np.meshgrid(np.arange(d), np.arange(h), np.arange(w), indexing="ij")
    idx = np.stack([x, y, z], -1).astype(np.float64) + 0.5
    diffs = [(positions_fn(idx - e) - positions_fn(idx + e)) * 0.5 for e in np.eye(3)]
    lens = [np.sqrt(diffs[0][..., a] ** 2 + diffs[1][..., a] ** 2 + diffs[2][..., a] ** 2) for a in range(3)]
    dmax = np.maximum(np.maximum(lens[0], lens[1]), lens[2])
    with np.errstate(divide="ignore"):
        return np.maximum(0.0, np.log2(dmax))
