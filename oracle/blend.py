"""Front-to-back blend reduction along depth and its gradient (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Blend functors (blending.hpp:9-243), applied per ray over cells k = 0..D-1
with accumulator ``acc`` starting at 0:

* C == 1, or mode ADDITIVE: ``acc += cell`` (the generic BlendState; float1
  BEER_LAMBERT / ALPHA are plain sums, blending.hpp:9-19,80,129,179,227);
* BEER_LAMBERT (C in {2,4}, last channel = density d, others = light l):
  ``d = max(d + sigma, 0); L += exp(-d) * l`` (inclusive self-attenuation);
* ALPHA: ``t = 1 - a; L += t*l; a += t*alpha``;
* ALPHA_ADDITIVE: ``t = 1 - alpha; L += t*l; a += alpha``.

``cells`` arrays are ``[..., K, C]`` (K = depth axis, second to last).
"""
import numpy as np

BEER_LAMBERT, ALPHA, ALPHA_ADDITIVE, ADDITIVE = 0, 2, 3, 4


def blend_step(mode, acc, cell):
    C = cell.shape[-1]
    if C == 1 or mode == ADDITIVE:
        return acc + cell
    l, s = cell[..., :-1], cell[..., -1:]
    L, a = acc[..., :-1], acc[..., -1:]
    if mode == BEER_LAMBERT:
        d = np.maximum(a + s, 0.0)
        return np.concatenate([L + np.exp(-d) * l, d], axis=-1)
    if mode == ALPHA:
        t = 1.0 - a
        return np.concatenate([L + t * l, a + t * s], axis=-1)
    if mode == ALPHA_ADDITIVE:
        t = 1.0 - s
        return np.concatenate([L + t * l, a + s], axis=-1)
    raise ValueError(mode)


def blend_scan(cells, mode, mask=None):
    """All prefix accumulators ``[..., K, C]`` (inclusive).  ``mask[..., K]``
    (optional) skips cells (used for rays with fewer samples)."""
    K = cells.shape[-2]
    acc = np.zeros(cells.shape[:-2] + cells.shape[-1:])
    out = np.zeros_like(cells)
    for k in range(K):
        nxt = blend_step(mode, acc, cells[..., k, :])
        if mask is not None:
            nxt = np.where(mask[..., k, None], nxt, acc)
        acc = nxt
        out[..., k, :] = acc
    return out


def reduce_grid_blend(inp, mode, keep_dims=False):
    """ReduceGridBlend: input ``[N, D, H, W, C]`` -> ``[N, H, W, C]`` or the full
    scan ``[N, D, H, W, C]`` (reduce_blend.cu.cc:34-52, reduce_blend.cc:96-101)."""
    cells = np.moveaxis(np.asarray(inp, dtype=np.float64), 1, -2)  # N H W D C
    scan = blend_scan(cells, mode)
    if keep_dims:
        return np.moveaxis(scan, -2, 1)
    return scan[..., -1, :]


def blend_grad_exact(cells, scan, g_out, mode, keep_dims=False, mask=None):
    """Exact adjoint of ``blend_scan`` by reverse-mode over the recurrence.

    g_out: ``[..., C]`` on the final accumulator, or ``[..., K, C]`` on every
    prefix when keep_dims.  The derivative of BEER_LAMBERT's ``max(., 0)`` is
    taken as 1 (the reference's convention, blending.hpp:40-54; identical for
    non-negative densities).
    """
    K = cells.shape[-2]
    C = cells.shape[-1]
    g_cells = np.zeros_like(cells)
    zero = np.zeros(cells.shape[:-2] + (C,))
    gacc = zero.copy() if keep_dims else np.asarray(g_out, dtype=np.float64).copy()
    for k in range(K - 1, -1, -1):
        if keep_dims:
            gacc = gacc + g_out[..., k, :]
        prev = scan[..., k - 1, :] if k > 0 else zero
        if mask is not None and k > 0:
            # previous accumulator of a masked ray = last unmasked prefix
            pass
        cell = cells[..., k, :]
        if C == 1 or mode == ADDITIVE:
            gc = gacc.copy()
            gprev = gacc
        else:
            l, s = cell[..., :-1], cell[..., -1:]
            gL, ga = gacc[..., :-1], gacc[..., -1:]
            Lp, ap = prev[..., :-1], prev[..., -1:]
            if mode == BEER_LAMBERT:
                d = np.maximum(ap + s, 0.0)
                e = np.exp(-d)
                gd = ga - e * np.sum(l * gL, axis=-1, keepdims=True)
                gc = np.concatenate([e * gL, gd], axis=-1)
                gprev = np.concatenate([gL, gd], axis=-1)
            elif mode == ALPHA:
                t = 1.0 - ap
                gc = np.concatenate([t * gL, t * ga], axis=-1)
                gap = (1.0 - s) * ga - np.sum(l * gL, axis=-1, keepdims=True)
                gprev = np.concatenate([gL, gap], axis=-1)
            elif mode == ALPHA_ADDITIVE:
                gc = np.concatenate([(1.0 - s) * gL, ga - np.sum(l * gL, axis=-1, keepdims=True)], axis=-1)
                gprev = gacc
            else:
                raise ValueError(mode)
        if mask is not None:
            m = mask[..., k, None]
            gc = np.where(m, gc, 0.0)
            gprev = np.where(m, gprev, gacc)
        g_cells[..., k, :] = gc
        gacc = gprev
    return g_cells


def blend_grad_reference(cells, g_out, out, mode, keep_dims=False):
    """ReduceGridBlendGrad verbatim (reduce_blend.cu.cc:55-80 with
    blending.hpp blendGradients): reverse loop that reconstructs the previous
    accumulator from the current one (``d -= sigma`` clamped at 0 for
    BEER_LAMBERT; ``a = (1-a)/(1-alpha)`` for ALPHA, whose transmissivity term
    then uses the reconstructed value as in the reference)."""
    K = cells.shape[-2]
    C = cells.shape[-1]
    g_cells = np.zeros_like(cells)
    if keep_dims:
        grads = np.zeros(cells.shape[:-2] + (C,))
        prev_out = np.zeros(cells.shape[:-2] + (C,))
    else:
        grads = np.asarray(g_out, dtype=np.float64).copy()
        prev_out = np.asarray(out, dtype=np.float64).copy()
    for k in range(K - 1, -1, -1):
        if keep_dims:
            grads = grads + g_out[..., k, :]
            prev_out = np.asarray(out[..., k, :], dtype=np.float64).copy()
        cell = cells[..., k, :]
        if C == 1 or mode == ADDITIVE:
            g_cells[..., k, :] = grads
            continue
        l, s = cell[..., :-1], cell[..., -1:]
        gL, ga = grads[..., :-1], grads[..., -1:]
        if mode == BEER_LAMBERT:
            e = np.exp(-prev_out[..., -1:])
            prev_out[..., -1:] = np.maximum(prev_out[..., -1:] - s, 0.0)
            dd = ga - e * np.sum(l * gL, axis=-1, keepdims=True)
            g_cells[..., k, :] = np.concatenate([e * gL, dd], axis=-1)
            grads = np.concatenate([gL, dd], axis=-1)
        elif mode == ALPHA:
            prev_out[..., -1:] = (1.0 - prev_out[..., -1:]) / (1.0 - s)
            t = 1.0 - prev_out[..., -1:]
            dacc = (1.0 - s) * ga - np.sum(l * gL, axis=-1, keepdims=True)
            g_cells[..., k, :] = np.concatenate([t * gL, t * ga], axis=-1)
            grads = np.concatenate([gL, dacc], axis=-1)
        elif mode == ALPHA_ADDITIVE:
            prev_out[..., -1:] = prev_out[..., -1:] - s
            dd = ga - np.sum(l * gL, axis=-1, keepdims=True)
            g_cells[..., k, :] = np.concatenate([(1.0 - s) * gL, dd], axis=-1)
            grads = np.concatenate([gL, ga], axis=-1)
        else:
            raise ValueError(mode)
    return g_cells


def reduce_grid_blend_grad(out_grad, out, inp, mode, keep_dims=False, semantics="exact"):
    cells = np.moveaxis(np.asarray(inp, dtype=np.float64), 1, -2)
    if keep_dims:
        g = np.moveaxis(np.asarray(out_grad, dtype=np.float64), 1, -2)
        o = np.moveaxis(np.asarray(out, dtype=np.float64), 1, -2)
    else:
        g = np.asarray(out_grad, dtype=np.float64)
        o = np.asarray(out, dtype=np.float64)
    if semantics == "reference":
        gc = blend_grad_reference(cells, g, o, mode, keep_dims)
    else:
        scan = blend_scan(cells, mode)
        gc = blend_grad_exact(cells, scan, g, mode, keep_dims)
    return np.moveaxis(gc, -2, 1)
