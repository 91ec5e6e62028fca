"""RaymarchGridTransform (fused frustum sampling + blend) and its gradient (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Per pixel (x, y) of a view with output dims (W, H, D) (raymarch_grid.cu.cc:
173-200): ``start = IDXtoOS(x+.5, y+.5, 0.5)``, ``end = IDXtoOS(x+.5, y+.5,
zMax+.5)`` with ``zMax = max(D-1, 1)`` (projective NDC depth, not linearised),
``step = (end-start)/zMax``, ``steps = D``.  BORDER clips the ray to the grid's
cell-centre AABB (half size dims/2 - 0.5) widened by one cell on both ends,
``steps = int(|end-start| / |step|) + 2``, and a ray that misses yields 0.
Samples ``p_k = start + k*step`` are read at ``p_k - 0.5`` (sampling_v2.hpp:
358-372) and blended front to back.

The reference accumulates ``samplePos += step`` in FP32; this restatement
evaluates ``start + k*step`` directly (the HIP kernel does the same).
"""
import numpy as np

from . import BORDER, LINEAR, NEAREST
from . import sampling as S
from . import geometry as G
from . import blend as B


def ray_setup(cx, out_dhw, in_dhw, boundary):
    """Returns start [H,W,3], end [H,W,3], step [H,W,3], steps [H,W] (int),
    hit [H,W] (bool)."""
    D, H, W = out_dhw
    out_dims = np.array([W, H, D], dtype=np.float64)
    in_dims = np.array([in_dhw[2], in_dhw[1], in_dhw[0]], dtype=np.float64)
    zmax = float(max(D - 1, 1))
    y, x = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    i0 = np.stack([x + 0.5, y + 0.5, np.full(x.shape, 0.5)], axis=-1).astype(np.float64)
    i1 = np.stack([x + 0.5, y + 0.5, np.full(x.shape, zmax + 0.5)], axis=-1).astype(np.float64)
    start = G.idx_to_os(cx, i0, out_dims)
    end = G.idx_to_os(cx, i1, out_dims)
    step = (end - start) / zmax
    steps = np.full((H, W), D, dtype=np.int64)
    hit = np.ones((H, W), dtype=bool)
    if boundary == BORDER:
        seg = end - start
        rdir = seg / np.linalg.norm(seg, axis=-1, keepdims=True)
        half = in_dims * 0.5 - 0.5
        centered = start - in_dims * 0.5
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1.0 / rdir
            n = inv * centered
            k = np.abs(inv) * half
            t1 = -n - k
            t2 = k - n
        # fmaxf / fminf semantics: a NaN operand (axis-parallel ray, 0*inf) is
        # ignored (raymarch_grid.cu.cc:152-153)
        tmin = np.fmax(np.fmax(t1[..., 0], t1[..., 1]), t1[..., 2])
        tmax = np.fmin(np.fmin(t2[..., 0], t2[..., 1]), t2[..., 2])
        hit = ~((tmin > tmax) | (tmax < 0))
        end = start + rdir * (tmax[..., None] + 1.0)
        start = start + rdir * np.maximum(tmin - 1.0, 0.0)[..., None]
        ssize = np.linalg.norm(step, axis=-1)
        length = np.linalg.norm(start - end, axis=-1)
        with np.errstate(invalid="ignore"):
            steps = np.where(hit, np.floor(length / ssize).astype(np.int64) + 2, 0)
    return start, end, step, steps, hit


def _views(N, nV, separate_camera_batch):
    if separate_camera_batch:
        return [(n, v, v) for n in range(N) for v in range(nV)]
    if nV != N:
        raise ValueError("camera batch must match data batch when not using global sampling.")
    return [(n, n, 0) for n in range(N)]


def _ray_samples(grid, start, step, steps, filter_mode, boundary, reverse_from=None):
    """Sample positions [H,W,K,3], mask [H,W,K] and samples [H,W,K,C]."""
    K = int(steps.max()) if steps.size else 0
    k = np.arange(K, dtype=np.float64)
    if reverse_from is None:
        pos = start[..., None, :] + k[None, None, :, None] * step[..., None, :]
    else:
        # reference backward lattice: end - j*step, reordered front to back
        kk = (steps[..., None] - 1 - np.arange(K)[None, None, :]).astype(np.float64)
        pos = reverse_from[..., None, :] - kk[..., None] * step[..., None, :]
    mask = np.arange(K)[None, None, :] < steps[..., None]
    q = pos - 0.5
    samp = S.sample(grid, q, filter_mode, boundary)
    samp = np.where(mask[..., None], samp, 0.0)
    return pos, mask, samp


def raymarch_grid_transform(inp, M, V, P, F, out_dhw, filter_mode=LINEAR, boundary=BORDER,
                            blend_mode=B.BEER_LAMBERT, separate_camera_batch=True):
    """Output ``[N, V, H, W, C]`` (raymarch_grid.cc:273-278)."""
    inp = np.asarray(inp, dtype=np.float64)
    N = inp.shape[0]
    nV = len(V)
    Vo = nV if separate_camera_batch else 1
    D, H, W = out_dhw
    out = np.zeros((N, Vo, H, W, inp.shape[-1]))
    for n, v, vo in _views(N, nV, separate_camera_batch):
        cx = G.CameraXform.from_op_args(M[n], V[v], P[v], F[v])
        start, end, step, steps, hit = ray_setup(cx, out_dhw, inp.shape[1:4], boundary)
        _, mask, samp = _ray_samples(inp[n], start, step, steps, filter_mode, boundary)
        scan = B.blend_scan(samp, blend_mode, mask)
        res = scan[..., -1, :] if scan.shape[-2] > 0 else np.zeros((H, W, inp.shape[-1]))
        out[n, vo] = np.where(hit[..., None], res, 0.0)
    return out


def raymarch_grid_transform_grad(inp, out, out_grad, M, V, P, F, out_dhw, filter_mode=LINEAR, boundary=BORDER,
                                 blend_mode=B.BEER_LAMBERT, separate_camera_batch=True, semantics="exact"):
    """semantics="exact": the true adjoint of ``raymarch_grid_transform``
    (forward lattice, forward BORDER mask).

    semantics="reference": RaymarchGridTransformGrad (raymarch_grid.cu.cc:
    482-724): per ray a reverse walk from ``end`` (the shifted lattice under
    BORDER), accumulator reconstruction by subtraction (blending.hpp),
    trilinear scatter with the reference BORDER mask accumulating the corner
    weights, then per camera ``grad / sum(weights)`` where the sum is > 0,
    summed over cameras.  Only LINEAR scatters (other filters give 0).
    """
    inp = np.asarray(inp, dtype=np.float64)
    N, D_, H_, W_, C = inp.shape
    nV = len(V)
    in_grad = np.zeros_like(inp)
    for n, v, vo in _views(N, nV, separate_camera_batch):
        cx = G.CameraXform.from_op_args(M[n], V[v], P[v], F[v])
        start, end, step, steps, hit = ray_setup(cx, out_dhw, inp.shape[1:4], boundary)
        steps = np.where(hit, steps, 0)
        g = np.asarray(out_grad[n, vo], dtype=np.float64)
        if semantics == "exact":
            pos, mask, samp = _ray_samples(inp[n], start, step, steps, filter_mode, boundary)
            scan = B.blend_scan(samp, blend_mode, mask)
            gc = B.blend_grad_exact(samp, scan, g, blend_mode, mask=mask)
            q = pos[mask] - 0.5
            if filter_mode == LINEAR:
                S.scatter_linear(inp.shape[1:], q, gc[mask], boundary, mask="forward", out=in_grad[n])
            elif filter_mode == NEAREST:
                S.scatter_nearest(inp.shape[1:], q, gc[mask], boundary, out=in_grad[n])
            else:
                S.scatter_minmax(inp[n], q, gc[mask], boundary, filter_mode == 3, out=in_grad[n])
        else:
            if filter_mode != LINEAR:
                continue
            o = np.asarray(out[n, vo], dtype=np.float64)
            cam_grad = np.zeros_like(inp[n])
            wsum = np.zeros(inp.shape[1:4])
            # group rays by sample count so the reverse loop sees no padding
            for K in np.unique(steps[hit]):
                sel = hit & (steps == K)
                kk = (K - 1 - np.arange(K)).astype(np.float64)
                pos = end[sel][:, None, :] - kk[None, :, None] * step[sel][:, None, :]
                samp = S.sample(inp[n], pos - 0.5, filter_mode, boundary)
                gc = B.blend_grad_reference(samp, g[sel], o[sel], blend_mode)
                _scatter_weighted(inp.shape[1:], (pos - 0.5).reshape(-1, 3), gc.reshape(-1, C), boundary,
                                  cam_grad, wsum)
            nz = wsum > 0
            cam_grad[nz] = cam_grad[nz] / wsum[nz][:, None]
            in_grad[n] += cam_grad
    return in_grad


def _scatter_weighted(grid_shape, q, g, boundary, out, wsum):
    """Trilinear scatter with the reference BORDER mask, also accumulating the
    corner weights (sampling_v2.hpp:461-498)."""
    dims = S.dims_xyz(grid_shape)
    f = np.floor(q).astype(np.int64)
    c = np.ceil(q).astype(np.int64)
    cw = q - np.floor(q)
    fw = 1.0 - cw
    if boundary == BORDER:
        pos_ok = np.all((q > -1.0) & (q < dims), axis=-1)
    else:
        pos_ok = np.ones(q.shape[:-1], dtype=bool)
        from .sampling import _fix_index
        f = _fix_index(f, dims, boundary)
        c = _fix_index(c, dims, boundary)
    for bz in (0, 1):
        for by in (0, 1):
            for bx in (0, 1):
                ix = (c if bx else f)[..., 0]
                iy = (c if by else f)[..., 1]
                iz = (c if bz else f)[..., 2]
                wt = (cw if bx else fw)[..., 0] * (cw if by else fw)[..., 1] * (cw if bz else fw)[..., 2]
                valid = pos_ok & (ix >= 0) & (ix < dims[0]) & (iy >= 0) & (iy < dims[1]) & (iz >= 0) & (iz < dims[2])
                np.add.at(out, (iz[valid], iy[valid], ix[valid]), g[valid] * wt[valid][:, None])
                np.add.at(wsum, (iz[valid], iy[valid], ix[valid]), wt[valid])
