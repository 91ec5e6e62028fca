"""SampleGridLut / SampleGridTransform and their gradients (FP64).

Test infrastructure only (see ``oracle/__init__.py``).

Shapes follow the ops (resample_grid.cc:23-172): input ``[N, D, H, W, C]``,
lookup ``[V, D', H', W', 4]`` (x, y, z, LoD), output ``[N, V, D', H', W', C]``
when ``separate_camera_batch`` (the reference's ``globalSampling``) is true,
else ``[N, 1, D', H', W', C]`` with instance n using camera n
(resample_grid.cu.cc:728-739).
"""
import numpy as np

from . import BORDER, LINEAR, NEAREST, MIN, MAX, QUADRATIC
from . import sampling as S
from . import geometry as G
from . import mips as MP


def _index_grid(d, h, w):
    z, y, x = np.meshgrid(np.arange(d), np.arange(h), np.arange(w), indexing="ij")
    return np.stack([x, y, z], axis=-1).astype(np.float64)


def _pairs(N, V, separate_camera_batch):
    if separate_camera_batch:
        return [(n, v, v) for n in range(N) for v in range(V)]
    if V != N:
        raise ValueError("camera batch must match data batch when not using global sampling.")
    return [(n, n, 0) for n in range(N)]


def lut_positions(lut_v, relative, normalized):
    """Sampling position from one LUT slice (resample_grid.cu.cc:610-618):
    ``normalized`` scales by the *output* dims, ``relative`` adds the integer
    output index."""
    d, h, w = lut_v.shape[:3]
    pos = lut_v[..., :3].astype(np.float64).copy()
    if normalized:
        pos = pos * np.array([w, h, d], dtype=np.float64)
    if relative:
        pos = pos + _index_grid(d, h, w)
    return pos


def sample_grid_lut(inp, lut, filter_mode=LINEAR, boundary=BORDER, cell_center_offset=0.0,
                    separate_camera_batch=True, relative=True, normalized=False, mipmapping=MP.MIP_NONE,
                    num_mipmaps=0, mip_bias=0.0):
    """SampleGridLut (kSampleGrid3DLuT, resample_grid.cu.cc:604-625).  With
    ``mipmapping`` NEAREST / LINEAR the LUT's 4th channel is the raw level of
    detail (sampling.hpp:939-1010)."""
    inp = np.asarray(inp, dtype=np.float64)
    lut = np.asarray(lut, dtype=np.float64)
    N = inp.shape[0]
    V = lut.shape[0]
    Vo = V if separate_camera_batch else 1
    out = np.zeros((N, Vo) + lut.shape[1:4] + (inp.shape[-1],))
    levels = [MP.generate_mips(inp[n], num_mipmaps if mipmapping != MP.MIP_NONE else 0) for n in range(N)]
    for n, v, vo in _pairs(N, V, separate_camera_batch):
        pos = lut_positions(lut[v], relative, normalized)
        out[n, vo] = MP.sample_levels(levels[n], pos, lut[v][..., 3], mipmapping, mip_bias, cell_center_offset,
                                      filter_mode, boundary)
    return out


def sample_grid_lut_grad(inp, lut, out_grad, filter_mode=LINEAR, boundary=BORDER, cell_center_offset=0.0,
                         separate_camera_batch=True, relative=True, normalized=False, semantics="exact",
                         mipmapping=MP.MIP_NONE, num_mipmaps=0, mip_bias=0.0):
    """Returns (input_grad, lookup_grad).

    semantics="exact": the true adjoint of ``sample_grid_lut``.  The LUT
      gradient is accumulated over the data batch and carries the chain-rule
      factor of the ``normalized`` scaling; NEAREST / MIN / MAX route the input
      gradient to the cell the forward read from.
    semantics="reference": SampleGridLutGrad (resample_grid.cu.cc:1420-1665):
      input gradient = trilinear scatter with the reference's BORDER mask for
      every filter, divided per voxel by the number of corner hits summed over
      the cameras of that batch entry; LUT gradient (LINEAR only) written, not
      accumulated, per camera, so with a data batch the last batch entry wins,
      and it ignores the ``normalized`` scaling.  The count buffer is zeroed
      (the reference does not zero it before first use).  The reference
      gradient op accepts no mipmapping (resample_grid.cc:135), so mips with
      this semantics raise.

    With mips (exact semantics) the input gradient is scattered into the
    levels read and collapsed through the adjoint of the mip averaging; the
    LUT gradient carries the per-level data gradients scaled by 2^-m and, in
    its 4th channel, the derivative with respect to the level of detail.
    """
    if mipmapping != MP.MIP_NONE:
        if semantics == "reference":
            raise ValueError("the reference gradient op does not support mipmapping")
        return _lut_grad_mips(inp, lut, out_grad, filter_mode, boundary, cell_center_offset, separate_camera_batch,
                              relative, normalized, mipmapping, num_mipmaps, mip_bias)
    inp = np.asarray(inp, dtype=np.float64)
    lut = np.asarray(lut, dtype=np.float64)
    out_grad = np.asarray(out_grad, dtype=np.float64)
    N, D, H, W, C = inp.shape
    V = lut.shape[0]
    in_grad = np.zeros_like(inp)
    lut_grad = np.zeros_like(lut)
    d, h, w = lut.shape[1:4]
    for n in range(N):
        count = np.zeros((D, H, W), dtype=np.int64) if semantics == "reference" else None
        for nn, v, vo in _pairs(N, V, separate_camera_batch):
            if nn != n:
                continue
            q = lut_positions(lut[v], relative, normalized) - cell_center_offset
            g = out_grad[n, vo]
            if semantics == "reference":
                S.scatter_linear(inp.shape[1:], q, g, boundary, mask="reference", count=count, out=in_grad[n])
                if filter_mode == LINEAR:
                    dx, dy, dz = S.data_gradient(inp[n], q, boundary)
                    lut_grad[v, ..., 0] = np.sum(g * dx, axis=-1)
                    lut_grad[v, ..., 1] = np.sum(g * dy, axis=-1)
                    lut_grad[v, ..., 2] = np.sum(g * dz, axis=-1)
                    lut_grad[v, ..., 3] = 0.0
            else:
                if filter_mode == LINEAR:
                    S.scatter_linear(inp.shape[1:], q, g, boundary, mask="forward", out=in_grad[n])
                    dx, dy, dz = S.data_gradient(inp[n], q, boundary)
                    sc = np.array([w, h, d], dtype=np.float64) if normalized else np.ones(3)
                    lut_grad[v, ..., 0] += np.sum(g * dx, axis=-1) * sc[0]
                    lut_grad[v, ..., 1] += np.sum(g * dy, axis=-1) * sc[1]
                    lut_grad[v, ..., 2] += np.sum(g * dz, axis=-1) * sc[2]
                elif filter_mode == NEAREST:
                    S.scatter_nearest(inp.shape[1:], q, g, boundary, out=in_grad[n])
                else:
                    S.scatter_minmax(inp[n], q, g, boundary, filter_mode == MAX, out=in_grad[n])
        if semantics == "reference":
            nz = count > 0
            in_grad[n][nz] = in_grad[n][nz] / count[nz][:, None]
    return in_grad, lut_grad


def _lut_grad_mips(inp, lut, out_grad, filter_mode, boundary, cco, separate_camera_batch, relative, normalized,
                   mipmapping, num_mipmaps, mip_bias):
    inp = np.asarray(inp, dtype=np.float64)
    lut = np.asarray(lut, dtype=np.float64)
    out_grad = np.asarray(out_grad, dtype=np.float64)
    N = inp.shape[0]
    V = lut.shape[0]
    in_grad = np.zeros_like(inp)
    lut_grad = np.zeros_like(lut)
    d, h, w = lut.shape[1:4]
    sc = np.array([w, h, d], dtype=np.float64) if normalized else np.ones(3)
    for n in range(N):
        levels = MP.generate_mips(inp[n], num_mipmaps)
        shapes = [l.shape for l in levels]
        acc = [np.zeros(sh) for sh in shapes]
        for nn, v, vo in _pairs(N, V, separate_camera_batch):
            if nn != n:
                continue
            pos = lut_positions(lut[v], relative, normalized)
            lod = lut[v][..., 3]
            g = out_grad[n, vo]
            gl = MP.scatter_levels(shapes, pos, g, lod, mipmapping, mip_bias, cco, filter_mode, boundary, levels)
            for a, b in zip(acc, gl):
                a += b
            gp, glod = MP.position_gradient(levels, pos, g, lod, mipmapping, mip_bias, cco, filter_mode, boundary)
            lut_grad[v, ..., :3] += gp * sc
            lut_grad[v, ..., 3] += glod
        in_grad[n] = MP.collapse_adjoint(acc)
    return in_grad, lut_grad


def lod_transform(in_dhw, MV, P, F, out_dhw, inverse_transform=False, linearize_depth=True, relative_coords=False):
    """LodTransform (resample_grid.cc:152-172, kComputeLoD3D
    resample_grid.cu.cc:1706-1720): per camera a [D', H', W', 4] lookup table
    of sample positions (samplePos3D<CM, false> at the output cell centres,
    minus the integer output index when ``relative_coords``) and the
    central-difference level of detail in the 4th channel.  The model matrix
    is the identity; MV is the model-view matrix (op layout)."""
    if linearize_depth:
        cm = TRANSFORM_LINDEPTH_REVERSE if inverse_transform else TRANSFORM_LINDEPTH
    else:
        cm = TRANSFORM_REVERSE if inverse_transform else TRANSFORM
    d, h, w = (int(v) for v in out_dhw)
    eye = np.eye(4).reshape(-1)
    out = np.zeros((len(MV), d, h, w, 4))
    for v in range(len(MV)):
        cx = G.CameraXform.from_op_args(eye, MV[v], P[v], F[v])
        fn = lambda idx: _positions(cx, idx, out_dhw, in_dhw, cm)
        pos = fn(_index_grid(d, h, w) + 0.5)
        if relative_coords:
            pos = pos - _index_grid(d, h, w)
        out[v, ..., :3] = pos
        out[v, ..., 3] = MP.central_lod(fn, out_dhw)
    return out


# --- transform sampling -----------------------------------------------------

TRANSFORM_LINDEPTH, TRANSFORM_LINDEPTH_REVERSE, TRANSFORM, TRANSFORM_REVERSE = 0, 1, 3, 4
COORDINATE_MODE = {"TRANSFORM_LINDEPTH": 0, "TRANSFORM_LINDEPTH_REVERSE": 1, "TRANSFORM": 3, "TRANSFORM_REVERSE": 4}


def _positions(cx, idx, out_dhw, in_dhw, coordinate_mode):
    d, h, w = out_dhw
    out_dims = np.array([w, h, d], dtype=np.float64)
    in_dims = np.array([in_dhw[2], in_dhw[1], in_dhw[0]], dtype=np.float64)
    if coordinate_mode == TRANSFORM:
        return G.idx_to_os(cx, idx, out_dims)
    if coordinate_mode == TRANSFORM_LINDEPTH:
        return G.idx_to_os_lindepth(cx, idx, out_dims)
    if coordinate_mode == TRANSFORM_REVERSE:
        return G.os_to_idx(cx, idx, in_dims)
    if coordinate_mode == TRANSFORM_LINDEPTH_REVERSE:
        return G.os_to_idx_lindepth(cx, idx, in_dims)
    raise ValueError("Invalid coordinate_mode")


def transform_positions(cx, out_dhw, in_dhw, coordinate_mode):
    """samplePos3D<CM,false> on every output cell centre (sampling.hpp:151-228)."""
    d, h, w = out_dhw
    return _positions(cx, _index_grid(d, h, w) + 0.5, out_dhw, in_dhw, coordinate_mode)


def transform_lod(cx, out_dhw, in_dhw, coordinate_mode):
    """Raw level of detail per output cell (samplePos3D<CM,true>)."""
    return MP.transform_lod(lambda idx: _positions(cx, idx, out_dhw, in_dhw, coordinate_mode), out_dhw,
                            coordinate_mode in (TRANSFORM, TRANSFORM_LINDEPTH))


def _transform_pairs(N, V, separate_camera_batch):
    return _pairs(N, V, separate_camera_batch)


def _sample2(grid, q, filter_mode, boundary):
    """Sampler2 (sampling.hpp:802-826): BORDER/CLAMP only, WRAP reads 0;
    modes other than LINEAR / QUADRATIC fall to NEAREST."""
    from . import WRAP, QUADRATIC
    if boundary == WRAP:
        return np.zeros(q.shape[:-1] + (grid.shape[-1],))
    if filter_mode == LINEAR:
        return S.sample_linear(grid, q, boundary)
    if filter_mode == QUADRATIC:
        return S.sample_quadratic(grid, q, boundary)
    return S.sample_nearest(grid, q, boundary)


def sample_grid_transform(inp, M, V, P, F, out_dhw, filter_mode=LINEAR, boundary=BORDER,
                          coordinate_mode=TRANSFORM_LINDEPTH, separate_camera_batch=True, mipmapping=MP.MIP_NONE,
                          num_mipmaps=0, mip_bias=0.0):
    """SampleGridTransform (resample_grid.cu.cc:255-544).  M/V/P are op arguments
    (flat 16 floats each, per batch / per camera), F is [V, 6].  With
    ``mipmapping`` NEAREST / LINEAR the input's ``num_mipmaps`` mips are read
    at the cell's level of detail plus ``mip_bias`` (resample_grid.cu.cc:347-497)."""
    inp = np.asarray(inp, dtype=np.float64)
    N = inp.shape[0]
    nV = len(V)
    Vo = nV if separate_camera_batch else 1
    out = np.zeros((N, Vo) + tuple(out_dhw) + (inp.shape[-1],))
    for n, v, vo in _transform_pairs(N, nV, separate_camera_batch):
        cx = G.CameraXform.from_op_args(M[n], V[v], P[v], F[v])
        pos = transform_positions(cx, out_dhw, inp.shape[1:4], coordinate_mode)
        if mipmapping == MP.MIP_NONE:
            out[n, vo] = _sample2(inp[n], pos - 0.5, filter_mode, boundary)
        else:
            lod = transform_lod(cx, out_dhw, inp.shape[1:4], coordinate_mode)
            levels = MP.generate_mips(inp[n], num_mipmaps)
            out[n, vo] = MP.sample_levels(levels, pos, lod, mipmapping, mip_bias, 0.5, filter_mode, boundary,
                                          sampler=lambda g, q, fm, bm: _sample2(g, q, fm, bm))
    return out


def sample_grid_transform_grad(inp, out_grad, M, V, P, F, filter_mode=LINEAR, boundary=BORDER,
                               coordinate_mode=TRANSFORM_LINDEPTH, separate_camera_batch=True, semantics="exact",
                               mipmapping=MP.MIP_NONE, num_mipmaps=0, mip_bias=0.0):
    """SampleGridTransformGrad.  semantics="reference": trilinear scatter with
    the reference BORDER mask, divided by the per-voxel corner-hit count summed
    over the cameras of each batch entry (resample_grid.cu.cc:1160-1387); the
    reference op accepts no mipmapping (resample_grid.cc:79), so mips with this
    semantics raise.  semantics="exact": true adjoint of
    ``sample_grid_transform``, mips included (scatter into the levels read,
    then the adjoint of the mip averaging)."""
    if semantics == "reference" and mipmapping != MP.MIP_NONE:
        raise ValueError("the reference gradient op does not support mipmapping")
    if mipmapping != MP.MIP_NONE:
        from . import WRAP
        inp = np.asarray(inp, dtype=np.float64)
        out_grad = np.asarray(out_grad, dtype=np.float64)
        N = inp.shape[0]
        in_grad = np.zeros_like(inp)
        out_dhw = out_grad.shape[2:5]
        if boundary == WRAP:
            return in_grad
        for n, v, vo in _transform_pairs(N, len(V), separate_camera_batch):
            cx = G.CameraXform.from_op_args(M[n], V[v], P[v], F[v])
            pos = transform_positions(cx, out_dhw, inp.shape[1:4], coordinate_mode)
            lod = transform_lod(cx, out_dhw, inp.shape[1:4], coordinate_mode)
            levels = MP.generate_mips(inp[n], num_mipmaps)
            fm = filter_mode if filter_mode in (LINEAR, QUADRATIC) else NEAREST
            gl = MP.scatter_levels([l.shape for l in levels], pos, out_grad[n, vo], lod, mipmapping, mip_bias, 0.5,
                                   fm, boundary)
            in_grad[n] += MP.collapse_adjoint(gl)
        return in_grad
    from . import WRAP
    inp = np.asarray(inp, dtype=np.float64)
    out_grad = np.asarray(out_grad, dtype=np.float64)
    N, D, H, W, C = inp.shape
    nV = len(V)
    out_dhw = out_grad.shape[2:5]
    in_grad = np.zeros_like(inp)
    for n in range(N):
        count = np.zeros((D, H, W), dtype=np.int64)
        for nn, v, vo in _transform_pairs(N, nV, separate_camera_batch):
            if nn != n:
                continue
            cx = G.CameraXform.from_op_args(M[n], V[v], P[v], F[v])
            q = transform_positions(cx, out_dhw, inp.shape[1:4], coordinate_mode) - 0.5
            g = out_grad[n, vo]
            if semantics == "reference":
                # kScatter3DGradients: the quadratic scatter for QUADRATIC, the
                # trilinear one for every other filter (resample_grid.cu.cc:1160-1185)
                if filter_mode == QUADRATIC:
                    S.scatter_quadratic(inp.shape[1:], q, g, boundary, mask="reference", count=count, out=in_grad[n])
                else:
                    S.scatter_linear(inp.shape[1:], q, g, boundary, mask="reference", count=count, out=in_grad[n])
            else:
                if boundary == WRAP:
                    continue
                if filter_mode == LINEAR:
                    S.scatter_linear(inp.shape[1:], q, g, boundary, mask="forward", out=in_grad[n])
                elif filter_mode == QUADRATIC:
                    S.scatter_quadratic(inp.shape[1:], q, g, boundary, mask="forward", out=in_grad[n])
                else:
                    S.scatter_nearest(inp.shape[1:], q, g, boundary, out=in_grad[n])
        if semantics == "reference":
            nz = count > 0
            in_grad[n][nz] = in_grad[n][nz] / count[nz][:, None]
    return in_grad


# --- aligned resampling (orthographic transform sampling) ------------------------

def aligned_scale(align, in_size, out_size):
    """Per-axis input scale of resample_grid3D_aligned (renderer.py:864-876):
    BORDER aligns the outer cell borders, CENTER the outer cell centres,
    STAGGER_INPUT input centres to output borders, STAGGER_OUTPUT input
    borders to output centres."""
    a = align.upper()
    if a == "BORDER":
        return 2.0 / in_size
    if a == "CENTER":
        return 2.0 / out_size * (out_size - 1.0) / (in_size - 1.0)
    if a == "STAGGER_INPUT":
        return 2.0 / (in_size - 1.0)
    if a == "STAGGER_OUTPUT":
        return 2.0 / (in_size + 1.0)
    raise ValueError("unknown alignment {}".format(align))


def aligned_args(in_dhw, out_dhw, align_xyz):
    """Op arguments (M, V, P, F) of resample_grid3D_aligned: the input grid,
    centred and scaled per axis, seen by an orthographic identity camera over
    [-1, 1]^3 whose frustum grid is the target (renderer.py:222-238,884-905)."""
    sc = [aligned_scale(align_xyz[0], in_dhw[2], out_dhw[2]), aligned_scale(align_xyz[1], in_dhw[1], out_dhw[1]),
          aligned_scale(align_xyz[2], in_dhw[0], out_dhw[0])]
    M = G.grid_transform(in_dhw, scale_=sc, center=True)
    V = np.linalg.inv(G.grid_transform(out_dhw))
    return G.op_matrix(M), G.op_matrix(V), G.op_matrix(np.eye(4)), np.array([-1.0, 1.0, -1.0, 1.0, 1.0, -1.0])


def resample_grid3D_aligned(data, out_dhw, align_xyz=("BORDER", "BORDER", "BORDER"), filter_mode=LINEAR,
                            boundary=BORDER):
    """[N, D, H, W, C] -> [N, D', H', W', C] (renderer.py:843-905)."""
    data = np.asarray(data, dtype=np.float64)
    M, V, P, F = aligned_args(data.shape[1:4], out_dhw, align_xyz)
    N = data.shape[0]
    return sample_grid_transform(data, [M] * N, [V], [P], [F], out_dhw, filter_mode, boundary, TRANSFORM)[:, 0]


def resample_grid3D_aligned_grad(data_shape, out_grad, align_xyz=("BORDER", "BORDER", "BORDER"), filter_mode=LINEAR,
                                 boundary=BORDER):
    """Exact adjoint of ``resample_grid3D_aligned`` w.r.t. its data."""
    out_grad = np.asarray(out_grad, dtype=np.float64)
    M, V, P, F = aligned_args(data_shape[1:4], out_grad.shape[1:4], align_xyz)
    N = data_shape[0]
    return sample_grid_transform_grad(np.zeros(data_shape), out_grad[:, None], [M] * N, [V], [P], [F], filter_mode,
                                      boundary, TRANSFORM)
