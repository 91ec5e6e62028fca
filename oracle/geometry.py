"""Camera / grid-transform matrices and coordinate mappings in FP64.

Test infrastructure only (see ``oracle/__init__.py``).

Matrix conventions follow the reference exactly:
  * ``projection(near, far, trbl)`` is the OpenGL perspective matrix
    (camera.py:75-95), ``set_fov`` the horizontal-FOV frustum (camera.py:67-73);
  * a grid transform is ``T @ R @ S_norm @ S @ C`` (transform.py:370-384);
  * the ops receive every matrix *transposed* (renderer.py:117-121,
    ``get_transform_matrix().transpose()``) and read the 16 floats
    column-major (transformations.hpp:42-47, vectormath.hpp:307-314), so the
    kernel sees the mathematical matrix.  ``op_matrix(A)`` converts a math
    matrix to that flat op argument and ``math_matrix(flat)`` back.

Coordinate spaces (transformations.hpp:70-80): IDX = frustum-grid index
coordinates in [0, dims]; NDC in [-1, 1]; VS = view space; OS = object space
of the volume grid in cell units (cell i spans [i, i+1]).
"""
import numpy as np
from scipy.spatial.transform import Rotation


def op_matrix(A):
    """Flat 16-float argument the reference ops expect for math matrix A."""
    return np.asarray(A, dtype=np.float64).T.reshape(16).copy()


def math_matrix(flat):
    """Inverse of ``op_matrix``: the matrix the kernel reads column-major."""
    return np.asarray(flat, dtype=np.float64).reshape(4, 4).T.copy()


def translation(t):
    M = np.eye(4)
    M[:3, 3] = t
    return M


def scale(s):
    return np.diag([s[0], s[1], s[2], 1.0])


def rotation_xyz_deg(r):
    """``Rotation.from_euler('xyz', r, degrees=True)`` as 4x4 (transform.py:60-70)."""
    M = np.eye(4)
    M[:3, :3] = Rotation.from_euler("xyz", r, degrees=True).as_matrix()
    return M


def grid_transform(grid_size_dhw, translation_=(0, 0, 0), rotation_deg=(0, 0, 0), scale_=(1, 1, 1),
                   center=False, normalize="NONE", parent=None):
    """GridTransform.get_transform_matrix (transform.py:370-384, 138-141)."""
    size = np.asarray(grid_size_dhw, dtype=np.float64)[::-1]  # xyz
    M_center = translation(-size / 2.0)
    M = scale(scale_) @ M_center if center else scale(scale_)
    if normalize == "ALL":
        M = scale(1.0 / size) @ M
    elif normalize == "MIN":
        M = scale([1.0 / np.min(size)] * 3) @ M
    elif normalize == "MAX":
        M = scale([1.0 / np.max(size)] * 3) @ M
    M = translation(translation_) @ (rotation_xyz_deg(rotation_deg) @ M)
    if parent is not None:
        M = parent @ M
    return M


def transform(translation_=(0, 0, 0), rotation_deg=(0, 0, 0), scale_=(1, 1, 1), parent=None):
    """Transform.get_local_transform (transform.py:247-251) with parent."""
    M = translation(translation_) @ (rotation_xyz_deg(rotation_deg) @ scale(scale_))
    if parent is not None:
        M = parent @ M
    return M


def set_fov(fov_deg, aspect, near, vertical=False):
    """Camera.set_fov -> [top, right, bottom, left] (camera.py:67-73)."""
    half_h = np.tan(fov_deg / 2.0 * np.pi / 180.0) * near
    half_v = half_h / aspect
    if vertical:
        return [half_h, half_v, -half_h, -half_v]
    return [half_v, half_h, -half_v, -half_h]


def projection(near, far, trbl, perspective=True):
    """Camera.projection_matrix as a math matrix (camera.py:75-95)."""
    t, r, b, l = trbl
    n, f = near, far
    if not perspective:
        A = np.array([[2 / (r - l), 0, 0, 0],
                      [0, 2 / (t - b), 0, 0],
                      [0, 0, -2 / (f - n), 0],
                      [-(r + l) / (r - l), -(t + b) / (t - b), -(f + n) / (f - n), 1]], dtype=np.float64)
    else:
        A = np.array([[2 * n / (r - l), 0, 0, 0],
                      [0, 2 * n / (t - b), 0, 0],
                      [(r + l) / (r - l), (t + b) / (t - b), -(f + n) / (f - n), -1],
                      [0, 0, -2 * f * n / (f - n), 0]], dtype=np.float64)
    return A.T


def frustum_params(near, far, trbl):
    """Camera.frustum: [near, far, left, right, top, bottom] (camera.py:64-65)."""
    return np.array([near, far, trbl[3], trbl[1], trbl[0], trbl[2]], dtype=np.float64)


class CameraXform:
    """Per (model, camera) pair: everything the sampling kernels read."""

    def __init__(self, M, V, P, frustum):
        self.M = np.asarray(M, dtype=np.float64)
        self.V = np.asarray(V, dtype=np.float64)
        self.P = np.asarray(P, dtype=np.float64)
        self.MV = self.V @ self.M  # transformations.hpp:53
        self.MV_inv = np.linalg.inv(self.MV)
        self.P_inv = np.linalg.inv(self.P)
        self.near = float(frustum[0])
        self.far = float(frustum[1])

    @classmethod
    def from_op_args(cls, m16, v16, p16, frustum6):
        return cls(math_matrix(m16), math_matrix(v16), math_matrix(p16), frustum6)


def _apply(M, p4):
    return np.einsum("ij,...j->...i", M, p4)


def _h(p3, w=1.0):
    return np.concatenate([p3, np.full(p3.shape[:-1] + (1,), w)], axis=-1)


def idx_to_ndc(idx, out_dims_xyz):
    return 2.0 * idx / out_dims_xyz - 1.0


def ndc_to_os(cx, ndc):
    vs = _apply(cx.P_inv, _h(ndc))
    vs = vs / vs[..., 3:4]
    os_ = _apply(cx.MV_inv, vs)
    return os_[..., :3]


def idx_to_os(cx, idx, out_dims_xyz):
    """IDXtoOS (transformations.hpp:124-138)."""
    return ndc_to_os(cx, idx_to_ndc(idx, out_dims_xyz))


def idx_to_os_lindepth(cx, idx, out_dims_xyz):
    """IDXtoOSlinearDepth (transformations.hpp:139-149): depth slices uniform in
    view-space z between near and far."""
    ndc = idx_to_ndc(idx, out_dims_xyz)
    t = idx[..., 2] / out_dims_xyz[2]
    z = -(cx.near + t * (cx.far - cx.near))
    ndc = ndc.copy()
    ndc[..., 2] = -cx.P[2, 2] - cx.P[2, 3] / z
    return ndc_to_os(cx, ndc)


def os_to_ndc(cx, pos):
    vs = _apply(cx.MV, _h(pos))
    ndc = _apply(cx.P, vs)
    return ndc[..., :3] / ndc[..., 3:4]


def os_to_idx(cx, pos, in_dims_xyz):
    """OStoIDX (transformations.hpp:100-108)."""
    return (os_to_ndc(cx, pos) * 0.5 + 0.5) * in_dims_xyz


def os_to_idx_lindepth(cx, pos, in_dims_xyz):
    """OStoIDXlinearDepth (transformations.hpp:110-121)."""
    ndc = os_to_ndc(cx, pos)
    idx = (ndc * 0.5 + 0.5) * in_dims_xyz
    z = -cx.P[2, 3] / (ndc[..., 2] + cx.P[2, 2])
    t = -(z + cx.near) / (cx.far - cx.near)
    idx = idx.copy()
    idx[..., 2] = t * in_dims_xyz[2]
    return idx


def look_at_ring_cameras(num_views, radius, fov_deg, aspect, near, far, height=0.0, start_deg=0.0):
    """Pinhole cameras on a horizontal ring around the origin, all looking at
    it.  Returns a list of (V, P, frustum) math matrices.  The camera transform
    is Transform(rotation=(0, yaw, 0)) @ translation(0, height, radius) so its
    local -z axis points at the origin (camera.py:58-62: V = inverse)."""
    cams = []
    trbl = set_fov(fov_deg, aspect, near)
    P = projection(near, far, trbl)
    F = frustum_params(near, far, trbl)
    for i in range(num_views):
        yaw = start_deg + 360.0 * i / num_views
        C = rotation_xyz_deg([0.0, yaw, 0.0]) @ translation([0.0, height, radius])
        V = np.linalg.inv(C)
        cams.append((V, P, F))
    return cams
