"""Pinhole / orthographic cameras (host side).

Mirror of phitest/render/camera.py:24-140: ``Camera(grid_transform, nearFar,
topRightBottomLeft, fov, aspect, perspective, static, jitter)`` with the same
OpenGL projection matrix (camera.py:75-95), horizontal-FOV frustum
(``set_fov``, camera.py:67-73), frustum parameter order
[near, far, left, right, top, bottom] (camera.py:64-65) and view matrix =
inverse of the camera transform (camera.py:58-62).  ``transform.grid_size``
is the frustum grid [D, H, W] (depth slices, image height, image width).
"""
import numpy as np

from .transform import GridTransform, MatrixTransform


class Camera(object):
    def __init__(self, grid_transform, nearFar=(1, 10), topRightBottomLeft=(1, 1, -1, -1), fov=None, aspect=1,
                 perspective=True, static=None, jitter=None):
        assert isinstance(grid_transform, MatrixTransform)
        self.transform = grid_transform
        self.jitter = jitter
        self.trbl = list(topRightBottomLeft)
        self.clip = list(nearFar)
        self.perspective = perspective
        if fov is not None:
            self.set_fov(fov, aspect)
        self.static = static
        self.scissor_pad = None
        # lookup tables of a static camera (grid -> frustum and back), built on demand by the Renderer
        self.LuT = None
        self.inverseLuT = None

    def view_transform_inverse(self, with_jitter=True):
        if self.jitter is not None and with_jitter:
            jitter_position = np.random.uniform(-self.jitter, self.jitter, [3]).astype(np.float32)
            return GridTransform(self.transform.grid_size, translation=jitter_position, parent=self.transform)
        return self.transform

    def view_matrix_inverse(self, with_jitter=True):
        return self.view_transform_inverse(with_jitter).get_transform_matrix()

    def view_matrix(self, with_jitter=True):
        return self.view_transform_inverse(with_jitter).get_inverse_transform()

    def frustum(self):
        return np.asarray([self.clip[0], self.clip[1], self.trbl[3], self.trbl[1], self.trbl[0], self.trbl[2]],
                          dtype=np.float32)

    def set_fov(self, fov, aspect=1, vertical=False):
        half_h = np.tan(fov / 2.0 * np.pi / 180.0) * self.clip[0]
        half_v = half_h / aspect
        if vertical:
            self.trbl = [half_h, half_v, -half_h, -half_v]
        else:
            self.trbl = [half_v, half_h, -half_v, -half_h]

    def projection_matrix(self):
        """OpenGL projection matrix, returned in the reference's layout (the
        transpose of the mathematical matrix, camera.py:81-95)."""
        t, r, b, l = self.trbl
        n, f = self.clip
        if not self.perspective:
            A = np.asarray([[2 / (r - l), 0, 0, 0],
                            [0, 2 / (t - b), 0, 0],
                            [0, 0, -2 / (f - n), 0],
                            [-(r + l) / (r - l), -(t + b) / (t - b), -(f + n) / (f - n), 1]], dtype=np.float32)
        else:
            A = np.asarray([[2 * n / (r - l), 0, 0, 0],
                            [0, 2 * n / (t - b), 0, 0],
                            [(r + l) / (r - l), (t + b) / (t - b), -(f + n) / (f - n), -1],
                            [0, 0, -2 * f * n / (f - n), 0]], dtype=np.float32)
        return A.transpose()

    @property
    def depth_step(self):
        """step size in view space (camera.py:97-100)"""
        return (self.far - self.near) / self.transform.grid_size[0]

    @property
    def near(self):
        return self.clip[0]

    @property
    def far(self):
        return self.clip[1]

    @property
    def top(self):
        return self.trbl[0]

    @property
    def right(self):
        return self.trbl[1]

    @property
    def bottom(self):
        return self.trbl[2]

    @property
    def left(self):
        return self.trbl[3]

    @property
    def position_global(self):
        return self.transform.position_global()

    @property
    def forward_global(self):
        return self.transform.forward_global()


def ring_cameras(num_views, radius, fov_deg, image_hw, depth, near, far, height=0.0, start_deg=0.0):
    """Cameras on a horizontal ring around the origin looking at it: yaw
    rotation about +y, camera at local (0, height, radius), frustum grid
    [depth, H, W]."""
    H, W = image_hw
    cams = []
    for i in range(num_views):
        yaw = start_deg + 360.0 * i / num_views
        parent = GridTransform([depth, H, W], rotation_deg=[0.0, yaw, 0.0])
        t = GridTransform([depth, H, W], translation=[0.0, height, radius], parent=parent)
        cams.append(Camera(t, nearFar=[near, far], fov=fov_deg, aspect=W / H))
    return cams
