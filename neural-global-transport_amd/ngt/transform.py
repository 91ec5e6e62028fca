"""Grid / camera transforms (host side, numpy, FP64 math).

Mirror of phitest/render/transform.py: ``MatrixTransform``, ``Transform`` and
``GridTransform`` with the same constructor arguments, matrix composition
(``get_local_transform`` = T @ R @ S_norm @ S @ C, transform.py:370-384) and
parent chaining (transform.py:135-141).  Data attached to a GridTransform is a
torch tensor (NDHWC) instead of a tf.Tensor.
"""
import copy

import numpy as np
from scipy.spatial.transform import Rotation


class MatrixTransform(object):
    def __init__(self, transform_matrix=None, parent=None, grid_size=None, static=False):
        self._matrix = transform_matrix if transform_matrix is not None else self.identity_matrix()
        if parent is not None and not isinstance(parent, MatrixTransform):
            raise TypeError("parent must be a Transform object or None.")
        self.parent = parent
        self.grid_size = grid_size
        self.static = static

    @staticmethod
    def translation_matrix(translation):
        M = np.eye(4, dtype=np.float64)
        M[:3, 3] = np.asarray(translation, dtype=np.float64)
        return M

    @staticmethod
    def rotation_matrix(rotation):
        if isinstance(rotation, Rotation):
            rot = rotation.as_matrix()
        else:
            rotation = np.asarray(rotation, dtype=np.float64)
            if rotation.shape == (3, 3):
                rot = rotation
            elif rotation.shape == (3,):
                rot = Rotation.from_euler("xyz", rotation, degrees=True).as_matrix()
            else:
                raise ValueError("Can't use %s as rotation" % (rotation,))
        M = np.eye(4, dtype=np.float64)
        M[:3, :3] = rot
        return M

    @staticmethod
    def scale_matrix(scale):
        return np.diag([float(scale[0]), float(scale[1]), float(scale[2]), 1.0])

    @staticmethod
    def identity_matrix():
        return np.eye(4, dtype=np.float64)

    def set_parent(self, parent_transform):
        if parent_transform is not None and not isinstance(parent_transform, MatrixTransform):
            raise TypeError("parent must be a Transform object or None.")
        self.parent = parent_transform

    def copy_no_data(self):
        return copy.deepcopy(self)

    def get_local_transform(self):
        return self._matrix

    def get_transform_matrix(self):
        if self.parent is not None:
            return self.parent.get_transform_matrix() @ self.get_local_transform()
        return self.get_local_transform()

    def get_inverse_transform(self):
        return np.linalg.inv(self.get_transform_matrix())

    def inverse(self):
        return MatrixTransform(self.get_inverse_transform())

    def position_global(self):
        return self.get_transform_matrix() @ np.asarray([0, 0, 0, 1.0])

    def forward_global(self):
        v = self.get_transform_matrix() @ np.asarray([0, 0, 1, 0.0])
        return v / np.linalg.norm(v)

    def up_global(self):
        v = self.get_transform_matrix() @ np.asarray([0, 1, 0, 0.0])
        return v / np.linalg.norm(v)

    def right_global(self):
        v = self.get_transform_matrix() @ np.asarray([1, 0, 0, 0.0])
        return v / np.linalg.norm(v)

    def transform(self, vector):
        return self.get_transform_matrix() @ np.asarray(vector, dtype=np.float64)

    def __eq__(self, other):
        return np.array_equal(self.get_transform_matrix(), other.get_transform_matrix())


class Transform(MatrixTransform):
    def __init__(self, translation=(0, 0, 0), rotation_deg=(0, 0, 0), scale=(1, 1, 1), parent=None, static=False,
                 rotation_rotvec=None, rotation_quat=None):
        self.translation = list(translation)
        if rotation_quat is not None:
            self._rotation = Rotation.from_quat(rotation_quat)
        elif rotation_rotvec is not None:
            self._rotation = Rotation.from_rotvec(rotation_rotvec)
        else:
            self._rotation = Rotation.from_euler("xyz", rotation_deg, degrees=True)
        self.scale = list(scale)
        self.parent = parent
        self.static = static
        self.grid_size = None

    @property
    def rotation_deg(self):
        return self._rotation.as_euler("xyz", degrees=True)

    @rotation_deg.setter
    def rotation_deg(self, value):
        self._rotation = Rotation.from_euler("xyz", value, degrees=True)

    def set_rotation_angle(self, rotation_deg):
        self.rotation_deg = [0, 0, 0] if rotation_deg is None else rotation_deg

    def set_translation(self, translation):
        self.translation = [0, 0, 0] if translation is None else list(translation)

    def set_scale(self, scale):
        self.scale = [1, 1, 1] if scale is None else list(scale)

    def get_local_transform(self):
        return self.translation_matrix(self.translation) @ (self.rotation_matrix(self._rotation) @
                                                            self.scale_matrix(self.scale))


class GridTransform(Transform):
    """Transform of a [D, H, W] grid; ``center`` shifts the grid centre to the
    origin, ``normalize`` scales by 1/size ('ALL'), 1/min or 1/max."""

    def __init__(self, grid_size, translation=(0, 0, 0), rotation_deg=(0, 0, 0), scale=(1, 1, 1), center=False,
                 normalize="NONE", parent=None, static=False, rotation_rotvec=None, rotation_quat=None):
        super().__init__(translation, rotation_deg=rotation_deg, scale=scale, parent=parent,
                         rotation_rotvec=rotation_rotvec, rotation_quat=rotation_quat)
        self._data = None
        self._grid_size = [int(v) for v in grid_size]
        self.center = center
        self.normalize = normalize

    @property
    def grid_size(self):
        return list(self._grid_size)

    @grid_size.setter
    def grid_size(self, value):
        if getattr(self, "_data", None) is not None:
            raise ValueError("Can't set grid_size if GridTransform that has assoziated data set.")
        if value is not None:
            self._grid_size = [int(v) for v in value]

    @property
    def data(self):
        return self._data

    @data.setter
    def data(self, value):
        self.set_data(value)

    def set_data(self, data, format="NDHWC"):
        if data is None:
            self._data = None
            return
        self._data = data
        self._grid_size = [int(v) for v in data.shape[1:4]]

    @property
    def has_data(self):
        return self._data is not None

    def get_grid(self):
        return self._data

    def get_grid_size(self):
        return np.asarray(self._grid_size)

    def get_batch_size(self):
        if self._data is None:
            raise ValueError("data is not set")
        return int(self._data.shape[0])

    def get_channel(self):
        if self._data is None:
            raise ValueError("data is not set")
        return int(self._data.shape[-1])

    def copy_no_data(self):
        d = self._data
        self._data = None
        gt = copy.deepcopy(self)
        self._data = d
        return gt

    def copy_new_data(self, data):
        gt = self.copy_no_data()
        gt.set_data(data)
        return gt

    def copy_same_data(self):
        return self.copy_new_data(self._data)

    def get_local_transform(self):
        size = np.flip(self.get_grid_size()).astype(np.float64)  # shape is zyx, coordinates xyz
        M_center = self.translation_matrix(-size / 2.0)
        M = self.scale_matrix(self.scale) @ M_center if self.center else self.scale_matrix(self.scale)
        if self.normalize == "ALL":
            M = self.scale_matrix(1.0 / size) @ M
        elif self.normalize == "MIN":
            M = self.scale_matrix([1.0 / np.min(size)] * 3) @ M
        elif self.normalize == "MAX":
            M = self.scale_matrix([1.0 / np.max(size)] * 3) @ M
        return self.translation_matrix(self.translation) @ (self.rotation_matrix(self._rotation) @ M)

    def cell_size_world(self):
        M = self.get_transform_matrix()
        return np.linalg.norm(M[:3, :3], axis=0)

    def __eq__(self, other):
        return (np.array_equal(self.get_transform_matrix(), other.get_transform_matrix())
                and np.array_equal(self.get_grid_size(), other.get_grid_size()))
