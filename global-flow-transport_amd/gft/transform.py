"""Grid and camera transforms with the reference's matrix conventions.

Restates phitest/render/transform.py:8-302: 4x4 matrices in row-major numpy
form (column vectors, ``M @ p``), composed as ``T @ (R @ S)``; a GridTransform
adds the centring translation and size normalisation of
transform.py:253-267. The ops receive these matrices transposed (column-major
storage), exactly like the reference's ``get_transform_matrix().transpose()``
call sites (renderer.py:108-112,321-333).
"""
import copy

import numpy as np
from scipy.spatial.transform import Rotation


class MatrixTransform(object):
    """transform.py:8-145: a 4x4 matrix with an optional parent transform."""

    def __init__(self, transform_matrix=None, parent=None, grid_size=None, static=False):
        self._matrix = transform_matrix if transform_matrix is not None else self.identity_matrix()
        if parent is not None and not isinstance(parent, MatrixTransform):
            raise TypeError("parent must be a Transform object or None.")
        self.parent = parent
        self.grid_size = grid_size

    @classmethod
    def from_fwd_up_right_pos(cls, fwd, up, right, pos, parent=None):
        mat = np.asarray([[right[0], up[0], fwd[0], pos[0]],
                          [right[1], up[1], fwd[1], pos[1]],
                          [right[2], up[2], fwd[2], pos[2]],
                          [0, 0, 0, 1]], dtype=np.float32)
        return cls(mat, parent)

    @staticmethod
    def translation_matrix(translation):
        m = np.eye(4, dtype=np.float32)
        m[:3, 3] = np.asarray(translation, dtype=np.float32)[:3]
        return m

    @staticmethod
    def rotation_matrix(rotation):
        # scipy 'xyz' extrinsic Euler angles in degrees (transform.py:50-55)
        rot = Rotation.from_euler('xyz', rotation, degrees=True).as_matrix()
        m = np.eye(4, dtype=np.float64)
        m[:3, :3] = rot
        return m

    @staticmethod
    def scale_matrix(scale):
        m = np.eye(4, dtype=np.float32)
        m[0, 0], m[1, 1], m[2, 2] = scale[0], scale[1], scale[2]
        return m

    @staticmethod
    def identity_matrix():
        return np.eye(4, dtype=np.float32)

    def get_local_transform(self):
        return self._matrix

    def get_transform_matrix(self):
        if self.parent is not None:
            return self.parent.get_transform_matrix() @ self.get_local_transform()
        return self.get_local_transform()

    def get_inverse_transform(self):
        return np.linalg.inv(self.get_transform_matrix())

    def transform(self, vector):
        return self.get_transform_matrix() @ np.asarray(vector)

    def position_global(self):
        return self.get_transform_matrix() @ np.asarray([0, 0, 0, 1])

    def copy_no_data(self):
        return copy.deepcopy(self)


class Transform(MatrixTransform):
    """transform.py:147-192: translation, Euler rotation (deg) and scale."""

    def __init__(self, translation=(0, 0, 0), rotation_deg=(0, 0, 0), scale=(1, 1, 1), parent=None, static=False):
        self.translation = list(translation)
        self.rotation_deg = list(rotation_deg)
        self.scale = list(scale)
        self.parent = parent
        self.grid_size = None

    def get_local_transform(self):
        M_scale = Transform.scale_matrix(self.scale)
        M_rot = Transform.rotation_matrix(self.rotation_deg)
        M_trans = Transform.translation_matrix(self.translation)
        return M_trans @ (M_rot @ M_scale)


class GridTransform(Transform):
    """transform.py:195-302: grid_size is (D,H,W); ``center`` moves the grid
    centre to the origin, ``normalize`` scales by ALL/MIN/MAX of the size."""

    def __init__(self, grid_size, translation=(0, 0, 0), rotation_deg=(0, 0, 0), scale=(1, 1, 1), center=False,
                 normalize='NONE', parent=None, static=False):
        super().__init__(translation, rotation_deg, scale, parent)
        self.grid_size = list(grid_size)
        self.center = center
        self.normalize = normalize
        self.data = None

    def get_grid_size(self):
        return np.asarray(self.grid_size)

    def get_local_transform(self):
        size = np.flip(self.get_grid_size()).astype(np.float64)  # zyx -> xyz
        M_center = Transform.translation_matrix(-size / 2.0)
        M_norm_scale = None
        if self.normalize == 'ALL':
            M_norm_scale = Transform.scale_matrix(1.0 / size)
        if self.normalize == 'MIN':
            M_norm_scale = Transform.scale_matrix([1.0 / np.min(size)] * 3)
        if self.normalize == 'MAX':
            M_norm_scale = Transform.scale_matrix([1.0 / np.max(size)] * 3)
        M_scale = Transform.scale_matrix(self.scale)
        M_rot = Transform.rotation_matrix(self.rotation_deg)
        M_trans = Transform.translation_matrix(self.translation)
        M = M_scale @ M_center if self.center else M_scale
        M = M_norm_scale @ M if self.normalize != 'NONE' else M
        return M_trans @ (M_rot @ M)

    @classmethod
    def from_transform(cls, transform, grid_size, center=False, normalize='NONE'):
        """transform.py:211-212"""
        return cls(grid_size, transform.translation, transform.rotation_deg, transform.scale, center, normalize,
                   transform.parent)

    def grid_size_world(self):
        """World-space edge lengths of the grid along its x, y, z axes
        (transform.py:273-278)."""
        D, H, W = [float(v) for v in self.grid_size]
        M = np.asarray(self.get_transform_matrix(), dtype=np.float64)
        return np.array([np.linalg.norm(M[:3, 0] * W), np.linalg.norm(M[:3, 1] * H), np.linalg.norm(M[:3, 2] * D)])

    def cell_size_world(self):
        """World-space size of one cell along x, y, z (transform.py:280-284)."""
        M = np.asarray(self.get_transform_matrix(), dtype=np.float64)
        return np.array([np.linalg.norm(M[:3, 0]), np.linalg.norm(M[:3, 1]), np.linalg.norm(M[:3, 2])])

    def set_data(self, data):
        shape = list(data.shape)
        self.grid_size = shape[1:4]
        self.data = data
