"""Camera model with the reference's OpenGL conventions.

Restates phitest/render/camera.py:24-136: the camera looks along -z of its
transform, the view matrix is the inverse of the camera transform
(camera.py:56-59), the projection is the OpenGL perspective/orthographic
matrix (camera.py:71-92), ``frustum()`` packs [near, far, left, right, top,
bottom] (camera.py:61-62) and ``set_fov`` takes a horizontal fov with
aspect = width/height (camera.py:64-70).
"""
import numpy as np

from .transform import GridTransform


class Camera(object):
    def __init__(self, grid_transform, nearFar=(1, 10), topRightBottomLeft=(1, 1, -1, -1), fov=None, aspect=1,
                 perspective=True, static=None, jitter=None):
        self.transform = grid_transform
        self.jitter = jitter
        self.trbl = list(topRightBottomLeft)
        self.clip = list(nearFar)
        self.perspective = perspective
        if fov is not None:
            self.set_fov(fov, aspect)
        self.static = static
        self.scissor_pad = None

    def view_transform_inverse(self, with_jitter=True, rng=None):
        if self.jitter is not None and with_jitter:
            rng = rng if rng is not None else np.random
            jitter_position = rng.uniform(-self.jitter, self.jitter, [3]).astype(np.float32)
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
        half_h = np.tan(fov / 2. * np.pi / 180.) * self.clip[0]
        half_v = half_h / aspect
        if vertical:
            self.trbl = [half_h, half_v, -half_h, -half_v]
        else:
            self.trbl = [half_v, half_h, -half_v, -half_h]

    def projection_matrix(self):
        """Row-major OpenGL projection (the reference builds it transposed and
        transposes it back, camera.py:78-92)."""
        t, r, b, l = self.trbl
        n, f = self.clip
        if not self.perspective:
            return np.asarray([[2 / (r - l), 0, 0, -(r + l) / (r - l)],
                               [0, 2 / (t - b), 0, -(t + b) / (t - b)],
                               [0, 0, -2 / (f - n), -(f + n) / (f - n)],
                               [0, 0, 0, 1]], dtype=np.float32)
        return np.asarray([[2 * n / (r - l), 0, (r + l) / (r - l), 0],
                           [0, 2 * n / (t - b), (t + b) / (t - b), 0],
                           [0, 0, -(f + n) / (f - n), -2 * f * n / (f - n)],
                           [0, 0, -1, 0]], dtype=np.float32)

    @property
    def depth_step(self):
        """View-space step per depth cell (camera.py:94-97)."""
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


def camera_params_batch(cameras, with_jitter=False):
    """Column-major V, P and frustum arrays for a camera list, as packed by
    renderer.py:108-112 (``view_matrix().transpose()`` etc.)."""
    V = np.stack([np.asarray(c.view_matrix(with_jitter=with_jitter), dtype=np.float32).T for c in cameras])
    P = np.stack([np.asarray(c.projection_matrix(), dtype=np.float32).T for c in cameras])
    F = np.stack([c.frustum() for c in cameras])
    return np.ascontiguousarray(V), np.ascontiguousarray(P), np.ascontiguousarray(F)


def model_matrices(transforms):
    """Column-major model matrices ``get_transform_matrix().transpose()``."""
    return np.ascontiguousarray(
        np.stack([np.asarray(t.get_transform_matrix(), dtype=np.float32).T for t in transforms]))
