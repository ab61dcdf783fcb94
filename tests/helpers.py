"""Shared scene builders and error metrics for the tests."""
import numpy as np

from gft import synthetic as syn
from gft.camera import Camera, camera_params_batch, model_matrices
from gft.transform import GridTransform, Transform


def scene(grid=(12, 20, 14), image=(18, 11), views=3, seed=0, blobs=5, channels=1, yaw_range=(-60.0, 60.0)):
    """Small blob density seen from an arc of cameras; returns a dict with the
    grid (NDHWC float32), matrices and the per-camera depth sample count."""
    d = syn.gaussian_blobs(grid, blobs, seed=seed)
    if channels > 1:
        rng = np.random.Generator(np.random.PCG64(seed + 77))
        extra = rng.uniform(0.2, 1.0, size=(1, 1, 1, 1, channels - 1)).astype(np.float32)
        d = np.concatenate([d * extra, d], axis=-1)
    cams, S = syn.arc_cameras(grid, image, views, yaw_range=yaw_range)
    M = model_matrices([syn.grid_transform(grid)])
    V, P, F = camera_params_batch(cams)
    return dict(data=d.astype(np.float32), M=M, V=V, P=P, F=F, S=S, out_dhw=(S,) + tuple(image), cams=cams,
                step=cams[0].depth_step)


def norm_err(a, b):
    """max |a-b| / max |b| (normalised max error); 0 when both vanish."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max()
    num = np.abs(a - b).max()
    if den == 0:
        return float(num)
    return float(num / den)


def ortho_camera(grid, image_hw, depth, yaw=0.0):
    """Orthographic camera looking down -z at the grid centre, covering the
    grid exactly (pixel size = one cell when image matches the grid)."""
    D, H, W = grid
    t = GridTransform([depth, image_hw[0], image_hw[1]], translation=[0, 0, D], parent=Transform(rotation_deg=[0, yaw, 0]))
    cam = Camera(t, nearFar=[D / 2.0, 1.5 * D], topRightBottomLeft=[H / 2.0, W / 2.0, -H / 2.0, -W / 2.0],
                 perspective=False)
    return cam
