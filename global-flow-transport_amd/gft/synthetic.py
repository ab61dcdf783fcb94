"""Seeded synthetic workloads for the transport/render path.

Density is a sum of Gaussian blobs, velocity the curl of three octaves of
value noise, and the cameras sit on a yaw arc around the grid centre with one
depth sample per cell. All generators use numpy PCG64 so that every consumer
(tests, the oracle, the benchmark) sees identical inputs for a seed.

Grid sizes are (D, H, W) = (z, y, x) and tensors are NDHWC float32, as in the
reference (phitest/render/vector.py:256-279 GridShape).
"""
import math

import numpy as np

from .camera import Camera
from .transform import GridTransform, Transform


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def gaussian_blobs(grid_dhw, num_blobs, seed=0, sigma_range=(3.0, 8.0), peak_range=(0.2, 1.0)):
    """[1,D,H,W,1] float32 sum of ``num_blobs`` Gaussian blobs; centres are
    uniform in the interior at least 2 sigma from every face."""
    D, H, W = grid_dhw
    rng = _rng(seed)
    z = np.arange(D, dtype=np.float64)[:, None, None]
    y = np.arange(H, dtype=np.float64)[None, :, None]
    x = np.arange(W, dtype=np.float64)[None, None, :]
    out = np.zeros((D, H, W), dtype=np.float64)
    for _ in range(num_blobs):
        s = rng.uniform(*sigma_range)
        peak = rng.uniform(*peak_range)
        lo = np.minimum(2 * s, np.array([W, H, D]) / 2 - 0.5)
        cx = rng.uniform(lo[0], W - 1 - lo[0])
        cy = rng.uniform(lo[1], H - 1 - lo[1])
        cz = rng.uniform(lo[2], D - 1 - lo[2])
        g = np.exp(-((x - cx) ** 2) / (2 * s * s)) * np.exp(-((y - cy) ** 2) / (2 * s * s)) * \
            np.exp(-((z - cz) ** 2) / (2 * s * s))
        out += peak * g
    out = np.maximum(out, 0.0)
    return out.astype(np.float32)[None, ..., None]


def _value_noise(shape_dhw, period, rng):
    """Smoothstep-interpolated random lattice values with lattice spacing
    ``period`` cells."""
    D, H, W = shape_dhw
    n = [int(math.ceil(s / period)) + 2 for s in (D, H, W)]
    lat = rng.uniform(-1.0, 1.0, size=n)

    def axis(s):
        t = np.arange(s, dtype=np.float64) / period
        i = np.floor(t).astype(np.int64)
        f = t - i
        f = f * f * (3 - 2 * f)
        return i, f

    iz, fz = axis(D)
    iy, fy = axis(H)
    ix, fx = axis(W)
    fz = fz[:, None, None]
    fy = fy[None, :, None]
    fx = fx[None, None, :]

    def L(dz, dy, dx):
        return lat[(iz + dz)[:, None, None], (iy + dy)[None, :, None], (ix + dx)[None, None, :]]

    c00 = L(0, 0, 0) * (1 - fx) + L(0, 0, 1) * fx
    c01 = L(0, 1, 0) * (1 - fx) + L(0, 1, 1) * fx
    c10 = L(1, 0, 0) * (1 - fx) + L(1, 0, 1) * fx
    c11 = L(1, 1, 0) * (1 - fx) + L(1, 1, 1) * fx
    c0 = c00 * (1 - fy) + c01 * fy
    c1 = c10 * (1 - fy) + c11 * fy
    return c0 * (1 - fz) + c1 * fz


def curl_noise_velocity(grid_dhw, seed=1, max_speed=1.5, periods=(32, 16, 8)):
    """[1,D,H,W,3] float32 centred velocity (x,y,z in cells/step): the
    central-difference curl of a vector potential made of three value-noise
    fields with octaves at ``periods`` and halving amplitude, scaled so that
    max |v| == max_speed."""
    rng = _rng(seed)
    pot = []
    for _ in range(3):
        f = np.zeros(grid_dhw, dtype=np.float64)
        amp = 1.0
        for p in periods:
            f += amp * _value_noise(grid_dhw, p, rng)
            amp *= 0.5
        pot.append(f)
    Ax, Ay, Az = pot

    def d(a, axis):
        return np.gradient(a, axis=axis)

    # axes: 0=z, 1=y, 2=x
    vx = d(Az, 1) - d(Ay, 0)
    vy = d(Ax, 0) - d(Az, 2)
    vz = d(Ay, 2) - d(Ax, 1)
    v = np.stack([vx, vy, vz], axis=-1)
    m = np.sqrt((v * v).sum(-1)).max()
    if m > 0:
        v *= max_speed / m
    return v.astype(np.float32)[None]


def centered_to_staggered(vel):
    """Staggered MAC components (x:[1,D,H,W+1,1], y:[1,D,H+1,W,1],
    z:[1,D+1,H,W,1]) whose face averages reproduce ``vel`` in the interior;
    boundary faces copy the adjacent centred value."""
    v = vel[0]
    vx = np.concatenate([v[:, :, :1, 0], 0.5 * (v[:, :, 1:, 0] + v[:, :, :-1, 0]), v[:, :, -1:, 0]], axis=2)
    vy = np.concatenate([v[:, :1, :, 1], 0.5 * (v[:, 1:, :, 1] + v[:, :-1, :, 1]), v[:, -1:, :, 1]], axis=1)
    vz = np.concatenate([v[:1, :, :, 2], 0.5 * (v[1:, :, :, 2] + v[:-1, :, :, 2]), v[-1:, :, :, 2]], axis=0)
    return (vx[None, ..., None].astype(np.float32), vy[None, ..., None].astype(np.float32),
            vz[None, ..., None].astype(np.float32))


def grid_transform(grid_dhw):
    """Model transform with world units = cells and the grid centred at the
    origin (GridTransform(center=True), transform.py:253-267)."""
    return GridTransform(list(grid_dhw), center=True)


def arc_cameras(grid_dhw, image_hw, num_views, yaw_range=(-60.0, 60.0), fov=40.0, radius_scale=1.5,
                extra_yaws=()):
    """Cameras on a yaw arc in the xz-plane around the grid centre at radius
    ``radius_scale * H`` cells, horizontal fov ``fov``, aspect W/H of the image,
    near/far = radius -+ sqrt(W^2+D^2)/2 and S = ceil(far-near) depth samples
    (one per cell). Returns (cameras, S)."""
    D, H, W = grid_dhw
    ih, iw = image_hw
    radius = radius_scale * H
    half = 0.5 * math.sqrt(W * W + D * D)
    near, far = radius - half, radius + half
    S = int(math.ceil(far - near - 1e-9))
    yaws = [0.0] if num_views == 1 else list(np.linspace(yaw_range[0], yaw_range[1], num_views))
    yaws += list(extra_yaws)
    cams = []
    for yaw in yaws:
        t = GridTransform([S, ih, iw], translation=[0, 0, radius], parent=Transform(rotation_deg=[0, float(yaw), 0]))
        cams.append(Camera(t, nearFar=[near, far], fov=fov, aspect=iw / ih))
    return cams, S


# Named workloads (grid D,H,W; views; image H,W; blobs; max speed)
WORKLOADS = {
    "c1": dict(grid=(64, 112, 64), views=5, image=(228, 128), blobs=24, max_speed=1.5),
    "c2": dict(grid=(128, 224, 128), views=5, image=(456, 256), blobs=48, max_speed=1.5),
    "c3": dict(grid=(100, 178, 100), views=5, image=(480, 270), blobs=48, max_speed=2.0, unseen_yaw=90.0),
    "c5": dict(grid=(200, 356, 200), views=8, image=(960, 540), blobs=48, max_speed=2.0),
}

# Batched uneven workload c4: grid sizes (D,H,W) and the image (H,W) used with each
C4_GRIDS = [(32, 32, 32), (48, 86, 48), (64, 112, 64), (96, 170, 96), (128, 224, 128)]
C4_IMAGES = [(214, 120), (274, 154), (335, 188), (395, 222), (456, 256)]


def batched_uneven(count=32, seed=7):
    """``count`` independent reconstructions for the batched-uneven workload:
    each draws a grid size from C4_GRIDS (with its image size from C4_IMAGES,
    120x214 to 256x456 px) and 1-8 views, seeded. Returns a list of dicts
    (grid, image, views, blobs, max_speed, seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        i = int(rng.integers(0, len(C4_GRIDS)))
        out.append(dict(grid=C4_GRIDS[i], image=C4_IMAGES[i], views=int(rng.integers(1, 9)),
                        blobs=24 if i < 3 else 48, max_speed=1.5, seed=1000 + 10 * k))
    return out
