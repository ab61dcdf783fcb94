"""Batched density-reconstruction step over many independent reconstructions
of uneven grid sizes and view counts (gft_batch_render_*).

Per reconstruction: the reference's non-fused render of its density under an
ambient light (Renderer.render_density with fused=False,
phitest/render/renderer.py:555-568,603-651: SampleGridTransform
TRANSFORM_LINDEPTH with LINEAR mipmapping and a per-sample LoD, then the
BEER_LAMBERT depth blend), the mean squared rgb error against its targets and
the exact density gradient. All reconstructions run together: per mip level
a pass building the level and one taking its corner-quad copy, one forward
launch over every pixel, one gather launch per level (the gradient summed per
voxel block, no atomics in global memory: images, losses and gradients repeat
bit for bit), one for the mip adjoint and the losses. ``run()`` only launches
kernels, so a ``torch.cuda.graph`` can capture it.

``autotune()`` picks the backward's brick (the voxel block of its gather)
per grid-size class by timing the whole batch: for each class in turn, every
candidate with the other classes at their current choice.
"""
import ctypes

import numpy as np
import torch

from . import _lib

BRICKS = {0: "16x16x2-voxel gather blocks", 1: "8x8x8-voxel gather blocks", 2: "16x8x4-voxel gather blocks"}


class BatchRecon(ctypes.Structure):
    """gft_batch_recon (include/gft.h)."""
    _fields_ = [("grid_dhw", ctypes.c_int32 * 3), ("density", ctypes.c_void_p), ("targets", ctypes.c_void_p),
                ("images", ctypes.c_void_p), ("density_grad", ctypes.c_void_p), ("loss", ctypes.c_void_p),
                ("M", ctypes.c_void_p), ("V", ctypes.c_void_p), ("P", ctypes.c_void_p), ("frustum", ctypes.c_void_p),
                ("num_views", ctypes.c_int32), ("output_shape", ctypes.c_int32 * 3), ("depth_step", ctypes.c_float),
                ("light", ctypes.c_float * 3), ("loss_scale", ctypes.c_float), ("brick", ctypes.c_int32)]


def _f32(a, n):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1, n))
    return a


class BatchedRenderStep(object):
    """``recons``: list of dicts with grid (D, H, W), M (one column-major 4x4),
    V, P (views x 4x4), F (views x 6), out_dhw (S, H, W), depth_step, and
    optionally light (rgb), loss_scale (default: the mean over the
    reconstruction's rgb values) and brick (0, 1 or 2). ``bind(densities,
    targets)`` takes the [1,D,H,W,1] densities and [V,H,W,3] targets (device
    float32; they stay bound), after which ``run()`` fills ``images``,
    ``grads`` and ``losses``."""

    def __init__(self, recons, num_mipmaps=3, device="cuda"):
        self.n = len(recons)
        if self.n < 1:
            raise _lib.InvalidArgumentError("the batch needs at least one reconstruction")
        self.num_mipmaps = int(num_mipmaps)
        self.device = torch.device(device)
        self._host = []
        self.images, self.grads, self.losses = [], [], []
        self._arr = (BatchRecon * self.n)()
        for k, r in enumerate(recons):
            D, H, W = (int(x) for x in r["grid"])
            S, h, w = (int(x) for x in r["out_dhw"])
            M, V, P, F = _f32(r["M"], 16)[:1], _f32(r["V"], 16), _f32(r["P"], 16), _f32(r["F"], 6)
            nv = V.shape[0]
            self._host.append((M, V, P, F))
            c = self._arr[k]
            c.grid_dhw[:] = [D, H, W]
            c.output_shape[:] = [S, h, w]
            c.num_views = nv
            c.depth_step = float(r["depth_step"])
            c.light[:] = [float(x) for x in r.get("light", (1.0, 1.0, 1.0))]
            c.loss_scale = float(r.get("loss_scale", 1.0 / (nv * h * w * 3)))
            c.brick = int(r.get("brick", 0))
            c.M, c.V, c.P, c.frustum = (x.ctypes.data for x in (M, V, P, F))
            img = torch.empty((nv, h, w, 4), device=self.device)
            g = torch.empty((1, D, H, W, 1), device=self.device)
            lo = torch.zeros((1,), device=self.device)
            self.images.append(img)
            self.grads.append(g)
            self.losses.append(lo)
            c.images, c.density_grad, c.loss = img.data_ptr(), g.data_ptr(), lo.data_ptr()
        self._arr_p = ctypes.cast(self._arr, ctypes.c_void_p)
        self._bound = None
        self.ws = None

    @property
    def bricks(self):
        return [int(self._arr[k].brick) for k in range(self.n)]

    def set_bricks(self, bricks):
        for k, b in enumerate(bricks):
            self._arr[k].brick = int(b)
        if self._bound is not None:
            self._prepare()

    def bind(self, densities, targets):
        if len(densities) != self.n or len(targets) != self.n:
            raise _lib.InvalidArgumentError("one density and one target per reconstruction")
        ds, ts = [], []
        for k, (d, t) in enumerate(zip(densities, targets)):
            c = self._arr[k]
            D, H, W = c.grid_dhw
            S, h, w = c.output_shape
            if d.dtype != torch.float32 or not d.is_cuda or tuple(d.shape) != (1, D, H, W, 1):
                raise _lib.InvalidArgumentError("reconstruction %d: density must be a [1,%d,%d,%d,1] float32 CUDA "
                                                "tensor" % (k, D, H, W))
            if t.dtype != torch.float32 or not t.is_cuda or tuple(t.shape) != (c.num_views, h, w, 3):
                raise _lib.InvalidArgumentError("reconstruction %d: targets must be [%d,%d,%d,3] float32 CUDA" %
                                                (k, c.num_views, h, w))
            d, t = d.contiguous(), t.contiguous()
            c.density, c.targets = d.data_ptr(), t.data_ptr()
            ds.append(d)
            ts.append(t)
        self._bound = (ds, ts)
        self._prepare()

    def _prepare(self):
        L = _lib.lib()
        need = int(L.gft_batch_render_workspace(self.n, self._arr_p, self.num_mipmaps))
        if need == 0:  # an invalid record: the set-up call reports which
            _lib.check(L.gft_batch_render_prepare(None, self.n, self._arr_p, self.num_mipmaps, None, 0))
        if self.ws is None or self.ws.numel() < need:
            self.ws = torch.empty((need,), dtype=torch.uint8, device=self.device)
        self._nws = need
        s = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _lib.check(L.gft_batch_render_prepare(s, self.n, self._arr_p, self.num_mipmaps, _lib.ptr(self.ws), self._nws))

    def run(self):
        if self._bound is None:
            raise _lib.InvalidArgumentError("bind() the densities and targets first")
        s = ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        _lib.check(_lib.lib().gft_batch_render_run(s, self.n, self._arr_p, self.num_mipmaps, _lib.ptr(self.ws),
                                                   self._nws))
        return self.losses, self.grads

    def voxel_samples(self):
        """Trilinear evaluations of one step: every frustum sample once forward
        and once in the backward walk (a LINEAR mip sample between two levels
        counts once)."""
        total = 0
        for k in range(self.n):
            c = self._arr[k]
            total += 2 * c.num_views * c.output_shape[0] * c.output_shape[1] * c.output_shape[2]
        return total

    def time_ms(self, reps=5):
        self.run()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            self.run()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    def autotune(self, candidates=(0, 1, 2), reps=5):
        """Per grid-size class, the brick whose whole-batch time is lowest with
        the other classes fixed. Returns {grid: (brick, {brick: ms})}."""
        classes = {}
        for k in range(self.n):
            classes.setdefault(tuple(self._arr[k].grid_dhw), []).append(k)
        result = {}
        for grid, members in sorted(classes.items(), key=lambda kv: -np.prod(kv[0])):
            times = {}
            for b in candidates:
                cur = self.bricks
                for k in members:
                    cur[k] = b
                self.set_bricks(cur)
                times[b] = self.time_ms(reps)
            best = min(times, key=times.get)
            cur = self.bricks
            for k in members:
                cur[k] = best
            self.set_bricks(cur)
            result[grid] = (best, times)
        return result
