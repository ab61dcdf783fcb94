"""Fused transport/render path: advect -> render -> L2 loss, forward and
exact backward, on the HIP kernels of libgft.

``advect`` and ``render_density`` are differentiable building blocks
(torch.autograd.Function). ``TransportRenderStep`` runs one optimisation
step's fused forward and backward on preallocated buffers, issuing only HIP
kernel launches on torch's current stream, so it can be captured in a
hipGraph (torch.cuda.CUDAGraph).

Semantics: the density warp of State.density_advected / VelocityGrid.warp
(phitest/render/data_structures.py:263-266,870-905) with a centred velocity
LUT (lut = -v*dt, data_structures.py:864-868), followed by
Renderer.render_density with an ambient light (renderer.py:603-651) and the
density target loss (optimization.py:396-412) with an L2 loss function.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .ops import BOUNDARY, _cams, _enum, _np_ptr, _shape3

CLAMP = {"NONE": 0, "MC_SMOOTH": 1, "NEGATIVE": 2}


def _scalar_grid(t, name, dtypes=(torch.float32,)):
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype not in dtypes:
        raise _lib.InvalidArgumentError("%s must be a %s CUDA tensor" % (name, " or ".join(str(d)[6:] for d in dtypes)))
    if t.dim() != 5 or t.shape[0] != 1 or t.shape[4] != 1:
        raise _lib.InvalidArgumentError("%s must be [1,D,H,W,1], got %s" % (name, list(t.shape)))
    return t.contiguous()


def _velocity(v, dens):
    """A centred velocity of the density's grid and storage type."""
    if not isinstance(v, torch.Tensor) or not v.is_cuda or v.dtype != dens.dtype:
        raise _lib.InvalidArgumentError("velocity must be a %s CUDA tensor (the density's type)" % str(dens.dtype)[6:])
    if tuple(v.shape) != tuple(dens.shape[:4]) + (3,):
        raise _lib.InvalidArgumentError("Shape mismatch: velocity %s for grid %s" % (list(v.shape), list(dens.shape)))
    return v.contiguous()


STORAGE = {"float32": torch.float32, "bfloat16": torch.bfloat16}


def _advect_calls(dtype):
    """The forward and backward entry points for density/velocity storage."""
    L = _lib.lib()
    if dtype == torch.bfloat16:
        return L.gft_advect_fwd_bf16, L.gft_advect_bwd_bf16
    return L.gft_advect_fwd, L.gft_advect_bwd


# -------------------------------------------------------------- advection
def advect_forward(d, vel, dt=1.0, order=2, clamp="MC_SMOOTH", out=None, saved=None):
    """d, vel float32 or bfloat16 (both the same); the output is float32."""
    L = _lib.lib()
    shape = _lib.i64_array(d.shape)
    nsaved = L.gft_advect_saved_bytes(shape)
    if saved is None:
        saved = torch.empty((nsaved,), dtype=torch.uint8, device=d.device)
    if out is None:
        out = torch.empty(d.shape, dtype=torch.float32, device=d.device)
    fwd, _ = _advect_calls(d.dtype)
    _lib.check(fwd(_lib.stream_of(d), _lib.ptr(d), shape, _lib.ptr(vel), float(dt), int(order), CLAMP[clamp],
                   _lib.ptr(out), _lib.ptr(saved), nsaved))
    return out, saved


def advect_backward(d, vel, saved, grad_out, dt=1.0, order=2, clamp="MC_SMOOTH", gd=None, gv=None, ws=None):
    """Gradients are float32 whatever the storage of d and vel."""
    L = _lib.lib()
    shape = _lib.i64_array(d.shape)
    nws = L.gft_advect_bwd_workspace(shape)
    if ws is None:
        ws = torch.empty((nws,), dtype=torch.uint8, device=d.device)
    gd = torch.empty(d.shape, dtype=torch.float32, device=d.device) if gd is None else gd
    gv = torch.empty(vel.shape, dtype=torch.float32, device=d.device) if gv is None else gv
    _, bwd = _advect_calls(d.dtype)
    _lib.check(bwd(_lib.stream_of(d), _lib.ptr(d), shape, _lib.ptr(vel), float(dt), int(order), CLAMP[clamp],
                   _lib.ptr(saved), _lib.ptr(grad_out.float().contiguous()), _lib.ptr(gd), _lib.ptr(gv),
                   _lib.ptr(ws), nws))
    return gd, gv


class _AdvectFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, d, vel, dt, order, clamp):
        out, saved = advect_forward(d, vel, dt, order, clamp)
        ctx.save_for_backward(d, vel, saved)
        ctx.cfg = (dt, order, clamp)
        return out

    @staticmethod
    def backward(ctx, g):
        d, vel, saved = ctx.saved_tensors
        dt, order, clamp = ctx.cfg
        gd, gv = advect_backward(d, vel, saved, g, dt, order, clamp)
        return gd, gv, None, None, None


def advect(d, velocity, dt=1.0, order=2, clamp="MC_SMOOTH"):
    """Warp a scalar grid [1,D,H,W,1] by a centred velocity [1,D,H,W,3]
    (VelocityGrid.warp with a CLAMP LuT sampler); exact gradients for both.
    Both may be stored in bfloat16 (then sampled in float32); the result is
    float32 and autograd casts the float32 gradients to the inputs' type."""
    d = _scalar_grid(d, "density", (torch.float32, torch.bfloat16))
    velocity = _velocity(velocity, d)
    if order not in (1, 2):
        raise _lib.InvalidArgumentError("Unsupported warp order '%s'" % order)
    clamp = str(clamp).upper()
    if clamp == "MC":
        raise NotImplementedError("MIM and MAX warp sampling have wrong gradients.")
    if clamp not in CLAMP:
        raise _lib.InvalidArgumentError("invalid warp clamp '%s'" % clamp)
    return _AdvectFn.apply(d, velocity, float(dt), int(order), clamp)


# ----------------------------------------------------------------- render
class _Cameras(object):
    """Host camera arrays kept alive for the C calls."""

    def __init__(self, M, V, P, F, out_dhw):
        self.M, self.nM, self.V, self.P, self.F, self.nCam = _cams(M, V, P, F)
        self.out_dhw = _shape3(out_dhw)
        self.c_out = _lib.i32_array(self.out_dhw)


def render_saved_bytes(shape, cams, boundary="BORDER"):
    """Bytes of the buffer a forward fills for its backward (0 unless BORDER)."""
    return int(_lib.lib().gft_render_density_saved_bytes(_lib.i64_array(shape), cams.nCam, cams.c_out,
                                                         _enum(BOUNDARY, boundary, "boundary")))


def render_density_forward(d, cams, depth_step, light=(1.0, 1.0, 1.0), boundary="BORDER", out=None, ws=None,
                           saved=None):
    """``saved``: optional uint8 device buffer of render_saved_bytes() that
    receives what the backward of this forward needs."""
    L = _lib.lib()
    if out is None:
        out = torch.empty((1, cams.nCam, cams.out_dhw[1], cams.out_dhw[2], 4), device=d.device, dtype=torch.float32)
    lt = (ctypes.c_float * 3)(*[float(x) for x in light])
    shape = _lib.i64_array(d.shape)
    bm = _enum(BOUNDARY, boundary, "boundary")
    nws = L.gft_render_density_fwd_workspace(shape, bm)
    if ws is None:
        ws = torch.empty((max(int(nws), 1),), dtype=torch.uint8, device=d.device)
    sp, sn = (_lib.ptr(saved), saved.numel()) if saved is not None else (None, 0)
    _lib.check(L.gft_render_density_fwd(_lib.stream_of(d), _lib.ptr(d), shape, _np_ptr(cams.M), cams.nM,
                                        _np_ptr(cams.V), _np_ptr(cams.P), _np_ptr(cams.F), cams.nCam, cams.c_out, lt,
                                        float(depth_step), bm, _lib.ptr(ws), nws, sp, sn, _lib.ptr(out)))
    return out


def _align256(n):
    return (int(n) + 255) // 256 * 256


def _rewritten_tail(saved, n_ambient, cams):
    """The part of a saved buffer that a render backward rewrites in place: the
    per-pixel records and the per-tile loss partials, which end the ambient
    render's layout at byte ``n_ambient`` (the per-sample values before them
    are only read). Bounded from above by 16 + 4 bytes per pixel, each block
    256-byte aligned, so the slice may also cover the last per-sample values.
    A view through ``.data``: restoring it must not count as an in-place change
    of a tensor autograd saved."""
    plane = int(cams.out_dhw[1]) * int(cams.out_dhw[2]) * int(cams.nCam)
    tail = min(int(n_ambient), _align256(plane * 16) + _align256(plane * 4))
    return saved.data[int(n_ambient) - tail:int(n_ambient)]


def _keep_or_restore(ctx, tail):
    """The first backward keeps a copy of what it will overwrite (a few bytes
    per pixel, not the per-sample values); a later backward of a retained
    graph puts it back first."""
    if getattr(ctx, "pristine", None) is None:
        ctx.pristine = tail.clone()
    else:
        tail.copy_(ctx.pristine)


class _RenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, d, cams, depth_step, light, boundary):
        nsaved = render_saved_bytes(d.shape, cams, boundary)
        saved = torch.empty((max(nsaved, 1),), dtype=torch.uint8, device=d.device)
        img = render_density_forward(d, cams, depth_step, light, boundary, saved=saved)
        ctx.save_for_backward(d, img, saved)
        ctx.cfg = (cams, depth_step, light, boundary)
        return img

    @staticmethod
    def backward(ctx, g):
        d, img, saved = ctx.saved_tensors
        cams, depth_step, light, boundary = ctx.cfg
        gd = torch.empty_like(d)
        lt = (ctypes.c_float * 3)(*[float(x) for x in light])
        L = _lib.lib()
        shape = _lib.i64_array(d.shape)
        bm = _enum(BOUNDARY, boundary, "boundary")
        if saved.numel() > 1:
            _keep_or_restore(ctx, _rewritten_tail(saved, saved.numel(), cams))
        nws = int(L.gft_render_density_bwd_workspace(shape, cams.nCam, cams.c_out, bm))
        ws = torch.empty((max(nws, 1),), dtype=torch.uint8, device=d.device)
        _lib.check(L.gft_render_density_bwd(
            _lib.stream_of(d), _lib.ptr(d), shape, _np_ptr(cams.M), cams.nM, _np_ptr(cams.V),
            _np_ptr(cams.P), _np_ptr(cams.F), cams.nCam, cams.c_out, lt, float(depth_step), bm, _lib.ptr(img),
            _lib.ptr(g.contiguous()), _lib.ptr(saved), saved.numel(), _lib.ptr(ws), nws, _lib.ptr(gd)))
        return gd, None, None, None, None


def render_density(d, matrix_m, matrix_v, matrix_p, frustum_params, output_shape, depth_step,
                   light=(1.0, 1.0, 1.0), boundary="BORDER"):
    """Images [1,V,H,W,4] = (light * E, optical depth) of a density [1,D,H,W,1]
    under an ambient light; differentiable w.r.t. the density (exact)."""
    d = _scalar_grid(d, "density")
    cams = _Cameras(matrix_m, matrix_v, matrix_p, frustum_params, output_shape)
    return _RenderFn.apply(d, cams, float(depth_step), tuple(float(x) for x in light), boundary)


# ----------------------------------------------------------- lit render
def render_lit_saved_bytes(shape, cams):
    return int(_lib.lib().gft_render_density_lit_saved_bytes(_lib.i64_array(shape), cams.nCam, cams.c_out,
                                                             BOUNDARY["BORDER"]))


def render_lit_forward(d, lam, cams, depth_step, color=(1.0, 1.0, 1.0), ambient=0.0, out=None, ws=None, saved=None):
    """Images of a density under a light with per-cell factor ``lam`` plus an
    ambient term (gft_render_density_lit_fwd, BORDER)."""
    L = _lib.lib()
    if out is None:
        out = torch.empty((1, cams.nCam, cams.out_dhw[1], cams.out_dhw[2], 4), device=d.device, dtype=torch.float32)
    lt = (ctypes.c_float * 3)(*[float(x) for x in color])
    shape = _lib.i64_array(d.shape)
    nws = int(L.gft_render_density_lit_fwd_workspace(shape, BOUNDARY["BORDER"]))
    if ws is None:
        ws = torch.empty((max(nws, 1),), dtype=torch.uint8, device=d.device)
    sp, sn = (_lib.ptr(saved), saved.numel()) if saved is not None else (None, 0)
    _lib.check(L.gft_render_density_lit_fwd(_lib.stream_of(d), _lib.ptr(d), _lib.ptr(lam), shape, _np_ptr(cams.M),
                                            cams.nM, _np_ptr(cams.V), _np_ptr(cams.P), _np_ptr(cams.F), cams.nCam,
                                            cams.c_out, lt, float(ambient), float(depth_step), BOUNDARY["BORDER"],
                                            _lib.ptr(ws), nws, sp, sn, _lib.ptr(out)))
    return out


class _LitRenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, d, lam, cams, depth_step, color, ambient):
        saved = torch.empty((render_lit_saved_bytes(d.shape, cams),), dtype=torch.uint8, device=d.device)
        img = render_lit_forward(d, lam, cams, depth_step, color, ambient, saved=saved)
        ctx.save_for_backward(d, lam, img, saved)
        ctx.cfg = (cams, depth_step, color, ambient)
        return img

    @staticmethod
    def backward(ctx, g):
        d, lam, img, saved = ctx.saved_tensors
        cams, depth_step, color, ambient = ctx.cfg
        gd, gl = torch.empty_like(d), torch.empty_like(lam)
        lt = (ctypes.c_float * 3)(*[float(x) for x in color])
        _keep_or_restore(ctx, _rewritten_tail(saved, render_saved_bytes(d.shape, cams, "BORDER"), cams))
        _lib.check(_lib.lib().gft_render_density_lit_bwd(
            _lib.stream_of(d), _lib.ptr(d), _lib.ptr(lam), _lib.i64_array(d.shape), _np_ptr(cams.M), cams.nM,
            _np_ptr(cams.V), _np_ptr(cams.P), _np_ptr(cams.F), cams.nCam, cams.c_out, lt, float(ambient),
            float(depth_step), BOUNDARY["BORDER"], _lib.ptr(img), _lib.ptr(g.contiguous()), _lib.ptr(saved),
            saved.numel(), _lib.ptr(gd), _lib.ptr(gl)))
        return gd, gl, None, None, None, None


def render_density_lit(d, light_factor, matrix_m, matrix_v, matrix_p, frustum_params, output_shape, depth_step,
                       color=(1.0, 1.0, 1.0), ambient=0.0):
    """Images [1,V,H,W,4] = (color * E, optical depth) of a density [1,D,H,W,1]
    whose light grid is color * d * (light_factor + ambient) (a spot or point
    light with its per-cell factor plus an ambient light of the same colour);
    BORDER; differentiable (exact) w.r.t. the density and the light factor."""
    d = _scalar_grid(d, "density")
    lam = _scalar_grid(light_factor, "light_factor")
    if tuple(lam.shape) != tuple(d.shape):
        raise _lib.InvalidArgumentError("light_factor shape %s differs from the density shape %s" %
                                        (tuple(lam.shape), tuple(d.shape)))
    cams = _Cameras(matrix_m, matrix_v, matrix_p, frustum_params, output_shape)
    return _LitRenderFn.apply(d, lam, cams, float(depth_step), tuple(float(x) for x in color), float(ambient))


# ------------------------------------------------------------- fused step
LOSS = {"SE": 0, "AE": 1}


class TransportRenderStep(object):
    """advect (MacCormack + MC_SMOOTH) -> ambient render -> image loss,
    forward and exact backward, on fixed buffers. ``run(d0, vel, target)``
    leaves the loss in ``self.loss`` and the gradients in
    ``self.grad_density`` / ``self.grad_velocity``.

    The image loss is folded into the render backward: ``loss`` "SE" (squared
    error, the default) or "AE" (absolute error, the reference's default for
    the target losses, optimization.py:114-116), summed and multiplied by
    ``loss_scale`` (default 1 / (views * pixels * 3), the mean). With a
    ``background`` ([1,V,H,W,3] tensor or an rgb triple) the loss compares
    rgb + background * exp(-depth) with the targets (loss_dens_target_raw,
    optimization.py:414-430).

    ``reduction`` picks how the BORDER render backward reduces the sample
    gradients into the density gradient: "gather" (the forward saves two
    things per sample and pixel, the backward sums per voxel block in a
    fixed order: deterministic) or "scatter" (nothing saved: the backward
    walks every ray from far to near recomputing the transmittance and
    scatters with an LDS pre-reduction and float atomics, whose order varies
    between runs). CLAMP and WRAP always scatter."""

    def __init__(self, grid_dhw, matrix_m, matrix_v, matrix_p, frustum_params, output_shape, depth_step,
                 light=(1.0, 1.0, 1.0), order=2, clamp="MC_SMOOTH", dt=1.0, boundary="BORDER", device="cuda",
                 loss="SE", background=None, loss_scale=None, storage="float32", reduction="gather",
                 render_storage=None):
        D, H, W = grid_dhw
        if reduction not in ("gather", "scatter"):
            raise _lib.InvalidArgumentError("reduction must be 'gather' or 'scatter'")
        self.reduction = reduction
        if storage not in STORAGE:
            raise _lib.InvalidArgumentError("storage must be one of %s" % sorted(STORAGE))
        self.storage = STORAGE[storage]
        self._advect_fwd, self._advect_bwd = _advect_calls(self.storage)
        self.shape = (1, D, H, W, 1)
        self.cams = _Cameras(matrix_m, matrix_v, matrix_p, frustum_params, output_shape)
        self.depth_step = float(depth_step)
        self.light = (ctypes.c_float * 3)(*[float(x) for x in light])
        self.order, self.clamp, self.dt = int(order), CLAMP[str(clamp).upper()], float(dt)
        self.boundary = _enum(BOUNDARY, boundary, "boundary")
        L = _lib.lib()
        self._L = L
        self._shape = _lib.i64_array(self.shape)
        dev = torch.device(device)
        nV = self.cams.nCam
        h, w = self.cams.out_dhw[1], self.cams.out_dhw[2]
        self.n_pixels = nV * h * w
        self.loss_scale = 1.0 / (self.n_pixels * 3) if loss_scale is None else float(loss_scale)
        self.loss_kind = _enum(LOSS, loss, "loss")
        self.background = None
        if background is not None:
            if isinstance(background, torch.Tensor):
                bg = background.to(device=dev, dtype=torch.float32)
            else:
                bg = torch.tensor([float(c) for c in background], dtype=torch.float32, device=dev)
            self.background = bg.expand(1, nV, h, w, 3).contiguous()
        self.d1 = torch.empty(self.shape, device=dev)
        self.images = torch.empty((1, nV, h, w, 4), device=dev)
        self.loss = torch.zeros((1,), device=dev)
        self.grad_d1 = torch.empty(self.shape, device=dev)
        self.grad_density = torch.empty(self.shape, device=dev)
        self.grad_velocity = torch.empty((1, D, H, W, 3), device=dev)
        self._nsaved = L.gft_advect_saved_bytes(self._shape)
        self._nws = L.gft_advect_bwd_workspace(self._shape)
        self.saved = torch.empty((self._nsaved,), dtype=torch.uint8, device=dev)
        self.ws = torch.empty((max(self._nws, 1),), dtype=torch.uint8, device=dev)
        self._nrsaved = L.gft_render_density_saved_bytes(self._shape, self.cams.nCam, self.cams.c_out, self.boundary)
        if reduction == "scatter":
            self._nrsaved = 0
        self.render_saved = torch.empty((max(self._nrsaved, 1),), dtype=torch.uint8, device=dev)
        self._rsp = _lib.ptr(self.render_saved) if self._nrsaved else None
        # one workspace for the render forward and backward (stream-ordered): the
        # corner-quad copy each of them writes
        self._nfws = max(L.gft_render_density_fwd_workspace(self._shape, self.boundary),
                         L.gft_render_density_bwd_workspace(self._shape, self.cams.nCam, self.cams.c_out,
                                                            self.boundary))
        self.render_fwd_ws = torch.empty((max(self._nfws, 1),), dtype=torch.uint8, device=dev)
        # bfloat16 storage: the BORDER render with the gather reduction samples a bfloat16 copy of the
        # advected density (the stored precision) unless render_storage is "float32"; CLAMP, WRAP and the
        # scatter walk sample float32
        bf16_ok = self.boundary == BOUNDARY["BORDER"] and reduction == "gather"
        if render_storage is None:
            render_storage = "bfloat16" if (self.storage == torch.bfloat16 and bf16_ok) else "float32"
        if render_storage not in ("float32", "bfloat16") or (render_storage == "bfloat16" and not bf16_ok):
            raise _lib.InvalidArgumentError("render_storage '%s' needs the BORDER boundary and the gather reduction "
                                            "for bfloat16" % render_storage)
        self.render_storage = render_storage
        self._render_fwd = L.gft_render_density_fwd_bf16 if render_storage == "bfloat16" else \
            L.gft_render_density_fwd

    PHASES = ("advect_fwd", "render_fwd", "render_bwd", "advect_bwd")

    def _mark(self, timers, name, begin):
        if timers is not None:
            ev = torch.cuda.Event(enable_timing=True)
            ev.record()
            timers.setdefault(name, []).append(ev) if begin else timers[name].append(ev)

    def forward(self, d0, vel, timers=None):
        L, c, s = self._L, self.cams, _lib.stream_of(d0)
        self._mark(timers, "advect_fwd", True)
        if d0.dtype != self.storage or vel.dtype != self.storage:
            raise _lib.InvalidArgumentError("d0 and vel must be %s (the step's storage)" % str(self.storage)[6:])
        _lib.check(self._advect_fwd(s, _lib.ptr(d0), self._shape, _lib.ptr(vel), self.dt, self.order, self.clamp,
                                    _lib.ptr(self.d1), _lib.ptr(self.saved), self._nsaved))
        self._mark(timers, "advect_fwd", False)
        self._mark(timers, "render_fwd", True)
        _lib.check(self._render_fwd(s, _lib.ptr(self.d1), self._shape, _np_ptr(c.M), c.nM, _np_ptr(c.V),
                                    _np_ptr(c.P), _np_ptr(c.F), c.nCam, c.c_out, self.light, self.depth_step,
                                    self.boundary, _lib.ptr(self.render_fwd_ws), self._nfws,
                                    self._rsp, self._nrsaved, _lib.ptr(self.images)))
        self._mark(timers, "render_fwd", False)
        return self.images

    def backward(self, d0, vel, target, timers=None):
        """Needs the forward just before it: with BORDER the render backward
        consumes (overwrites) the per-sample record the forward saved."""
        L, c, s = self._L, self.cams, _lib.stream_of(d0)
        self._mark(timers, "render_bwd", True)
        bg = _lib.ptr(self.background) if self.background is not None else None
        _lib.check(L.gft_render_density_bwd_loss(s, _lib.ptr(self.d1), self._shape, _np_ptr(c.M), c.nM, _np_ptr(c.V),
                                                 _np_ptr(c.P), _np_ptr(c.F), c.nCam, c.c_out, self.light,
                                                 self.depth_step, self.boundary, _lib.ptr(self.images),
                                                 _lib.ptr(target), bg, self.loss_kind, self.loss_scale,
                                                 self._rsp, self._nrsaved,
                                                 _lib.ptr(self.render_fwd_ws), self._nfws,
                                                 _lib.ptr(self.grad_d1), _lib.ptr(self.loss)))
        self._mark(timers, "render_bwd", False)
        self._mark(timers, "advect_bwd", True)
        _lib.check(self._advect_bwd(s, _lib.ptr(d0), self._shape, _lib.ptr(vel), self.dt, self.order, self.clamp,
                                    _lib.ptr(self.saved), _lib.ptr(self.grad_d1), _lib.ptr(self.grad_density),
                                    _lib.ptr(self.grad_velocity), _lib.ptr(self.ws), self._nws))
        self._mark(timers, "advect_bwd", False)

    def run(self, d0, vel, target, timers=None):
        """Fused forward + backward. ``timers`` (a dict) collects HIP event
        pairs around each phase on the current stream."""
        self.forward(d0, vel, timers)
        self.backward(d0, vel, target, timers)
        return self.loss, self.grad_density, self.grad_velocity

    def algorithmic_bytes(self):
        """Compulsory HBM bytes per phase: each persistent tensor read once
        and written once (intermediates excluded); d0, v and the render's read
        of d' at the storage size s (4 or 2 bytes), the advection's write of
        d', the images and every gradient float32."""
        V = int(np.prod(self.shape))
        N = self.n_pixels
        st = self.storage.itemsize
        return {"advect_fwd": (4 * st + 4) * V, "render_fwd": st * V + 16 * N, "render_bwd": 8 * V + 28 * N,
                "advect_bwd": (20 + 4 * st) * V}

    def kernel_algorithmic_bytes(self):
        """The per-phase compulsory bytes split over the kernels that touch
        those tensors: the first advection pass reads d and v, the second
        writes d'; the render forward reads d' and writes the images; the
        BORDER pixel pass reads images and targets, the gather writes g_d' and
        stands for the read of d' (it consumes the forward's per-sample record
        instead); the CLAMP/WRAP scatter does both; advect_bwd_a reads g_d' and
        v, advect_bwd_b reads d and writes g_d and g_v."""
        V = int(np.prod(self.shape))
        N = self.n_pixels
        st = self.storage.itemsize
        return {"advect_sl": 4 * st * V, "advect_mc": 4 * V, "advect_o1_out": 4 * V, "render_fwd": st * V + 16 * N,
                "render_bwd_pixel": 28 * N, "render_bwd_gather": 8 * V, "render_bwd_scatter": 8 * V + 28 * N,
                "advect_bwd_a": (4 + 3 * st) * V, "advect_bwd_b": (16 + st) * V, "quad_copy": 0, "fill": 0}

    def voxel_samples(self):
        """Trilinear evaluations of one fwd+bwd step: advection fwd a*V
        (a = 4 for MacCormack+MC_SMOOTH: forward, back, min, max; 1 for
        semi-Lagrangian), bwd b*V (b = 2 or 1), raymarch sum W*H*S each way."""
        V = int(np.prod(self.shape))
        ray = self.n_pixels * self.cams.out_dhw[0]
        a, b = (4, 2) if self.order == 2 else (1, 1)
        return a * V + ray + b * V + ray
