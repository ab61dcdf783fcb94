"""Global transport chain: one density advected through per-frame velocity
fields, every frame rendered and compared with its targets, the summed loss
back-propagated through the whole chain with checkpointing.

Frames f = 0..T-1 hold densities d_f with d_{f+1} = advect(d_f, v_f)
(State.density_advected, phitest/render/data_structures.py:1173-1176);
frame f's loss is the mean squared rgb error of its render against its
targets. The backward follows the density-gradient warping of
optStep_sequence in BWD order (phitest/render/optimization.py:592-602,
974-1047): lambda_f, the gradient arriving at d_f, is

  EXACT      lambda_f = g_f + A_f^T lambda_{f+1}                 (the true chain rule)
  REFERENCE  lambda_f = (1-a) g_f + a A_f^T lambda_{f+1}, lambda_{T-1} = (1-a) g_{T-1}
             for a decay a > 0; g_f + A_f^T lambda_{f+1} for a = 0

with g_f the frame's own render-loss gradient, A_f^T the advection adjoint
and a the decay (warp_dens_grads_decay). The gradient applied to a frame is
its own g_f plus the warped term; in REFERENCE mode at decay 0 the warped term
is divided by num_step = T-1-f, the number of frames already visited in BWD
order (optimization.py:1017-1018), while the accumulator stays unscaled
(:1031). So grad d_0 = g_0 + s_0 (A_0^T lambda_1) and the velocity gradient of
frame f is s_f times the velocity part of A_f^T lambda_{f+1}, with s_f =
1/(T-1-f) at decay 0 in REFERENCE mode and 1 otherwise; the velocity uses its
own decay (warp_vel_grads_decay, :1019-1020), which defaults to the density's.

Checkpointing: the forward keeps d_f only at every k-th frame; the backward
recomputes each k-frame segment (advection with its saved MacCormack state,
then render forward and backward per frame) from its checkpoint. Peak device
memory therefore holds ceil(T/k) checkpoints plus one segment instead of all T
frames.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .ops import BOUNDARY, _enum, _np_ptr
from .pipeline import CLAMP, _Cameras


class TransportChain(object):
    def __init__(self, grid_dhw, matrix_m, matrix_v, matrix_p, frustum_params, output_shape, depth_step,
                 light=(1.0, 1.0, 1.0), order=2, clamp="MC_SMOOTH", dt=1.0, boundary="BORDER", checkpoint_every=4,
                 grad_mode="EXACT", decay=0.0, device="cuda", vel_decay=None):
        D, H, W = [int(s) for s in grid_dhw]
        self.shape = (1, D, H, W, 1)
        self.cams = _Cameras(matrix_m, matrix_v, matrix_p, frustum_params, output_shape)
        self.depth_step = float(depth_step)
        self.light = (ctypes.c_float * 3)(*[float(x) for x in light])
        self.order, self.clamp, self.dt = int(order), CLAMP[str(clamp).upper()], float(dt)
        self.boundary = _enum(BOUNDARY, boundary, "boundary")
        self.k = max(1, int(checkpoint_every))
        self.grad_mode = str(grad_mode).upper()
        if self.grad_mode not in ("EXACT", "REFERENCE"):
            raise _lib.InvalidArgumentError("invalid grad_mode '%s'" % grad_mode)
        self.decay = float(decay)
        self.vel_decay = self.decay if vel_decay is None else float(vel_decay)
        self.device = torch.device(device)
        L = _lib.lib()
        self._L = L
        self._shape = _lib.i64_array(self.shape)
        self._nsaved = L.gft_advect_saved_bytes(self._shape)
        self._nws = L.gft_advect_bwd_workspace(self._shape)
        self._nrsaved = L.gft_render_density_saved_bytes(self._shape, self.cams.nCam, self.cams.c_out, self.boundary)
        nV, h, w = self.cams.nCam, self.cams.out_dhw[1], self.cams.out_dhw[2]
        self.loss_scale = 1.0 / (nV * h * w * 3)
        dev = self.device
        self.images = torch.empty((1, nV, h, w, 4), device=dev)
        self.ws = torch.empty((max(self._nws, 1),), dtype=torch.uint8, device=dev)
        self.render_saved = torch.empty((max(self._nrsaved, 1),), dtype=torch.uint8, device=dev)
        self._nfws = max(L.gft_render_density_fwd_workspace(self._shape, self.boundary),
                         L.gft_render_density_bwd_workspace(self._shape, self.cams.nCam, self.cams.c_out,
                                                            self.boundary))
        self.render_fwd_ws = torch.empty((max(self._nfws, 1),), dtype=torch.uint8, device=dev)
        self.frame_loss = None

    # ---------------------------------------------------------------- kernels
    def _advect(self, d, v, out, saved):
        L, s = self._L, _lib.stream_of(d)
        _lib.check(L.gft_advect_fwd(s, _lib.ptr(d), self._shape, _lib.ptr(v), self.dt, self.order, self.clamp,
                                    _lib.ptr(out), _lib.ptr(saved), self._nsaved))

    def _advect_bwd(self, d, v, saved, lam, gd, gv):
        L, s = self._L, _lib.stream_of(d)
        _lib.check(L.gft_advect_bwd(s, _lib.ptr(d), self._shape, _lib.ptr(v), self.dt, self.order, self.clamp,
                                    _lib.ptr(saved), _lib.ptr(lam), _lib.ptr(gd), _lib.ptr(gv), _lib.ptr(self.ws),
                                    self._nws))

    def _render_loss_grad(self, d, target, grad, loss):
        L, c, s = self._L, self.cams, _lib.stream_of(d)
        _lib.check(L.gft_render_density_fwd(s, _lib.ptr(d), self._shape, _np_ptr(c.M), c.nM, _np_ptr(c.V),
                                            _np_ptr(c.P), _np_ptr(c.F), c.nCam, c.c_out, self.light, self.depth_step,
                                            self.boundary, _lib.ptr(self.render_fwd_ws), self._nfws,
                                            _lib.ptr(self.render_saved), self._nrsaved, _lib.ptr(self.images)))
        _lib.check(L.gft_render_density_bwd_l2(s, _lib.ptr(d), self._shape, _np_ptr(c.M), c.nM, _np_ptr(c.V),
                                               _np_ptr(c.P), _np_ptr(c.F), c.nCam, c.c_out, self.light,
                                               self.depth_step, self.boundary, _lib.ptr(self.images),
                                               _lib.ptr(target), self.loss_scale, _lib.ptr(self.render_saved),
                                               self._nrsaved, _lib.ptr(self.render_fwd_ws), self._nfws,
                                               _lib.ptr(grad), _lib.ptr(loss)))

    # ------------------------------------------------------------------- run
    def run(self, d0, velocities, targets):
        """Returns (total loss [1] device tensor, grad of d0, list of velocity
        gradients, one per advection). ``targets`` holds one [1,V,H,W,3]
        tensor per frame; ``velocities`` at least len(targets)-1 centred
        [1,D,H,W,3] fields."""
        T = len(targets)
        if T < 1:
            raise _lib.InvalidArgumentError("the chain needs at least one frame")
        if len(velocities) < T - 1:
            raise _lib.InvalidArgumentError("%d frames need %d velocity fields" % (T, T - 1))
        dev = self.device
        k = self.k
        # forward: keep every k-th density
        ckpt = {0: d0}
        d = d0
        scratch_saved = torch.empty((self._nsaved,), dtype=torch.uint8, device=dev)
        for f in range(T - 1):
            nxt = torch.empty(self.shape, device=dev)
            self._advect(d, velocities[f], nxt, scratch_saved)
            if (f + 1) % k == 0:
                ckpt[f + 1] = nxt
            d = nxt
        del scratch_saved
        losses = torch.zeros((T,), device=dev)
        grad_v = [None] * (T - 1)
        lam = None
        gd0 = None
        ref = self.grad_mode == "REFERENCE"
        a = self.decay
        decayed = ref and a > 0  # optimization.py:1031: lerp(l, d, a) if a > 0 else l + d
        g_own = torch.empty(self.shape, device=dev)
        gd = torch.empty(self.shape, device=dev)
        starts = list(range(0, T, k))
        for s0 in reversed(starts):
            s1 = min(s0 + k, T)
            # recompute the segment's densities with the saved state of every advection leaving it
            ds = [ckpt[s0]]
            saved = []
            for f in range(s0, s1):
                if f == T - 1:
                    break
                sv = torch.empty((self._nsaved,), dtype=torch.uint8, device=dev)
                nxt = torch.empty(self.shape, device=dev)
                self._advect(ds[-1], velocities[f], nxt, sv)
                saved.append(sv)
                if f + 1 < s1:
                    ds.append(nxt)
            for f in reversed(range(s0, s1)):
                df = ds[f - s0]
                self._render_loss_grad(df, targets[f], g_own, losses[f:f + 1])
                if f == T - 1:
                    if T == 1:
                        gd0 = g_own.clone()
                    lam = (1.0 - a) * g_own if decayed else g_own.clone()
                    continue
                gv = torch.empty(tuple(self.shape[:4]) + (3,), device=dev)
                self._advect_bwd(df, velocities[f], saved[f - s0], lam, gd, gv)
                num_step = T - 1 - f
                if ref and self.vel_decay == 0:
                    gv.mul_(1.0 / num_step)
                grad_v[f] = gv
                if f == 0:
                    gd0 = g_own + gd * (1.0 / num_step if (ref and a == 0) else 1.0)
                    continue
                lam = (1.0 - a) * g_own + a * gd if decayed else g_own + gd
            del ds, saved
        self.frame_loss = losses
        return losses.sum(), gd0, grad_v
