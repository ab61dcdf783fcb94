"""Loss and optimizer epilogue of the reconstruction step on the HIP kernels.

``divergence_loss`` is loss_vel_divergence (phitest/render/optimization.py:
824-838) with its default SE loss function over VelocityGrid.divergence
(data_structures.py:929-936): scale * sum(div^2), loss and gradient in one
deterministic call. ``AdamOptimizer`` applies the update that
tf.train.AdamOptimizer documents (the reference's optimizer,
reconstruct_sequence.py:1899-1905) to up to 8 tensors per kernel launch. The
image losses with background compositing are folded into the fused render
backward (``TransportRenderStep(loss=..., background=...)``).
"""
import ctypes

import torch

from . import _lib

_MAX_ADAM_TENSORS = 8


def _components(velocity):
    if hasattr(velocity, "x") and hasattr(velocity, "centered_shape"):
        return velocity.x, velocity.y, velocity.z, list(velocity.centered_shape)
    x, y, z = velocity
    return x, y, z, [int(z.shape[1]) - 1, int(x.shape[2]), int(x.shape[3]) - 1]


def _check_components(x, y, z, shape):
    D, H, W = shape
    want = [(1, D, H, W + 1, 1), (1, D, H + 1, W, 1), (1, D + 1, H, W, 1)]
    for c, w, n in zip((x, y, z), want, "xyz"):
        if not isinstance(c, torch.Tensor) or not c.is_cuda or c.dtype != torch.float32:
            raise _lib.InvalidArgumentError("velocity %s must be a float32 CUDA (HIP) tensor" % n)
        if tuple(c.shape) != w:
            raise _lib.InvalidArgumentError("velocity %s has shape %s, expected %s" % (n, list(c.shape), list(w)))


def divergence_loss_into(velocity, world_scale=(1.0, 1.0, 1.0), scale=1.0, grads=None, accumulate=False):
    """loss = scale * sum(div^2) as a device scalar; with ``grads`` (three
    tensors of the component shapes) the gradient is written into them, or
    added when ``accumulate``. ``velocity`` is a VelocityGrid or (x, y, z)."""
    x, y, z, shape = _components(velocity)
    x, y, z = x.contiguous(), y.contiguous(), z.contiguous()
    _check_components(x, y, z, shape)
    L = _lib.lib()
    cs = _lib.i32_array(shape)
    wsb = L.gft_divergence_loss_workspace(cs)
    ws = torch.empty((max(wsb, 1),), dtype=torch.uint8, device=x.device)
    loss = torch.empty((1,), dtype=torch.float32, device=x.device)
    wsc = (ctypes.c_float * 3)(*[float(w) for w in world_scale])
    if grads is not None:
        for g, c in zip(grads, (x, y, z)):
            if tuple(g.shape) != tuple(c.shape) or not g.is_contiguous():
                raise _lib.InvalidArgumentError("gradient buffers must be contiguous and match the components")
        gp = [_lib.ptr(g) for g in grads]
    else:
        gp = [None, None, None]
    _lib.check(L.gft_divergence_loss(_lib.stream_of(x), _lib.ptr(x), _lib.ptr(y), _lib.ptr(z), cs, wsc, float(scale),
                                     int(bool(accumulate)), _lib.ptr(ws), wsb, gp[0], gp[1], gp[2],
                                     _lib.ptr(loss)))
    return loss[0]


class _DivergenceLossFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, z, world_scale, scale):
        if any(ctx.needs_input_grad[:3]):
            grads = [torch.empty_like(c) for c in (x, y, z)]
            loss = divergence_loss_into((x, y, z), world_scale, scale, grads=grads)
            ctx.save_for_backward(*grads)
        else:
            loss = divergence_loss_into((x, y, z), world_scale, scale)
        return loss

    @staticmethod
    def backward(ctx, g):
        gx, gy, gz = ctx.saved_tensors
        return gx * g, gy * g, gz * g, None, None


def divergence_loss(velocity, world_scale=(1.0, 1.0, 1.0), scale=1.0):
    """Differentiable scale * sum(div^2) of a staggered velocity (the gradient
    is computed in the forward call, by the same kernel pair)."""
    x, y, z, shape = _components(velocity)
    return _DivergenceLossFn.apply(x.contiguous(), y.contiguous(), z.contiguous(),
                                   tuple(float(w) for w in world_scale), float(scale))


class AdamOptimizer(object):
    """tf.train.AdamOptimizer(learning_rate, beta1, beta2, epsilon) on device
    tensors: ``apply_gradients([(grad, var), ...])`` updates every var in
    place, 8 tensors per launch; the moments live here, zero-initialised, and
    one step counter serves all vars as TF's beta-power accumulators do."""

    def __init__(self, learning_rate=0.001, beta1=0.9, beta2=0.999, epsilon=1e-8, device_step=False):
        """device_step: keep the update count in a device counter (read and
        advanced by the kernels), so that a captured graph of the optimisation
        step advances it on every replay; ``iterations`` then counts host
        calls only."""
        self.learning_rate = learning_rate
        self.beta1, self.beta2, self.epsilon = float(beta1), float(beta2), float(epsilon)
        self.iterations = 0
        self._slots = {}
        self.device_step = bool(device_step)
        self._counter = None

    def _slot(self, var):
        key = id(var)
        if key not in self._slots:
            self._slots[key] = (var, torch.zeros_like(var), torch.zeros_like(var))
        return self._slots[key][1], self._slots[key][2]

    def get_slot(self, var, name):
        m, v = self._slot(var)
        return m if name == "m" else v

    def apply_gradients(self, grads_and_vars):
        pairs = [(g, v) for g, v in grads_and_vars if g is not None]
        for g, v in pairs:
            if not (v.is_cuda and v.dtype == torch.float32 and v.is_contiguous()):
                raise _lib.InvalidArgumentError("variables must be contiguous float32 CUDA (HIP) tensors")
            if tuple(g.shape) != tuple(v.shape):
                raise _lib.InvalidArgumentError("gradient shape %s does not match variable %s"
                                                % (list(g.shape), list(v.shape)))
        self.iterations += 1
        lr = self.learning_rate() if callable(self.learning_rate) else self.learning_rate
        L = _lib.lib()
        for first in range(0, len(pairs), _MAX_ADAM_TENSORS):
            chunk = pairs[first:first + _MAX_ADAM_TENSORS]
            n = len(chunk)
            grads = [g.detach().contiguous() for g, _ in chunk]
            vs = [v for _, v in chunk]
            ms = [self._slot(v) for v in vs]
            arr = lambda ts: (ctypes.c_void_p * n)(*[t.data_ptr() for t in ts])  # noqa: E731
            sizes = (ctypes.c_int64 * n)(*[v.numel() for v in vs])
            with torch.no_grad():
                if self.device_step:
                    groups = (len(pairs) + _MAX_ADAM_TENSORS - 1) // _MAX_ADAM_TENSORS
                    if self._counter is None:
                        self._counter = torch.zeros(groups, dtype=torch.int64, device=vs[0].device)
                    if groups > self._counter.numel():
                        raise _lib.InvalidArgumentError("device_step keeps one counter per group of 8 variables: "
                                                        "pass the same variables on every call")
                    # one counter per launch group, all advancing together
                    cnt = self._counter[first // _MAX_ADAM_TENSORS:]
                    _lib.check(L.gft_adam_step_device(_lib.stream_of(vs[0]), n, arr(vs), arr(grads),
                                                      arr([m for m, _ in ms]), arr([s for _, s in ms]), sizes,
                                                      float(lr), self.beta1, self.beta2, self.epsilon,
                                                      _lib.ptr(cnt)))
                else:
                    _lib.check(L.gft_adam_step(_lib.stream_of(vs[0]), n, arr(vs), arr(grads),
                                               arr([m for m, _ in ms]), arr([s for _, s in ms]), sizes, float(lr),
                                               self.beta1, self.beta2, self.epsilon, self.iterations))
        return self.iterations

    @property
    def device_iterations(self):
        """The device update count (device_step only)."""
        return int(self._counter[0].item()) if self._counter is not None else 0
