"""PyTorch-ROCm ops with the reference's op signatures.

Mirrors phitest/render/cuda/ops_loader.py: ``raymarching_ops``,
``sampling_ops`` and ``blending_ops`` namespaces exposing the same op names,
argument names, attribute names and defaults, each differentiable through a
torch.autograd.Function whose backward calls the matching gradient kernel
(the gradient wiring of ops_loader.py:14-85).

Grids are CUDA float32 tensors in NDHWC layout. Matrices, frustum parameters
and shapes are host values (numpy arrays, lists or CPU tensors), as the
reference passes them in HostMemory. An extra keyword ``grad_mode`` selects
the backward flavour: "REFERENCE" (default; the reference's normalised
gradients) or "EXACT" (the true adjoint).
"""
import numpy as np
import torch

from . import _lib

FILTER = {"NEAREST": 0, "LINEAR": 1, "MIN": 2, "MAX": 3}
BOUNDARY = {"BORDER": 0, "CLAMP": 1, "WRAP": 2}
BLEND = {"BEER_LAMBERT": 0, "ALPHA": 1, "ALPHA_ADDITIVE": 2, "ADDITIVE": 3}
MIP = {"NONE": 0, "NEAREST": 1, "LINEAR": 2}
COORD = {"TRANSFORM_LINDEPTH": 0, "TRANSFORM_LINDEPTH_REVERSE": 1, "TRANSFORM": 2, "TRANSFORM_REVERSE": 3}
GRAD = {"REFERENCE": 0, "EXACT": 1}


def _enum(table, value, what):
    if isinstance(value, bytes):
        value = value.decode()
    key = str(value).upper()
    if key not in table:
        raise _lib.InvalidArgumentError("invalid %s '%s'" % (what, value))
    return table[key]


def _host_f32(x):
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    if isinstance(x, (list, tuple)):
        x = np.stack([np.asarray(a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a, dtype=np.float32)
                      for a in x]) if len(x) and not np.isscalar(x[0]) else np.asarray(x, dtype=np.float32)
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


def _matrices(m, name):
    """Validate a 4x4 matrix or batch thereof (raymarch_grid.cc:97-108)."""
    a = _host_f32(m)
    if a.ndim == 2 and a.shape == (4, 4):
        return a.reshape(1, 4, 4), 1
    if a.ndim == 3 and a.shape[1:] == (4, 4):
        return a, a.shape[0]
    raise _lib.InvalidArgumentError("transformation %s must be a 4x4 square matrix or a batch of 4x4 square matrices: %s"
                                    % (name, list(a.shape)))


def _frustum(f):
    """Validate 6 frustum values or a batch thereof (raymarch_grid.cc:110-121)."""
    a = _host_f32(f)
    if a.ndim == 1 and a.shape[0] == 6:
        return a.reshape(1, 6), 1
    if a.ndim == 2 and a.shape[1] == 6:
        return a, a.shape[0]
    raise _lib.InvalidArgumentError("frustum must be a 1D tensor of 6 elements or a batch thereof: %s" % list(a.shape))


def _shape3(s):
    if isinstance(s, torch.Tensor):
        s = s.detach().cpu().tolist()
    s = [int(v) for v in np.asarray(s).reshape(-1)]
    if len(s) != 3:
        raise _lib.InvalidArgumentError("Invalid output_shape")
    for i, v in enumerate(s):
        if v < 1:
            raise _lib.InvalidArgumentError("All shape dimensions must be positive, dimension %d is %d" % (i, v))
    return s


def _grid(t, name="input"):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise _lib.InvalidArgumentError("%s must be a CUDA (HIP) tensor" % name)
    if t.dtype != torch.float32:
        raise _lib.InvalidArgumentError("%s must be float32" % name)
    if t.dim() != 5:
        raise _lib.InvalidArgumentError("Invalid %s shape (NDHWC): %s" % (name, list(t.shape)))
    return t.contiguous()


def _cams(matrix_m, matrix_v, matrix_p, frustum_params):
    M, nM = _matrices(matrix_m, "matrix_m")
    V, nV = _matrices(matrix_v, "matrix_v")
    P, nP = _matrices(matrix_p, "matrix_p")
    F, nF = _frustum(frustum_params)
    if not (nV == nP == nF):
        raise _lib.InvalidArgumentError("camera batch size mismatch")
    return M, nM, V, P, F, nV


def _np_ptr(a):
    return a.ctypes.data_as(_lib.ctypes.c_void_p)


# ------------------------------------------------------------------- raymarch
def _raymarch_fwd_impl(input, M, nM, V, P, F, nCam, out_dhw, interpolation, boundary, blending_mode, keep_dims,
                       separate_camera_batch):
    N, D, H, W, C = input.shape
    if C > 4:
        raise _lib.InvalidArgumentError("Invalid input shape (NDHWC): %s" % list(input.shape))
    if nM != N:
        raise _lib.InvalidArgumentError("model matrix batch size mismatch")
    nv = nCam if separate_camera_batch else 1
    if not separate_camera_batch and nCam != N:
        raise _lib.InvalidArgumentError("camera batch must match data batch when not using global sampling.")
    if keep_dims:
        out = torch.empty((N, nv, 1, out_dhw[1], out_dhw[2], C), device=input.device, dtype=torch.float32)
    else:
        out = torch.empty((1, nv, out_dhw[1], out_dhw[2], C), device=input.device, dtype=torch.float32)
    L = _lib.lib()
    _lib.check(L.gft_raymarch_grid_transform_fwd(
        _lib.stream_of(input), _lib.ptr(input), _lib.i64_array(input.shape), _np_ptr(M), nM, _np_ptr(V), _np_ptr(P),
        _np_ptr(F), nCam, _lib.i32_array(out_dhw), interpolation, boundary, blending_mode, int(bool(keep_dims)),
        int(bool(separate_camera_batch)), _lib.ptr(out)))
    return out


def _raymarch_bwd_impl(input, output, output_grad, M, nM, V, P, F, nCam, out_dhw, interpolation, boundary,
                       blending_mode, keep_dims, separate_camera_batch, grad_mode):
    output = output.contiguous()
    output_grad = output_grad.contiguous()
    if tuple(output.shape) != tuple(output_grad.shape):
        raise _lib.InvalidArgumentError("Shapes of output grid and output gradients must match: %s %s"
                                        % (list(output.shape), list(output_grad.shape)))
    L = _lib.lib()
    shape = _lib.i64_array(input.shape)
    ws_bytes = L.gft_raymarch_grid_transform_bwd_workspace(shape, grad_mode)
    ws = torch.empty((max(int(ws_bytes), 1),), dtype=torch.uint8, device=input.device)
    grad = torch.empty_like(input)
    _lib.check(L.gft_raymarch_grid_transform_bwd(
        _lib.stream_of(input), _lib.ptr(input), shape, _lib.ptr(output), _lib.ptr(output_grad), _np_ptr(M), nM,
        _np_ptr(V), _np_ptr(P), _np_ptr(F), nCam, _lib.i32_array(out_dhw), interpolation, boundary, blending_mode,
        int(bool(keep_dims)), int(bool(separate_camera_batch)), grad_mode, _lib.ptr(ws), ws_bytes, _lib.ptr(grad)))
    return grad


class _RaymarchFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, input, cams, out_dhw, attrs):
        M, nM, V, P, F, nCam = cams
        out = _raymarch_fwd_impl(input, M, nM, V, P, F, nCam, out_dhw, *attrs[:5])
        ctx.save_for_backward(input, out)
        ctx.cams, ctx.out_dhw, ctx.attrs = cams, out_dhw, attrs
        return out

    @staticmethod
    def backward(ctx, grad_out):
        input, out = ctx.saved_tensors
        M, nM, V, P, F, nCam = ctx.cams
        interpolation, boundary, blending_mode, keep_dims, sep, grad_mode = ctx.attrs
        if interpolation != FILTER["LINEAR"]:
            raise _lib.InvalidArgumentError("raymarch gradient supports interpolation LINEAR only")
        g = _raymarch_bwd_impl(input, out, grad_out, M, nM, V, P, F, nCam, ctx.out_dhw, interpolation, boundary,
                               blending_mode, keep_dims, sep, grad_mode)
        return g, None, None, None


class _RaymarchingOps(object):
    """``raymarching_ops`` of ops_loader.py:64-85."""

    @staticmethod
    def raymarch_grid_transform(input, matrix_m, matrix_v, matrix_p, frustum_params, output_shape,
                                interpolation="LINEAR", boundary="BORDER", blending_mode="BEER_LAMBERT",
                                keep_dims=False, separate_camera_batch=True, grad_mode="REFERENCE"):
        """RaymarchGridTransform (raymarch_grid.cc:28-52): NDHWC -> [1,V,H,W,C]
        ([N,V,1,H,W,C] with keep_dims)."""
        input = _grid(input)
        cams = _cams(matrix_m, matrix_v, matrix_p, frustum_params)
        out_dhw = _shape3(output_shape)
        attrs = (_enum(FILTER, interpolation, "filter mode"), _enum(BOUNDARY, boundary, "boundary mode"),
                 _enum(BLEND, blending_mode, "blend mode"), bool(keep_dims), bool(separate_camera_batch),
                 _enum(GRAD, grad_mode, "grad_mode"))
        return _RaymarchFn.apply(input, cams, out_dhw, attrs)

    @staticmethod
    def raymarch_grid_transform_grad(input, output, output_grad, matrix_m, matrix_v, matrix_p, frustum_params,
                                     output_shape, interpolation="LINEAR", boundary="BORDER",
                                     blending_mode="BEER_LAMBERT", keep_dims=False, separate_camera_batch=True,
                                     grad_mode="REFERENCE"):
        """RaymarchGridTransformGrad (raymarch_grid.cc:56-76)."""
        input = _grid(input)
        M, nM, V, P, F, nCam = _cams(matrix_m, matrix_v, matrix_p, frustum_params)
        fm = _enum(FILTER, interpolation, "filter mode")
        if fm != FILTER["LINEAR"]:
            raise _lib.InvalidArgumentError("raymarch gradient supports interpolation LINEAR only")
        return _raymarch_bwd_impl(input, output.detach(), output_grad.detach(), M, nM, V, P, F, nCam,
                                  _shape3(output_shape), fm, _enum(BOUNDARY, boundary, "boundary mode"),
                                  _enum(BLEND, blending_mode, "blend mode"), bool(keep_dims),
                                  bool(separate_camera_batch), _enum(GRAD, grad_mode, "grad_mode"))


raymarching_ops = _RaymarchingOps()
raymarch_grid_transform = raymarching_ops.raymarch_grid_transform
