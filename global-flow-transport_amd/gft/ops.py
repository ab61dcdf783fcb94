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
    ws_bytes = L.gft_raymarch_grid_transform_bwd_workspace(shape, nCam, _lib.i32_array(out_dhw), boundary,
                                                           int(bool(separate_camera_batch)), grad_mode)
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


# ------------------------------------------------------------------- sampling
def _workspace(nbytes, device):
    return torch.empty((max(int(nbytes), 1),), dtype=torch.uint8, device=device)


def _check_transform_attrs(interpolation, boundary, coordinate_mode, cell_center_offset, mipmapping, num_mipmaps):
    fm = _enum(FILTER, interpolation, "interpolation")
    if fm not in (FILTER["NEAREST"], FILTER["LINEAR"]):
        raise _lib.InvalidArgumentError("interpolation must be NEAREST or LINEAR")
    bm = _enum(BOUNDARY, boundary, "boundary")
    if bm not in (BOUNDARY["BORDER"], BOUNDARY["CLAMP"]):
        raise _lib.InvalidArgumentError("boundary must be BORDER or CLAMP")
    cm = _enum(COORD, coordinate_mode, "coordinate_mode")
    if float(cell_center_offset) != 0.5:
        raise _lib.InvalidArgumentError("Invalid cell center offset for given coordinate mode.")
    mm = _enum(MIP, mipmapping, "mipmapping")
    if mm != MIP["NONE"] and int(num_mipmaps) <= 0:
        raise _lib.InvalidArgumentError("when using mipmaps num_mipmaps must be larger than 0.")
    return fm, bm, cm, mm


def _sample_transform_fwd_impl(input, cams, out_dhw, fm, bm, mm, num_mipmaps, mip_bias, cm, sep):
    M, nM, V, P, F, nCam = cams
    N, D, H, W, C = input.shape
    if nM != N:
        raise _lib.InvalidArgumentError("model matrix batch size mismatch")
    if not sep and nCam != N:
        raise _lib.InvalidArgumentError("camera batch must match data batch when not using global sampling.")
    nv = nCam if sep else 1
    out = torch.empty((N, nv) + tuple(out_dhw) + (C,), device=input.device, dtype=torch.float32)
    L = _lib.lib()
    shape = _lib.i64_array(input.shape)
    wsb = L.gft_sample_grid_transform_workspace(shape, mm, int(num_mipmaps))
    ws = _workspace(wsb, input.device)
    _lib.check(L.gft_sample_grid_transform_fwd(
        _lib.stream_of(input), _lib.ptr(input), shape, _np_ptr(M), nM, _np_ptr(V), _np_ptr(P), _np_ptr(F), nCam,
        _lib.i32_array(out_dhw), fm, bm, mm, int(num_mipmaps), float(mip_bias), cm, 0.5, int(bool(sep)),
        _lib.ptr(ws), wsb, _lib.ptr(out)))
    return out


def _sample_transform_bwd_impl(input, output_grad, cams, fm, bm, mm, num_mipmaps, mip_bias, cm, sep, gm):
    M, nM, V, P, F, nCam = cams
    output_grad = output_grad.contiguous()
    N = input.shape[0]
    if output_grad.dim() != 6 or output_grad.shape[5] > 4:
        raise _lib.InvalidArgumentError("Invalid output_grads shape (NVDHWC): %s" % list(output_grad.shape))
    if output_grad.shape[0] != N:
        raise _lib.InvalidArgumentError("output_grads batch size does not match input batch size.")
    if not sep and output_grad.shape[1] != 1:
        raise _lib.InvalidArgumentError("output_grads views size must be 1 if not using global sampling.")
    if sep and output_grad.shape[1] != nCam:
        raise _lib.InvalidArgumentError("output_grads views size does not match cameras.")
    L = _lib.lib()
    shape = _lib.i64_array(input.shape)
    wsb = L.gft_sample_grid_transform_bwd_workspace(shape, mm, int(num_mipmaps), gm)
    ws = _workspace(wsb, input.device)
    grad = torch.empty_like(input)
    _lib.check(L.gft_sample_grid_transform_bwd(
        _lib.stream_of(input), _lib.ptr(input), shape, _lib.ptr(output_grad), _np_ptr(M), nM, _np_ptr(V), _np_ptr(P),
        _np_ptr(F), nCam, _lib.i32_array(output_grad.shape[2:5]), fm, bm, mm, int(num_mipmaps), float(mip_bias), cm,
        0.5, int(bool(sep)), gm, _lib.ptr(ws), wsb, _lib.ptr(grad)))
    return grad


def _gather_grad_impl(output_grad, cams, in_dhw, fm, bm, mm, num_mipmaps, mip_bias, cm):
    M, nM, V, P, F, nCam = cams
    output_grad = output_grad.contiguous()
    N, Vc = output_grad.shape[0], output_grad.shape[1]
    C = output_grad.shape[5]
    L = _lib.lib()
    gshape = _lib.i64_array(output_grad.shape)
    wsb = L.gft_sample_grid_transform_gather_grad_workspace(gshape, mm, int(num_mipmaps))
    ws = _workspace(wsb, output_grad.device)
    grad = torch.empty((N,) + tuple(in_dhw) + (C,), device=output_grad.device, dtype=torch.float32)
    _lib.check(L.gft_sample_grid_transform_gather_grad(
        _lib.stream_of(output_grad), _lib.ptr(output_grad), gshape, _lib.i32_array(in_dhw), _np_ptr(M), nM,
        _np_ptr(V), _np_ptr(P), _np_ptr(F), nCam, fm, bm, mm, int(num_mipmaps), float(mip_bias), cm, _lib.ptr(ws),
        wsb, _lib.ptr(grad)))
    return grad


class _SampleTransformFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, input, cams, out_dhw, attrs):
        fm, bm, mm, nmip, bias, cm, sep, gm, gather_bias_add = attrs
        out = _sample_transform_fwd_impl(input, cams, out_dhw, fm, bm, mm, nmip, bias, cm, sep)
        ctx.save_for_backward(input)
        ctx.cams, ctx.attrs = cams, attrs
        return out

    @staticmethod
    def backward(ctx, grad_out):
        (input,) = ctx.saved_tensors
        fm, bm, mm, nmip, bias, cm, sep, gm, gather_bias_add = ctx.attrs
        if gm == "GATHER":
            g = _gather_grad_impl(grad_out, ctx.cams, tuple(input.shape[1:4]), fm, bm, mm, nmip,
                                  bias + gather_bias_add, cm)
        else:
            g = _sample_transform_bwd_impl(input, grad_out, ctx.cams, fm, bm, mm, nmip, bias, cm, sep, gm)
        return g, None, None, None


def _lut_check(input, lookup, cell_center_offset, separate_camera_batch, normalized_coords):
    lookup = _grid(lookup, "lookup")
    if lookup.shape[4] != 4:
        raise _lib.InvalidArgumentError("Invalid lut shape (NDHWC) with C=4: %s" % list(lookup.shape))
    if normalized_coords and float(cell_center_offset) != 0.5:
        raise _lib.InvalidArgumentError("Invalid cell center offset must be 0.5 when using normalized coordinates.")
    if not separate_camera_batch and lookup.shape[0] != input.shape[0]:
        raise _lib.InvalidArgumentError("camera batch must match data batch when not using global sampling.")
    return lookup


def _sample_lut_fwd_impl(input, lookup, fm, bm, mm, num_mipmaps, mip_bias, offset, sep, rel, norm):
    N, C = input.shape[0], input.shape[4]
    nv = lookup.shape[0] if sep else 1
    out = torch.empty((N, nv) + tuple(lookup.shape[1:4]) + (C,), device=input.device, dtype=torch.float32)
    L = _lib.lib()
    shape = _lib.i64_array(input.shape)
    wsb = L.gft_sample_grid_lut_workspace(shape, mm, int(num_mipmaps))
    ws = _workspace(wsb, input.device)
    _lib.check(L.gft_sample_grid_lut_fwd(
        _lib.stream_of(input), _lib.ptr(input), shape, _lib.ptr(lookup), _lib.i64_array(lookup.shape), fm, bm, mm,
        int(num_mipmaps), float(mip_bias), float(offset), int(bool(sep)), int(bool(rel)), int(bool(norm)),
        _lib.ptr(ws), wsb, _lib.ptr(out)))
    return out


def _sample_lut_bwd_impl(input, lookup, output_grad, fm, bm, mm, offset, sep, rel, norm, gm):
    output_grad = output_grad.contiguous()
    N = input.shape[0]
    if output_grad.dim() != 6 or output_grad.shape[0] != N:
        raise _lib.InvalidArgumentError("Invalid output_grads shape (NVDHWC): %s" % list(output_grad.shape))
    if tuple(output_grad.shape[2:5]) != tuple(lookup.shape[1:4]):
        raise _lib.InvalidArgumentError("Invalid lut shape (VDHWC): %s, DHW must macht output shape: %s"
                                        % (list(lookup.shape), list(output_grad.shape)))
    if not sep and output_grad.shape[1] != 1:
        raise _lib.InvalidArgumentError("output_grads views size must be 1 if not using global sampling.")
    if sep and output_grad.shape[1] != lookup.shape[0]:
        raise _lib.InvalidArgumentError("output_grads views size does not match cameras.")
    if mm != MIP["NONE"]:
        raise _lib.InvalidArgumentError("lut gradient supports mipmapping NONE only")
    L = _lib.lib()
    shape = _lib.i64_array(input.shape)
    wsb = L.gft_sample_grid_lut_bwd_workspace(shape, gm)
    ws = _workspace(wsb, input.device)
    gin = torch.empty_like(input)
    glut = torch.empty_like(lookup)
    _lib.check(L.gft_sample_grid_lut_bwd(
        _lib.stream_of(input), _lib.ptr(input), shape, _lib.ptr(lookup), _lib.i64_array(lookup.shape),
        _lib.ptr(output_grad), fm, bm, mm, float(offset), int(bool(sep)), int(bool(rel)), int(bool(norm)), gm,
        _lib.ptr(ws), wsb, _lib.ptr(gin), _lib.ptr(glut)))
    return gin, glut


class _SampleLutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, input, lookup, attrs):
        fm, bm, mm, nmip, bias, offset, sep, rel, norm, gm = attrs
        out = _sample_lut_fwd_impl(input, lookup, fm, bm, mm, nmip, bias, offset, sep, rel, norm)
        ctx.save_for_backward(input, lookup)
        ctx.attrs = attrs
        return out

    @staticmethod
    def backward(ctx, grad_out):
        input, lookup = ctx.saved_tensors
        fm, bm, mm, nmip, bias, offset, sep, rel, norm, gm = ctx.attrs
        gin, glut = _sample_lut_bwd_impl(input, lookup, grad_out, fm, bm, mm, offset, sep, rel, norm, gm)
        return gin, glut, None


class _SamplingOps(object):
    """``sampling_ops`` of ops_loader.py:11-51 (resample_grid.cc:23-154)."""

    @staticmethod
    def sample_grid_transform(input, matrix_m, matrix_v, matrix_p, frustum_params, output_shape,
                              interpolation="LINEAR", boundary="BORDER", mipmapping="NONE", num_mipmaps=0,
                              mip_bias=0.0, coordinate_mode="TRANSFORM_LINDEPTH", cell_center_offset=0.5,
                              separate_camera_batch=True, grad_mode="REFERENCE", gradient_mip_bias_add=0.0):
        """SampleGridTransform: NDHWC -> [N,V,Do,Ho,Wo,C] ([N,1,...] without
        separate_camera_batch). grad_mode "REFERENCE" registers the op's own
        gradient (ops_loader.py:14-31), "EXACT" the true adjoint, "GATHER"
        the renderer's view-summed reverse resample (renderer.py:246-272) with
        mip_bias + gradient_mip_bias_add."""
        input = _grid(input)
        fm, bm, cm, mm = _check_transform_attrs(interpolation, boundary, coordinate_mode, cell_center_offset,
                                                mipmapping, num_mipmaps)
        cams = _cams(matrix_m, matrix_v, matrix_p, frustum_params)
        gm = "GATHER" if str(grad_mode).upper() == "GATHER" else _enum(GRAD, grad_mode, "grad_mode")
        attrs = (fm, bm, mm, int(num_mipmaps), float(mip_bias), cm, bool(separate_camera_batch), gm,
                 float(gradient_mip_bias_add))
        return _SampleTransformFn.apply(input, cams, _shape3(output_shape), attrs)

    @staticmethod
    def sample_grid_transform_grad(input, output_grad, matrix_m, matrix_v, matrix_p, frustum_params,
                                   interpolation="LINEAR", boundary="BORDER", mipmapping="NONE", num_mipmaps=0,
                                   mip_bias=0.0, coordinate_mode="TRANSFORM_LINDEPTH", cell_center_offset=0.5,
                                   separate_camera_batch=True, grad_mode="REFERENCE"):
        """SampleGridTransformGrad (resample_grid.cc:64-84)."""
        input = _grid(input)
        fm, bm, cm, mm = _check_transform_attrs(interpolation, boundary, coordinate_mode, cell_center_offset,
                                                mipmapping, num_mipmaps)
        gm = _enum(GRAD, grad_mode, "grad_mode")
        if gm == GRAD["REFERENCE"] and mm != MIP["NONE"]:
            raise _lib.InvalidArgumentError("the REFERENCE gradient supports mipmapping NONE only")
        return _sample_transform_bwd_impl(input, output_grad.detach(), _cams(matrix_m, matrix_v, matrix_p,
                                                                               frustum_params),
                                          fm, bm, mm, num_mipmaps, mip_bias, cm, bool(separate_camera_batch), gm)

    @staticmethod
    def sample_grid_transform_gather_grad(output_grad, matrix_m, matrix_v, matrix_p, frustum_params, input_shape,
                                          interpolation="LINEAR", boundary="BORDER", mipmapping="NONE",
                                          num_mipmaps=0, mip_bias=0.0, coordinate_mode="TRANSFORM_LINDEPTH"):
        """The renderer's gather gradient (renderer.py:246-272) in one fused
        pass: dx[i] = sum_j resample(dy[i, j]) with the reverse coordinate
        mode of ``coordinate_mode``; ``mip_bias`` includes the gradient add."""
        fm, bm, cm, mm = _check_transform_attrs(interpolation, boundary, coordinate_mode, 0.5, mipmapping,
                                                num_mipmaps)
        og = output_grad.detach().contiguous()
        if og.dim() != 6:
            raise _lib.InvalidArgumentError("Invalid output_grads shape (NVDHWC): %s" % list(og.shape))
        return _gather_grad_impl(og, _cams(matrix_m, matrix_v, matrix_p, frustum_params), _shape3(input_shape), fm,
                                 bm, mm, num_mipmaps, mip_bias, cm)

    @staticmethod
    def sample_grid_lut(input, lookup, interpolation="LINEAR", boundary="BORDER", mipmapping="NONE", num_mipmaps=0,
                        mip_bias=0.0, coordinate_mode="LOOKUP", cell_center_offset=0.0, separate_camera_batch=True,
                        relative_coords=True, normalized_coords=False, grad_mode="REFERENCE"):
        """SampleGridLut (resample_grid.cc:86-109): NDHWC sampled at a lookup
        table [V,Do,Ho,Wo,4] -> [N,V,Do,Ho,Wo,C]."""
        input = _grid(input)
        if str(coordinate_mode).upper() != "LOOKUP":
            raise _lib.InvalidArgumentError("Invalid coorindate_mode")
        lookup = _lut_check(input, lookup, cell_center_offset, separate_camera_batch, normalized_coords)
        mm = _enum(MIP, mipmapping, "mipmapping")
        if mm != MIP["NONE"] and int(num_mipmaps) <= 0:
            raise _lib.InvalidArgumentError("when using mipmaps num_mipmaps must be larger than 0.")
        attrs = (_enum(FILTER, interpolation, "interpolation"), _enum(BOUNDARY, boundary, "boundary"), mm,
                 int(num_mipmaps), float(mip_bias), float(cell_center_offset), bool(separate_camera_batch),
                 bool(relative_coords), bool(normalized_coords), _enum(GRAD, grad_mode, "grad_mode"))
        return _SampleLutFn.apply(input, lookup, attrs)

    @staticmethod
    def sample_grid_lut_grad(input, output_grad, lookup, interpolation="LINEAR", boundary="BORDER",
                             mipmapping="NONE", num_mipmaps=0, mip_bias=0.0, coordinate_mode="LOOKUP",
                             cell_center_offset=0.0, separate_camera_batch=True, relative_coords=True,
                             normalized_coords=False, grad_mode="REFERENCE"):
        """SampleGridLutGrad (resample_grid.cc:113-134): (input_grad, lookup_grad)."""
        input = _grid(input)
        lookup = _lut_check(input, lookup, cell_center_offset, separate_camera_batch, normalized_coords)
        return _sample_lut_bwd_impl(input, lookup, output_grad.detach(), _enum(FILTER, interpolation, "interpolation"),
                                    _enum(BOUNDARY, boundary, "boundary"), _enum(MIP, mipmapping, "mipmapping"),
                                    float(cell_center_offset), bool(separate_camera_batch), bool(relative_coords),
                                    bool(normalized_coords), _enum(GRAD, grad_mode, "grad_mode"))

    @staticmethod
    def lod_transform(input_shape, matrix_mv, matrix_p, frustum_params, output_shape, inverse_transform=False,
                      device="cuda"):
        """LodTransform (resample_grid.cc:136-154): [V,Do,Ho,Wo,4] sampling
        positions (x,y,z) and LoD per output cell."""
        MV, nMV = _matrices(matrix_mv, "matrix_mv")
        P, nP = _matrices(matrix_p, "matrix_p")
        F, nF = _frustum(frustum_params)
        if not (nMV == nP == nF):
            raise _lib.InvalidArgumentError("camera batch size mismatch")
        in_dhw = _shape3(input_shape)
        out_dhw = _shape3(output_shape)
        out = torch.empty((nMV,) + tuple(out_dhw) + (4,), device=device, dtype=torch.float32)
        L = _lib.lib()
        _lib.check(L.gft_lod_transform(_lib.stream_of(out), _lib.i32_array(in_dhw), _np_ptr(MV), _np_ptr(P),
                                       _np_ptr(F), nMV, _lib.i32_array(out_dhw), int(bool(inverse_transform)),
                                       _lib.ptr(out)))
        return out

    @staticmethod
    def build_mips(input, num_mipmaps):
        """Mip pyramid (kGenerateMip3D): list of the levels 1..L as
        [N,D>>m,H>>m,W>>m,C] tensors (views into one atlas)."""
        input = _grid(input)
        L = _lib.lib()
        shape = _lib.i64_array(input.shape)
        nb = L.gft_mip_atlas_bytes(shape, int(num_mipmaps))
        atlas = torch.empty((max(int(nb) // 4, 1),), device=input.device, dtype=torch.float32)
        _lib.check(L.gft_build_mips(_lib.stream_of(input), _lib.ptr(input), shape, int(num_mipmaps), _lib.ptr(atlas),
                                    nb))
        N, D, H, W, C = input.shape
        levels, off, dims = [], 0, []
        m = 1
        while m <= int(num_mipmaps) and (D >> m) > 0 and (H >> m) > 0 and (W >> m) > 0 and m <= 15:
            dims.append((D >> m, H >> m, W >> m))
            m += 1
        per = sum(d * h * w * C for d, h, w in dims)
        stride = (per + 63) // 64 * 64
        for d, h, w in dims:
            n = d * h * w * C
            lv = torch.as_strided(atlas, (N, d, h, w, C), (stride, h * w * C, w * C, C, 1), off)
            levels.append(lv)
            off += n
        return levels


sampling_ops = _SamplingOps()


# ------------------------------------------------------------------- blending
def _blend_check(input, blending_mode):
    input = _grid(input)
    if input.shape[4] > 4:
        raise _lib.InvalidArgumentError("Invalid input shape %s" % list(input.shape))
    return input, _enum(BLEND, blending_mode, "blending mode")


def _reduce_blend_fwd_impl(input, bl, keep_dims):
    N, D, H, W, C = input.shape
    out = torch.empty(input.shape if keep_dims else (N, H, W, C), device=input.device, dtype=torch.float32)
    L = _lib.lib()
    _lib.check(L.gft_reduce_grid_blend_fwd(_lib.stream_of(input), _lib.ptr(input), _lib.i64_array(input.shape), bl,
                                           int(bool(keep_dims)), _lib.ptr(out)))
    return out


def _reduce_blend_bwd_impl(output_grad, output, input, bl, keep_dims, gm):
    output_grad = output_grad.contiguous()
    output = output.contiguous()
    if tuple(output_grad.shape) != tuple(output.shape):
        raise _lib.InvalidArgumentError("Mismatch between output gradients and output shape: %s %s"
                                        % (list(output_grad.shape), list(output.shape)))
    N, D, H, W, C = input.shape
    expect = tuple(input.shape) if keep_dims else (N, H, W, C)
    if tuple(output.shape) != expect:
        raise _lib.InvalidArgumentError("output gradients shape does not match input shape %s" % list(input.shape))
    grad = torch.empty_like(input)
    L = _lib.lib()
    _lib.check(L.gft_reduce_grid_blend_bwd(_lib.stream_of(input), _lib.ptr(input), _lib.i64_array(input.shape),
                                           _lib.ptr(output), _lib.ptr(output_grad), bl, int(bool(keep_dims)), gm,
                                           _lib.ptr(grad)))
    return grad


class _ReduceBlendFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, input, bl, keep_dims, gm):
        out = _reduce_blend_fwd_impl(input, bl, keep_dims)
        ctx.save_for_backward(input, out)
        ctx.attrs = (bl, keep_dims, gm)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        input, out = ctx.saved_tensors
        bl, keep_dims, gm = ctx.attrs
        return _reduce_blend_bwd_impl(grad_out, out, input, bl, keep_dims, gm), None, None, None


class _BlendingOps(object):
    """``blending_ops`` of ops_loader.py:53-60 (reduce_blend.cc:23-48)."""

    @staticmethod
    def reduce_grid_blend(input, blending_mode="BEER_LAMBERT", keep_dims=False, grad_mode="REFERENCE"):
        """ReduceGridBlend: blend NDHWC along D -> [N,H,W,C] (every prefix
        [N,D,H,W,C] with keep_dims)."""
        input, bl = _blend_check(input, blending_mode)
        return _ReduceBlendFn.apply(input, bl, bool(keep_dims), _enum(GRAD, grad_mode, "grad_mode"))

    @staticmethod
    def reduce_grid_blend_grad(output_grad, output, input, blending_mode="BEER_LAMBERT", keep_dims=False,
                               grad_mode="REFERENCE"):
        """ReduceGridBlendGrad (input order of reduce_blend.cc:37-47)."""
        input, bl = _blend_check(input, blending_mode)
        return _reduce_blend_bwd_impl(output_grad.detach(), output.detach(), input, bl, bool(keep_dims),
                                      _enum(GRAD, grad_mode, "grad_mode"))


blending_ops = _BlendingOps()


# ------------------------------------------------------------------- lighting
def _shadow_fwd_impl(density, cams, shadow_dhw, step, fm, bm, mm, num_mipmaps, mip_bias, bl, mask, save):
    M, nM, V, P, F, nCam = cams
    if nM != 1 or nCam != 1:
        raise _lib.InvalidArgumentError("the shadow map takes one model matrix and one light camera")
    L = _lib.lib()
    shape = _lib.i64_array(density.shape)
    wsb = L.gft_shadow_transmission_fwd_workspace(shape, mm, int(num_mipmaps))
    ws = _workspace(wsb, density.device)
    out = torch.empty((1,) + tuple(shadow_dhw) + (1,), device=density.device, dtype=torch.float32)
    saved = torch.empty(tuple(shadow_dhw), device=density.device, dtype=torch.float32) if save else None
    _lib.check(L.gft_shadow_transmission_fwd(
        _lib.stream_of(density), _lib.ptr(density), shape, _np_ptr(M), _np_ptr(V), _np_ptr(P), _np_ptr(F),
        _lib.i32_array(shadow_dhw), float(step), fm, bm, mm, int(num_mipmaps), float(mip_bias), bl,
        _lib.ptr(mask) if mask is not None else None, _lib.ptr(ws), wsb,
        _lib.ptr(saved) if saved is not None else None, _lib.ptr(out)))
    return out, saved


def _shadow_bwd_impl(saved, grad_t, shadow_dhw, bm, bl, mask):
    grad_t = grad_t.contiguous()
    if tuple(grad_t.shape) != (1,) + tuple(shadow_dhw) + (1,):
        raise _lib.InvalidArgumentError("transmission gradient shape %s does not match the shadow map %s"
                                        % (list(grad_t.shape), list(shadow_dhw)))
    gs = torch.empty((1, 1) + tuple(shadow_dhw) + (1,), device=saved.device, dtype=torch.float32)
    L = _lib.lib()
    _lib.check(L.gft_shadow_transmission_bwd(
        _lib.stream_of(saved), _lib.ptr(saved), _lib.i32_array(shadow_dhw), bm, bl,
        _lib.ptr(mask) if mask is not None else None, _lib.ptr(grad_t), _lib.ptr(gs)))
    return gs


class _ShadowFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, density, cams, shadow_dhw, attrs, mask):
        step, fm, bm, mm, nmip, bias, bl, gm, gather_bias_add = attrs
        out, saved = _shadow_fwd_impl(density, cams, shadow_dhw, step, fm, bm, mm, nmip, bias, bl, mask,
                                      save=ctx.needs_input_grad[0])
        if saved is not None:
            ctx.save_for_backward(density, saved)
        ctx.cams, ctx.shadow_dhw, ctx.attrs, ctx.mask = cams, shadow_dhw, attrs, mask
        return out

    @staticmethod
    def backward(ctx, grad_out):
        density, saved = ctx.saved_tensors
        step, fm, bm, mm, nmip, bias, bl, gm, gather_bias_add = ctx.attrs
        gs = _shadow_bwd_impl(saved, grad_out, ctx.shadow_dhw, bm, bl, ctx.mask)
        cm = COORD["TRANSFORM_LINDEPTH"]
        if gm == "GATHER":
            g = _gather_grad_impl(gs, ctx.cams, tuple(density.shape[1:4]), fm, bm, mm, nmip, bias + gather_bias_add,
                                  cm)
        else:
            g = _sample_transform_bwd_impl(density, gs, ctx.cams, fm, bm, mm, nmip, bias, cm, True, gm)
        # the op sequence scales the grid by the step before sampling: d/d density = step * d/d(scaled grid)
        g = g * torch.tensor([step], dtype=torch.float32, device=g.device)
        return g, None, None, None, None


class _LightingOps(object):
    """The fused spot-light shadow transmittance (lighting.py:106-141 up to
    the reverse resample) as one op."""

    @staticmethod
    def shadow_transmission(density, matrix_m, matrix_v, matrix_p, frustum_params, shadow_shape, depth_step,
                            interpolation="LINEAR", boundary="BORDER", mipmapping="NONE", num_mipmaps=0,
                            mip_bias=0.0, blending_mode="BEER_LAMBERT", cone_mask=None, grad_mode="REFERENCE",
                            gradient_mip_bias_add=0.0):
        """Transmittance [1,Ds,Hs,Ws,1] of a density [1,D,H,W,1] in the light
        frustum: sample_grid_transform (TRANSFORM_LINDEPTH) of density *
        depth_step, reduce_grid_blend keep_dims, the one-cell depth shift
        padded by 0 (BORDER) or the first cell (CLAMP), exp(-D) for
        BEER_LAMBERT or 1 - clip(D, 0, 1) for ALPHA, times cone_mask
        ([Hs,Ws] or any shape with Hs*Ws elements; None for no mask).
        grad_mode selects the sampling gradient as in sample_grid_transform
        ("REFERENCE", "EXACT" or "GATHER"); mipmapping, num_mipmaps and
        mip_bias as there (the REFERENCE gradient takes no mips)."""
        density = _grid(density, "density")
        fm, bm, cm, mm = _check_transform_attrs(interpolation, boundary, "TRANSFORM_LINDEPTH", 0.5, mipmapping,
                                                num_mipmaps)
        bl = _enum(BLEND, blending_mode, "blending mode")
        if bl not in (BLEND["BEER_LAMBERT"], BLEND["ALPHA"]):
            raise _lib.InvalidArgumentError("Unknown blend_mode '%s'" % blending_mode)
        dhw = _shape3(shadow_shape)
        if cone_mask is not None:
            cone_mask = cone_mask.detach().to(device=density.device, dtype=torch.float32).contiguous()
            if cone_mask.numel() != dhw[1] * dhw[2]:
                raise _lib.InvalidArgumentError("cone mask of %d values for a %dx%d shadow map"
                                                % (cone_mask.numel(), dhw[1], dhw[2]))
        gm = "GATHER" if str(grad_mode).upper() == "GATHER" else _enum(GRAD, grad_mode, "grad_mode")
        attrs = (float(depth_step), fm, bm, mm, int(num_mipmaps), float(mip_bias), bl, gm,
                 float(gradient_mip_bias_add))
        return _ShadowFn.apply(density, _cams(matrix_m, matrix_v, matrix_p, frustum_params), tuple(dhw), attrs,
                               cone_mask)


lighting_ops = _LightingOps()
sample_grid_transform = sampling_ops.sample_grid_transform
sample_grid_lut = sampling_ops.sample_grid_lut
reduce_grid_blend = blending_ops.reduce_grid_blend
