"""Loader of libgft.so (the HIP kernels behind a C ABI, include/gft.h).

The library is loaded after torch so that both share torch's HIP runtime
(both link libamdhip64.so.7). There is no fallback: a missing or unloadable
library raises, so a GPU run can never silently compute elsewhere.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgft.so")
_LIB = None

GFT_OK = 0
GFT_ERR_INVALID_ARGUMENT = 1
GFT_ERR_UNIMPLEMENTED = 2
GFT_ERR_RUNTIME = 3
GFT_ERR_WORKSPACE = 4


class GftError(RuntimeError):
    """HIP runtime or workspace failure reported by libgft."""


class InvalidArgumentError(GftError, ValueError):
    """Counterpart of TF's InvalidArgument raised by the reference ops."""


class UnimplementedError(GftError, NotImplementedError):
    """Counterpart of TF's Unimplemented raised by the reference ops."""


_i32 = ctypes.c_int32
_i64p = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_f32 = ctypes.c_float
_f32p = ctypes.POINTER(ctypes.c_float)

# name -> (restype, argtypes)
_SIGNATURES = {
    "gft_last_error": (ctypes.c_char_p, []),
    "gft_abi_version": (ctypes.c_int, []),
    "gft_raymarch_grid_transform_fwd": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _i32, _vp, _vp, _vp, _i32, _i32p, _i32,
                                                       _i32, _i32, _i32, _i32, _vp]),
    "gft_raymarch_grid_transform_bwd_workspace": (_sz, [_i64p, _i32, _i32p, _i32, _i32, _i32]),
    "gft_raymarch_grid_transform_bwd": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _i32, _i32p,
                                                       _i32, _i32, _i32, _i32, _i32, _i32, _vp, _sz, _vp]),
    "gft_advect_saved_bytes": (_sz, [_i64p]),
    "gft_advect_fwd": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _f32, _i32, _i32, _vp, _vp, _sz]),
    "gft_advect_bwd_workspace": (_sz, [_i64p]),
    "gft_advect_bwd": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _f32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _sz]),
    "gft_advect_bwd_noalloc": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _f32, _i32, _i32, _vp, _vp, _vp, _vp]),
    "gft_advect_fwd_bf16": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _f32, _i32, _i32, _vp, _vp, _sz]),
    "gft_advect_bwd_bf16": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _f32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _sz]),
    "gft_render_density_fwd_workspace": (_sz, [_i64p, _i32]),
    "gft_render_density_saved_bytes": (_sz, [_i64p, _i32, _i32p, _i32]),
    "gft_render_density_fwd": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _i32, _vp, _vp, _vp, _i32, _i32p, _f32p, _f32,
                                              _i32, _vp, _sz, _vp, _sz, _vp]),
    "gft_render_density_fwd_bf16": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _i32, _vp, _vp, _vp, _i32, _i32p, _f32p,
                                                   _f32, _i32, _vp, _sz, _vp, _sz, _vp]),
    "gft_render_density_bwd_workspace": (_sz, [_i64p, _i32, _i32p, _i32]),
    "gft_render_density_bwd_l2": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _i32, _vp, _vp, _vp, _i32, _i32p, _f32p, _f32,
                                                 _i32, _vp, _vp, _f32, _vp, _sz, _vp, _sz, _vp, _vp]),
    "gft_render_density_bwd_loss": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _i32, _vp, _vp, _vp, _i32, _i32p, _f32p,
                                                   _f32, _i32, _vp, _vp, _vp, _i32, _f32, _vp, _sz, _vp, _sz, _vp,
                                                   _vp]),
    "gft_render_density_bwd": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _i32, _vp, _vp, _vp, _i32, _i32p, _f32p, _f32,
                                              _i32, _vp, _vp, _vp, _sz, _vp, _sz, _vp]),
    "gft_render_density_lit_fwd_workspace": (_sz, [_i64p, _i32]),
    "gft_render_density_lit_saved_bytes": (_sz, [_i64p, _i32, _i32p, _i32]),
    "gft_render_density_lit_fwd": (ctypes.c_int, [_vp, _vp, _vp, _i64p, _vp, _i32, _vp, _vp, _vp, _i32, _i32p, _f32p,
                                                  _f32, _f32, _i32, _vp, _sz, _vp, _sz, _vp]),
    "gft_render_density_lit_bwd_loss": (ctypes.c_int, [_vp, _vp, _vp, _i64p, _vp, _i32, _vp, _vp, _vp, _i32, _i32p,
                                                       _f32p, _f32, _f32, _i32, _vp, _vp, _vp, _i32, _f32, _vp, _sz,
                                                       _vp, _vp, _vp]),
    "gft_render_density_lit_bwd": (ctypes.c_int, [_vp, _vp, _vp, _i64p, _vp, _i32, _vp, _vp, _vp, _i32, _i32p, _f32p,
                                                  _f32, _f32, _i32, _vp, _vp, _vp, _sz, _vp, _vp]),
    "gft_mip_atlas_bytes": (_sz, [_i64p, _i32]),
    "gft_build_mips": (ctypes.c_int, [_vp, _vp, _i64p, _i32, _vp, _sz]),
    "gft_sample_grid_transform_workspace": (_sz, [_i64p, _i32, _i32]),
    "gft_sample_grid_transform_fwd": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _i32, _vp, _vp, _vp, _i32, _i32p, _i32, _i32,
                                                     _i32, _i32, _f32, _i32, _f32, _i32, _vp, _sz, _vp]),
    "gft_sample_grid_transform_bwd_workspace": (_sz, [_i64p, _i32, _i32, _i32]),
    "gft_sample_grid_transform_bwd": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _vp, _i32, _vp, _vp, _vp, _i32, _i32p, _i32,
                                                     _i32, _i32, _i32, _f32, _i32, _f32, _i32, _i32, _vp, _sz, _vp]),
    "gft_sample_grid_transform_gather_grad_workspace": (_sz, [_i64p, _i32, _i32]),
    "gft_sample_grid_transform_gather_grad": (ctypes.c_int, [_vp, _vp, _i64p, _i32p, _vp, _i32, _vp, _vp, _vp, _i32,
                                                             _i32, _i32, _i32, _i32, _f32, _i32, _vp, _sz, _vp]),
    "gft_sample_grid_lut_workspace": (_sz, [_i64p, _i32, _i32]),
    "gft_sample_grid_lut_fwd": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _i64p, _i32, _i32, _i32, _i32, _f32, _f32, _i32,
                                               _i32, _i32, _vp, _sz, _vp]),
    "gft_sample_grid_lut_bwd_workspace": (_sz, [_i64p, _i32]),
    "gft_sample_grid_lut_bwd": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _i64p, _vp, _i32, _i32, _i32, _f32, _i32, _i32,
                                               _i32, _i32, _vp, _sz, _vp, _vp]),
    "gft_lod_transform": (ctypes.c_int, [_vp, _i32p, _vp, _vp, _vp, _i32, _i32p, _i32, _vp]),
    "gft_reduce_grid_blend_fwd": (ctypes.c_int, [_vp, _vp, _i64p, _i32, _i32, _vp]),
    "gft_reduce_grid_blend_bwd": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _vp, _i32, _i32, _i32, _vp]),
    "gft_shadow_transmission_saved_bytes": (_sz, [_i32p]),
    "gft_shadow_transmission_fwd_workspace": (_sz, [_i64p, _i32, _i32]),
    "gft_shadow_transmission_fwd": (ctypes.c_int, [_vp, _vp, _i64p, _vp, _vp, _vp, _vp, _i32p, _f32, _i32, _i32, _i32,
                                                   _i32, _f32, _i32, _vp, _vp, _sz, _vp, _vp]),
    "gft_shadow_transmission_bwd": (ctypes.c_int, [_vp, _vp, _i32p, _i32, _i32, _vp, _vp, _vp]),
    "gft_divergence_loss_workspace": (_sz, [_i32p]),
    "gft_divergence_loss": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32p, _f32p, _f32, _i32, _vp, _sz, _vp, _vp, _vp,
                                           _vp]),
    "gft_adam_step": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                     ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_int64), _f32, _f32, _f32, _f32,
                                     ctypes.c_int64]),
    "gft_adam_step_device": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                            ctypes.POINTER(_vp), ctypes.POINTER(ctypes.c_int64), _f32, _f32, _f32, _f32,
                                            _vp]),
    "gft_batch_render_workspace": (_sz, [_i32, _vp, _i32]),
    "gft_batch_render_prepare": (ctypes.c_int, [_vp, _i32, _vp, _i32, _vp, _sz]),
    "gft_batch_render_run": (ctypes.c_int, [_vp, _i32, _vp, _i32, _vp, _sz]),
    "gft_timer_start": (ctypes.c_int, [ctypes.c_char_p]),
    "gft_timer_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)]),
    "gft_timer_names": (ctypes.c_int, [ctypes.c_char_p, _sz]),
    "gft_timer_stop": (None, []),
}


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgft.so is not built (%s); run __graft_entry__.build() or "
                               "python global-flow-transport_amd/build.py" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                raise RuntimeError("libgft.so lacks symbol %s" % name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    return list(_SIGNATURES.keys())


def check(rc):
    if rc == GFT_OK:
        return
    msg = lib().gft_last_error().decode("utf-8", "replace")
    if rc == GFT_ERR_INVALID_ARGUMENT:
        raise InvalidArgumentError(msg)
    if rc == GFT_ERR_UNIMPLEMENTED:
        raise UnimplementedError(msg)
    raise GftError("libgft error %d: %s" % (rc, msg))


def i64_array(values):
    vals = [int(v) for v in values]
    return (ctypes.c_int64 * len(vals))(*vals)


def i32_array(values):
    vals = [int(v) for v in values]
    return (ctypes.c_int32 * len(vals))(*vals)


def stream_of(tensor):
    return ctypes.c_void_p(torch.cuda.current_stream(tensor.device).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class KernelTimers(object):
    """HIP-event timing of the library's named kernel launches on this thread
    (gft_timer_*): ``with KernelTimers("*") as t: ...`` then ``t.read()``
    -> {name: (launches, total_ms)}."""

    def __init__(self, kernels="*"):
        self.kernels = kernels
        self.result = None

    def __enter__(self):
        check(lib().gft_timer_start(self.kernels.encode()))
        return self

    def __exit__(self, *exc):
        try:
            if exc[0] is None:
                self.result = self.read()
        finally:
            lib().gft_timer_stop()
        return False

    def read(self):
        L = lib()
        n = L.gft_timer_names(None, 0)
        buf = ctypes.create_string_buffer(n + 1)
        L.gft_timer_names(buf, n + 1)
        out = {}
        for name in [x for x in buf.value.decode().split(",") if x]:
            cnt, tot = ctypes.c_int32(0), ctypes.c_double(0.0)
            check(L.gft_timer_read(name.encode(), ctypes.byref(cnt), ctypes.byref(tot)))
            out[name] = (int(cnt.value), float(tot.value))
        return out
