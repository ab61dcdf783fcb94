"""gft_advect_bwd_noalloc (the MacCormack adjoint with its accumulator in the
library's own device array) against gft_advect_bwd with a caller workspace,
both called through the C ABI.

Two calls in a row on different grids, the second smaller than the first, so
a second call that read the first call's accumulator contents would show.
Both entry points run the same kernels; the adjoint accumulates with atomics,
so the two results agree to rounding: normalised max error <= 1e-6 for the
density and velocity gradients (the FP64 oracle bar is 1e-3). Also checked:
the FP64 oracle at 1e-3, and the size limit rejected before any device work.
"""
import ctypes

import numpy as np
import pytest
import torch

from gft import _lib
from gft import synthetic as syn
from helpers import norm_err
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

MAX_VOXELS = 1 << 24  # GFT_ADVECT_NOALLOC_MAX_VOXELS


def _bwd_both(grid, seed, order=2, clamp=1):
    L = _lib.lib()
    d = syn.gaussian_blobs(grid, 6, seed=seed)
    v = syn.curl_noise_velocity(grid, seed=seed + 1, max_speed=1.5, periods=(8, 4, 2))
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    g = rng.uniform(-1, 1, size=d.shape).astype(np.float32)
    shape = _lib.i64_array((1,) + tuple(grid) + (1,))
    dd, vv, gg = (torch.from_numpy(x).cuda() for x in (d, v, g))
    s = _lib.stream_of(dd)
    nsaved = L.gft_advect_saved_bytes(shape)
    saved = torch.empty((nsaved,), dtype=torch.uint8, device="cuda")
    out = torch.empty_like(dd)
    _lib.check(L.gft_advect_fwd(s, _lib.ptr(dd), shape, _lib.ptr(vv), 1.0, order, clamp, _lib.ptr(out),
                                _lib.ptr(saved), nsaved))
    nws = L.gft_advect_bwd_workspace(shape)
    ws = torch.empty((nws,), dtype=torch.uint8, device="cuda")
    gd_ws, gv_ws = torch.empty_like(dd), torch.empty_like(vv)
    _lib.check(L.gft_advect_bwd(s, _lib.ptr(dd), shape, _lib.ptr(vv), 1.0, order, clamp, _lib.ptr(saved),
                                _lib.ptr(gg), _lib.ptr(gd_ws), _lib.ptr(gv_ws), _lib.ptr(ws), nws))
    gd_na, gv_na = torch.full_like(dd, 7.0), torch.full_like(vv, 7.0)
    _lib.check(L.gft_advect_bwd_noalloc(s, _lib.ptr(dd), shape, _lib.ptr(vv), 1.0, order, clamp, _lib.ptr(saved),
                                        _lib.ptr(gg), _lib.ptr(gd_na), _lib.ptr(gv_na)))
    torch.cuda.synchronize()
    return (d, v, g), (gd_ws.cpu().numpy(), gv_ws.cpu().numpy()), (gd_na.cpu().numpy(), gv_na.cpu().numpy())


def test_noalloc_equals_workspace_entry_point_twice(cuda):
    for grid, seed in (((24, 40, 32), 3), ((10, 14, 12), 11)):  # second call on a smaller grid
        print("noalloc advect bwd grid", grid, "seed", seed, flush=True)
        (d, v, g), (gd_ws, gv_ws), (gd_na, gv_na) = _bwd_both(grid, seed)
        assert norm_err(gd_na, gd_ws) <= 1e-6, grid
        assert norm_err(gv_na, gv_ws) <= 1e-6, grid
        gd_ref, gv_ref = orc.advect_bwd(d, v, g, 2, "MC_SMOOTH", 1.0, "CLAMP", "EXACT", prec=0)
        assert norm_err(gd_na, gd_ref) < 1e-3 and norm_err(gv_na, gv_ref) < 1e-3


def test_noalloc_semi_lagrangian(cuda):
    print("noalloc advect bwd order 1", flush=True)
    _, (gd_ws, gv_ws), (gd_na, gv_na) = _bwd_both((12, 16, 14), 5, order=1, clamp=0)
    assert norm_err(gd_na, gd_ws) <= 1e-6 and norm_err(gv_na, gv_ws) <= 1e-6


def test_noalloc_rejects_grids_beyond_its_scratch(cuda):
    L = _lib.lib()
    shape = _lib.i64_array((1, 256, 256, 257, 1))  # 16.8 M voxels > 2^24
    assert 256 * 256 * 257 > MAX_VOXELS
    p = ctypes.c_void_p(0)
    rc = L.gft_advect_bwd_noalloc(p, p, shape, p, 1.0, 2, 1, p, p, p, p)
    assert rc == _lib.GFT_ERR_INVALID_ARGUMENT
    assert b"no-allocation scratch" in L.gft_last_error()
