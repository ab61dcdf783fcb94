"""Host-side checks that need no GPU: the C ABI library exports every entry
point declared in include/gft.h, the ctypes binding covers them, workspace
queries, argument validation (which returns before any device work), the op
layer's refusal of host tensors, camera/matrix conventions and the workload
geometry."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gft.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gft_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from gft import _lib
    return _lib.lib()


def test_every_declared_symbol_is_exported_and_bound(lib):
    from gft import _lib
    names = _declared()
    assert len(names) >= 28
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.exported_symbols()) == names


def test_abi_version_and_error_slot(lib):
    assert lib.gft_abi_version() == 4
    assert isinstance(lib.gft_last_error(), bytes)


def _i64(*v):
    return (ctypes.c_int64 * len(v))(*v)


def _i32(*v):
    return (ctypes.c_int32 * len(v))(*v)


def test_workspace_queries(lib):
    # mip atlas: per batch levels 1..L of (D>>m, H>>m, W>>m, C), 64-float aligned blocks
    shape = _i64(2, 13, 18, 22, 4)
    per = sum((13 >> m) * (18 >> m) * (22 >> m) * 4 for m in (1, 2, 3))  # level 4 has D = 0
    stride = (per + 63) // 64 * 64
    assert lib.gft_mip_atlas_bytes(shape, 6) == ((stride * 4 * 2 + 255) // 256) * 256
    assert lib.gft_sample_grid_transform_workspace(shape, 0, 3) == 0
    assert lib.gft_sample_grid_transform_bwd_workspace(shape, 0, 0, 0) == ((13 * 18 * 22 * 4 + 255) // 256) * 256
    assert lib.gft_sample_grid_transform_bwd_workspace(shape, 2, 6, 1) == 2 * lib.gft_mip_atlas_bytes(shape, 6)
    assert lib.gft_sample_grid_lut_bwd_workspace(shape, 1) == 0
    out = _i32(7, 5, 6)
    CLAMP, BORDER = 1, 0
    assert lib.gft_raymarch_grid_transform_bwd_workspace(shape, 3, out, CLAMP, 1, 1) == 0
    vox = 13 * 18 * 22
    assert (lib.gft_raymarch_grid_transform_bwd_workspace(shape, 3, out, CLAMP, 1, 0)
            == ((vox * 4 * 4 + 255) // 256) * 256 + vox * 4)
    # BORDER: per-sample gradients of one batch's views (at most 16 per chunk) plus the sample ranges
    for cams, sep, views in ((3, 1, 3), (20, 1, 16), (3, 0, 1)):
        for gm in (0, 1):
            want = ((views * 7 * 30 * 4 * 4 + 255) // 256) * 256 + views * 30 * 8
            assert lib.gft_raymarch_grid_transform_bwd_workspace(shape, cams, out, BORDER, sep, gm) == want
    assert lib.gft_advect_saved_bytes(_i64(1, 4, 5, 6, 1)) > 0
    # shadow map: prefix sums of the light frustum; the pyramid of the step-scaled density
    assert lib.gft_shadow_transmission_saved_bytes(_i32(70, 13, 11)) == 70 * 13 * 11 * 4
    assert lib.gft_shadow_transmission_saved_bytes(_i32(0, 13, 11)) == 0
    dshape = _i64(1, 13, 18, 22, 1)
    assert lib.gft_shadow_transmission_fwd_workspace(dshape, 0, 4) == 0
    assert lib.gft_shadow_transmission_fwd_workspace(dshape, 2, 3) == lib.gft_mip_atlas_bytes(dshape, 3)


def _eye(n=1):
    return np.ascontiguousarray(np.broadcast_to(np.eye(4, dtype=np.float32), (n, 4, 4)))


def test_validation_returns_before_device_work(lib):
    from gft import _lib
    M, F = _eye(), np.array([[1, 10, -1, 1, 1, -1]], dtype=np.float32)
    vp = M.ctypes.data_as(ctypes.c_void_p)
    fp = F.ctypes.data_as(ctypes.c_void_p)
    null = ctypes.c_void_p(0)
    # C == 3 is Unimplemented (resample_grid.cc:330-331, raymarch_grid.cc)
    rc = lib.gft_raymarch_grid_transform_fwd(null, null, _i64(1, 4, 4, 4, 3), vp, 1, vp, vp, fp, 1, _i32(4, 4, 4), 1,
                                             0, 0, 0, 1, null)
    assert rc == _lib.GFT_ERR_UNIMPLEMENTED
    assert b"Channel" in lib.gft_last_error()
    # keep_dims=False with a data batch of 2 is rejected
    M2 = _eye(2)
    rc = lib.gft_raymarch_grid_transform_fwd(null, null, _i64(2, 4, 4, 4, 1), M2.ctypes.data_as(ctypes.c_void_p), 2,
                                             vp, vp, fp, 1, _i32(4, 4, 4), 1, 0, 0, 0, 1, null)
    assert rc == _lib.GFT_ERR_INVALID_ARGUMENT
    # model matrix batch mismatch
    rc = lib.gft_sample_grid_transform_fwd(null, null, _i64(2, 4, 4, 4, 1), vp, 1, vp, vp, fp, 1, _i32(4, 4, 4), 1, 0,
                                           0, 0, 0.0, 0, 0.5, 1, null, 0, null)
    assert rc == _lib.GFT_ERR_INVALID_ARGUMENT and b"model matrix" in lib.gft_last_error()
    # transform op: WRAP boundary and a cell-centre offset other than 0.5 are invalid
    rc = lib.gft_sample_grid_transform_fwd(null, null, _i64(1, 4, 4, 4, 1), vp, 1, vp, vp, fp, 1, _i32(4, 4, 4), 1, 2,
                                           0, 0, 0.0, 0, 0.5, 1, null, 0, null)
    assert rc == _lib.GFT_ERR_INVALID_ARGUMENT
    rc = lib.gft_sample_grid_transform_fwd(null, null, _i64(1, 4, 4, 4, 1), vp, 1, vp, vp, fp, 1, _i32(4, 4, 4), 1, 0,
                                           0, 0, 0.0, 0, 0.0, 1, null, 0, null)
    assert rc == _lib.GFT_ERR_INVALID_ARGUMENT and b"cell center" in lib.gft_last_error()
    # mips need num_mipmaps > 0
    rc = lib.gft_sample_grid_transform_fwd(null, null, _i64(1, 4, 4, 4, 1), vp, 1, vp, vp, fp, 1, _i32(4, 4, 4), 1, 0,
                                           2, 0, 0.0, 0, 0.5, 1, null, 0, null)
    assert rc == _lib.GFT_ERR_INVALID_ARGUMENT and b"num_mipmaps" in lib.gft_last_error()
    # REFERENCE transform gradient with mips is rejected (the op registers NONE only)
    rc = lib.gft_sample_grid_transform_bwd(null, null, _i64(1, 4, 4, 4, 1), null, vp, 1, vp, vp, fp, 1, _i32(4, 4, 4),
                                           1, 0, 2, 2, 0.0, 0, 0.5, 1, 0, null, 0, null)
    assert rc == _lib.GFT_ERR_INVALID_ARGUMENT
    # LuT: normalised coordinates need offset 0.5; the table must have 4 channels
    rc = lib.gft_sample_grid_lut_fwd(null, null, _i64(1, 4, 4, 4, 1), null, _i64(1, 3, 3, 3, 4), 1, 0, 0, 0, 0.0, 0.0,
                                     1, 1, 1, null, 0, null)
    assert rc == _lib.GFT_ERR_INVALID_ARGUMENT
    rc = lib.gft_sample_grid_lut_fwd(null, null, _i64(1, 4, 4, 4, 1), null, _i64(1, 3, 3, 3, 3), 1, 0, 0, 0, 0.0, 0.0,
                                     1, 1, 0, null, 0, null)
    assert rc == _lib.GFT_ERR_INVALID_ARGUMENT
    # reduce blend: REFERENCE has no ALPHA_ADDITIVE gradient
    rc = lib.gft_reduce_grid_blend_bwd(null, null, _i64(1, 4, 4, 4, 4), null, null, 2, 0, 0, null)
    assert rc == _lib.GFT_ERR_UNIMPLEMENTED
    # workspace too small is its own status
    rc = lib.gft_raymarch_grid_transform_bwd(null, ctypes.c_void_p(16), _i64(1, 4, 4, 4, 1), ctypes.c_void_p(16),
                                             ctypes.c_void_p(16), vp, 1, vp, vp, fp, 1, _i32(4, 4, 4), 1, 0, 0, 0, 1,
                                             0, null, 0, ctypes.c_void_p(16))
    assert rc == _lib.GFT_ERR_WORKSPACE
    # shadow map: one-channel, batch-1 density; BORDER/CLAMP; BEER_LAMBERT/ALPHA; workspace for mips
    def shadow(shape, boundary=0, blend=0, mm=0, nm=0, ws=0, out=(8, 8, 8)):
        return lib.gft_shadow_transmission_fwd(null, ctypes.c_void_p(16), _i64(*shape), vp, vp, vp, fp, _i32(*out),
                                               ctypes.c_float(0.5), 1, boundary, mm, nm, ctypes.c_float(0.0), blend,
                                               null, null, ws, null, ctypes.c_void_p(16))
    assert shadow((1, 4, 4, 4, 2)) == _lib.GFT_ERR_INVALID_ARGUMENT
    assert shadow((2, 4, 4, 4, 1)) == _lib.GFT_ERR_INVALID_ARGUMENT
    assert shadow((1, 4, 4, 4, 1), boundary=2) == _lib.GFT_ERR_INVALID_ARGUMENT
    assert shadow((1, 4, 4, 4, 1), blend=3) == _lib.GFT_ERR_INVALID_ARGUMENT and b"blend" in lib.gft_last_error()
    assert shadow((1, 4, 4, 4, 1), out=(8, 0, 8)) == _lib.GFT_ERR_INVALID_ARGUMENT
    assert shadow((1, 8, 8, 8, 1), mm=2, nm=2) == _lib.GFT_ERR_WORKSPACE
    rc = lib.gft_shadow_transmission_bwd(null, null, _i32(8, 8, 8), 0, 0, null, null, null)
    assert rc == _lib.GFT_ERR_INVALID_ARGUMENT
    # epilogue: divergence shape and gradient triple; Adam tensor count, step and sizes
    ws = (ctypes.c_float * 3)(1.0, 1.0, 1.0)
    assert lib.gft_divergence_loss(null, null, null, null, _i32(4, 0, 4), ws, ctypes.c_float(1.0), 0, null, 0, null,
                                   null, null, null) == _lib.GFT_ERR_INVALID_ARGUMENT
    p16 = ctypes.c_void_p(16)
    assert lib.gft_divergence_loss(null, p16, p16, p16, _i32(4, 4, 4), ws, ctypes.c_float(1.0), 0, p16, 1 << 20, p16,
                                   null, null, p16) == _lib.GFT_ERR_INVALID_ARGUMENT
    need = lib.gft_divergence_loss_workspace(_i32(4, 4, 4))
    assert need > 0 and lib.gft_divergence_loss(null, p16, p16, p16, _i32(4, 4, 4), ws, ctypes.c_float(1.0), 0, p16,
                                                need - 1, null, null, null, p16) == _lib.GFT_ERR_WORKSPACE
    ptrs = (ctypes.c_void_p * 9)()
    sizes = (ctypes.c_int64 * 9)()
    assert lib.gft_adam_step(null, 9, ptrs, ptrs, ptrs, ptrs, sizes, 1e-3, 0.9, 0.999, 1e-8, 1) == \
        _lib.GFT_ERR_INVALID_ARGUMENT
    assert lib.gft_adam_step(null, 1, ptrs, ptrs, ptrs, ptrs, sizes, 1e-3, 0.9, 0.999, 1e-8, 0) == \
        _lib.GFT_ERR_INVALID_ARGUMENT
    sizes[0] = 5
    assert lib.gft_adam_step(null, 1, ptrs, ptrs, ptrs, ptrs, sizes, 1e-3, 0.9, 0.999, 1e-8, 1) == \
        _lib.GFT_ERR_INVALID_ARGUMENT and b"tensor 0" in lib.gft_last_error()
    assert lib.gft_adam_step(null, 0, None, None, None, None, None, 1e-3, 0.9, 0.999, 1e-8, 1) == _lib.GFT_OK
    # the device-counter form needs its counter, and keeps the tensor limit
    assert lib.gft_adam_step_device(null, 0, None, None, None, None, None, 1e-3, 0.9, 0.999, 1e-8, None) == \
        _lib.GFT_ERR_INVALID_ARGUMENT and b"step counter" in lib.gft_last_error()
    cnt = (ctypes.c_int64 * 1)()
    assert lib.gft_adam_step_device(null, 9, ptrs, ptrs, ptrs, ptrs, sizes, 1e-3, 0.9, 0.999, 1e-8, cnt) == \
        _lib.GFT_ERR_INVALID_ARGUMENT
    # folded image loss: an unknown loss kind is rejected before any launch
    light = (ctypes.c_float * 3)(1.0, 1.0, 1.0)
    rc = lib.gft_render_density_bwd_loss(null, p16, _i64(1, 4, 4, 4, 1), vp, 1, vp, vp, fp, 1, _i32(4, 4, 4), light,
                                         ctypes.c_float(1.0), 1, p16, p16, null, 7, ctypes.c_float(1.0), null, 0, null,
                                         0, p16, p16)
    assert rc == _lib.GFT_ERR_INVALID_ARGUMENT and b"loss_kind" in lib.gft_last_error()
    # the CLAMP render needs its corner-quad copy in both directions, BORDER for the walk without a
    # saved buffer; WRAP needs none; the lit render holds two copies and a second per-sample plane
    sh = _i64(1, 4, 5, 6, 1)
    nq = lib.gft_render_density_fwd_workspace(sh, 1)
    assert nq > 0 and lib.gft_render_density_bwd_workspace(sh, 1, _i32(4, 4, 4), 1) == nq
    assert lib.gft_render_density_bwd_workspace(sh, 1, _i32(4, 4, 4), 0) == nq
    assert lib.gft_render_density_bwd_workspace(sh, 1, _i32(4, 4, 4), 2) == 0
    assert lib.gft_render_density_lit_fwd_workspace(sh, 0) == 2 * nq
    ns = lib.gft_render_density_saved_bytes(sh, 2, _i32(4, 4, 4), 0)
    # a second per-sample array: 2 views x 4 waves (one 16x16 tile) x 64 lanes x 4 samples x 4 B
    assert lib.gft_render_density_lit_saved_bytes(sh, 2, _i32(4, 4, 4), 0) == ns + 2 * 4 * 64 * 4 * 4
    assert lib.gft_render_density_fwd_workspace(sh, 2) == 0
    rc = lib.gft_render_density_bwd_loss(null, p16, sh, vp, 1, vp, vp, fp, 1, _i32(4, 4, 4), light,
                                         ctypes.c_float(1.0), 1, p16, p16, null, 0, ctypes.c_float(1.0), null, 0, null,
                                         0, p16, p16)
    assert rc == _lib.GFT_ERR_WORKSPACE
    # fused advection: scalar grids only
    rc = lib.gft_advect_fwd(null, null, _i64(1, 4, 4, 4, 2), null, ctypes.c_float(1.0), 2, 1, null, null, 0)
    assert rc == _lib.GFT_ERR_UNIMPLEMENTED
    # the bfloat16-storage forms validate as the float32 ones, before any launch
    rc = lib.gft_advect_fwd_bf16(null, null, _i64(1, 4, 4, 4, 2), null, ctypes.c_float(1.0), 2, 1, null, null, 0)
    assert rc == _lib.GFT_ERR_UNIMPLEMENTED
    rc = lib.gft_advect_fwd_bf16(null, null, _i64(1, 4, 4, 4, 1), null, ctypes.c_float(1.0), 2, 1, null, null, 0)
    assert rc == _lib.GFT_ERR_INVALID_ARGUMENT and b"null pointer" in lib.gft_last_error()
    rc = lib.gft_advect_fwd_bf16(null, p16, _i64(1, 4, 4, 4, 1), p16, ctypes.c_float(1.0), 3, 1, p16, p16, 1 << 20)
    assert rc == _lib.GFT_ERR_INVALID_ARGUMENT
    rc = lib.gft_advect_bwd_bf16(null, p16, _i64(1, 4, 4, 4, 1), p16, ctypes.c_float(1.0), 2, 1, p16, p16, p16, p16,
                                 null, 0)
    assert rc == _lib.GFT_ERR_WORKSPACE


def test_ops_refuse_host_tensors():
    import torch
    from gft import _lib
    from gft.ops import blending_ops, raymarching_ops, sampling_ops
    x = torch.zeros((1, 4, 4, 4, 1))
    M = np.eye(4, dtype=np.float32)[None]
    F = np.array([[1, 10, -1, 1, 1, -1]], dtype=np.float32)
    with pytest.raises(_lib.InvalidArgumentError):
        raymarching_ops.raymarch_grid_transform(x, M, M, M, F, (4, 4, 4))
    with pytest.raises(_lib.InvalidArgumentError):
        sampling_ops.sample_grid_transform(x, M, M, M, F, (4, 4, 4))
    with pytest.raises(_lib.InvalidArgumentError):
        blending_ops.reduce_grid_blend(x)


def test_projection_and_view_matrices_follow_opengl():
    from gft.camera import Camera, camera_params_batch, model_matrices
    from gft.transform import GridTransform, Transform
    t = GridTransform([10, 20, 30], translation=[1.0, 2.0, 3.0], rotation_deg=[0, 30, 0])
    cam = Camera(t, nearFar=[2.0, 12.0], fov=40.0, aspect=16 / 9.0)
    n, f = 2.0, 12.0
    r = np.tan(np.radians(20.0)) * n          # horizontal fov (camera.py:64-70)
    top = r / (16 / 9.0)
    P = cam.projection_matrix()
    assert np.allclose(P, [[n / r, 0, 0, 0], [0, n / top, 0, 0], [0, 0, -(f + n) / (f - n), -2 * f * n / (f - n)],
                           [0, 0, -1, 0]], atol=1e-6)
    assert np.allclose(cam.frustum(), [n, f, -r, r, top, -top], atol=1e-6)
    # view = inverse of the camera transform; the ops receive transposes
    V, Pc, F = camera_params_batch([cam])
    assert np.allclose(V[0].T @ t.get_transform_matrix(), np.eye(4), atol=1e-5)
    assert np.allclose(Pc[0], P.T)
    # grid model matrix: centring translation then rotation then translation
    g = GridTransform([4, 6, 8], translation=[5, 0, 0], center=True)
    Mg = model_matrices([g])[0].T
    assert np.allclose(Mg @ [4, 3, 2, 1], [5, 0, 0, 1])  # the grid centre (x,y,z) = (4,3,2) maps to the translation
    rot = Transform(rotation_deg=[0, 90, 0]).get_transform_matrix()
    assert np.allclose(rot @ [1, 0, 0, 1], [0, 0, -1, 1], atol=1e-12)


def test_workload_geometry():
    """Sample counts and pixel totals of the benchmark workloads (one depth
    sample per grid cell across the grid's diagonal in xz)."""
    from gft import synthetic as syn
    expect = {"c1": (91, 145920), "c2": (182, 583680)}
    for name, (S, npx) in expect.items():
        w = syn.WORKLOADS[name]
        cams, s = syn.arc_cameras(w["grid"], w["image"], w["views"])
        assert s == S
        assert w["views"] * w["image"][0] * w["image"][1] == npx
        assert abs(cams[0].far - cams[0].near - np.sqrt(w["grid"][2] ** 2 + w["grid"][0] ** 2)) < 1e-4


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_graft_entry_points_exist():
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    assert callable(ge.build) and callable(ge.smoke)


def _batch_records(specs, brick=0):
    """gft_batch_recon records of small synthetic reconstructions (host
    matrices real, device pointers placeholders: nothing here touches a GPU)."""
    from gft.batch import BatchRecon
    from gft.camera import camera_params_batch, model_matrices
    from gft import synthetic as syn
    arr = (BatchRecon * len(specs))()
    keep = []
    for k, (grid, image, views) in enumerate(specs):
        cams, S = syn.arc_cameras(grid, image, views)
        M = np.ascontiguousarray(model_matrices([syn.grid_transform(grid)]), dtype=np.float32)
        V, P, F = (np.ascontiguousarray(x, dtype=np.float32) for x in camera_params_batch(cams))
        keep += [M, V, P, F]
        c = arr[k]
        c.grid_dhw[:] = list(grid)
        c.output_shape[:] = [S, image[0], image[1]]
        c.num_views = views
        c.depth_step = float(cams[0].depth_step)
        c.light[:] = [1.0, 1.0, 1.0]
        c.loss_scale = 1.0
        c.brick = brick
        c.M, c.V, c.P, c.frustum = (x.ctypes.data for x in (M, V, P, F))
        c.density = c.targets = c.images = c.density_grad = c.loss = 0x1000  # never dereferenced here
    return arr, keep


def test_batch_workspace_and_validation(lib):
    from gft import _lib
    specs = [((12, 20, 14), (24, 14), 3), ((8, 9, 10), (16, 16), 1)]
    arr, keep = _batch_records(specs)
    p = ctypes.cast(arr, ctypes.c_void_p)
    ws = lib.gft_batch_render_workspace(2, p, 3)
    # at least 4 bytes of c_k per frustum sample and one 16-byte record per pixel of every tile
    samples = sum(v * ((h + 15) // 16) * ((w + 15) // 16) * 256 * arr[k].output_shape[0]
                  for k, (_, (h, w), v) in enumerate(specs))
    assert ws >= 4 * samples + 16 * samples // arr[0].output_shape[0]
    # the brick (the gather's voxel block) changes launch sizes, not the workspace
    for b in (1, 2):
        arr_b, keep_b = _batch_records(specs, brick=b)
        assert lib.gft_batch_render_workspace(2, ctypes.cast(arr_b, ctypes.c_void_p), 3) == ws
    # one more reconstruction needs more; fewer mip levels need less
    arr3, keep3 = _batch_records(specs + [((16, 16, 16), (20, 20), 2)])
    assert lib.gft_batch_render_workspace(3, ctypes.cast(arr3, ctypes.c_void_p), 3) > ws
    assert 0 < lib.gft_batch_render_workspace(2, p, 0) < ws
    # a set-up call with a too small workspace fails before any device work
    assert lib.gft_batch_render_prepare(None, 2, p, 3, None, 0) == _lib.GFT_ERR_WORKSPACE
    # invalid records: the query returns 0 and the set-up call names the problem
    bad = [("num_views", 0), ("num_views", 9), ("brick", 3)]
    for field, value in bad:
        arr_x, keep_x = _batch_records(specs)
        setattr(arr_x[1], field, value)
        px = ctypes.cast(arr_x, ctypes.c_void_p)
        assert lib.gft_batch_render_workspace(2, px, 3) == 0, field
        assert lib.gft_batch_render_prepare(None, 2, px, 3, None, 0) == _lib.GFT_ERR_INVALID_ARGUMENT, field
        assert b"reconstruction 1" in lib.gft_last_error(), field
    assert lib.gft_batch_render_workspace(2, p, 4) == 0  # at most 3 mip levels
    assert lib.gft_batch_render_workspace(0, p, 3) == 0
    arr_h, keep_h = _batch_records([((5000, 4000, 8), (16, 16), 1)])  # (H+2)(D+2) >= 2^24
    assert lib.gft_batch_render_workspace(1, ctypes.cast(arr_h, ctypes.c_void_p), 3) == 0


def test_render_saved_tail_covers_the_rewritten_records(lib):
    """The render backward rewrites the per-pixel records (16 B a pixel) and
    the per-tile loss partials at the end of the ambient saved buffer; the
    slice the autograd functions keep for a retained graph's second backward
    must reach back over both, whatever the shape (the per-sample values
    before them are 4 B per sample and stay untouched)."""
    import torch
    from gft import pipeline as pl
    from helpers import scene
    for grid, image, views in [((12, 20, 14), (24, 14), 3), ((8, 9, 10), (7, 33), 1), ((30, 36, 40), (96, 80), 18)]:
        sc = scene(grid=grid, image=image, views=views, seed=3)
        cams = pl._Cameras(sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"])
        n = pl.render_saved_bytes((1,) + tuple(grid) + (1,), cams, "BORDER")
        h, w = image
        plane = h * w * views
        tiles = ((w + 15) // 16) * ((h + 15) // 16) * views
        a256 = lambda b: (b + 255) // 256 * 256
        rewritten = a256(plane * 16) + a256(tiles * 4)
        samples = n - rewritten
        assert samples >= sc["out_dhw"][0] * plane * 4  # the per-sample values come first
        tail = pl._rewritten_tail(torch.zeros(n, dtype=torch.uint8), n, cams)
        assert tail.numel() >= rewritten and tail.numel() <= n
