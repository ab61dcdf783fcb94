#!/usr/bin/env python
"""Benchmark of the fused transport/render step on one MI355X.

One step = MacCormack advection (MC_SMOOTH) of the density by a curl-noise
velocity, ambient-light emission-absorption render of the advected density
into every view, L2 loss against target images, and the exact backward to the
density and the velocity. Workload c2: grid 128x224x128 (x,y,z), 48 Gaussian
blobs (seed 0), curl-noise velocity max 1.5 cells/step (seed 1), 5 views of
256x456 px on a -60..60 deg arc, 182 samples per ray; targets are renders of a
second blob density (seed 2).

Prints one JSON line: voxel-samples/s (one 8-corner trilinear evaluation = 1
voxel-sample; N_vs = 6V + 2 sum W*H*S per step), the HBM roofline of the
dominant kernel from HIP events around its launches in a second pass over the
same steps, the plain per-op FP32 path (the reference's op composition on
libgft's per-op kernels) timed on the same inputs, and a CPU baseline (the
FP32 oracle, one full step on the host cores).

``--copy-kernel`` instead runs a 1 GiB device copy: the calibration kernel of
known traffic for the HBM counters (FETCH_SIZE / WRITE_SIZE) under rocprofv3.

``--ops`` instead times every per-op kernel of the reference's op API
(resample, mips, LoD, their gradients, the gather gradient, raymarch, depth
blend, LuT sampling) at the c2 sizes, each against its compulsory bytes.

``--shadow`` instead times the spot-light shadow map at the production
settings (256x196x196, LINEAR mips, gather gradient) over the c2 density:
the fused shadow op against the op sequence it replaces.

``--epilogue`` instead times render + AE/background image loss + divergence loss
+ Adam at c2: the fused kernels against torch autograd / torch.optim.Adam over
libgft's render kernels.

``--precision`` instead runs workload c5 (200x356x200, 8 views of 540x960)
with the density and velocity stored in bfloat16: ms per fused step against
float32 storage, and the errors of images, loss and gradients against the
FP64 oracle on the bfloat16 values (the float32 arithmetic) and on the
unrounded inputs (arithmetic and storage). Variants: the render sampling the
advected density stored in bfloat16 (gather reduction), and the same BORDER
render gradient on a float32 render reduced two ways, the deterministic
segmented gather and the reverse walk with float32 atomics.

``--batched`` instead runs workload c4: 32 independent reconstructions of
uneven grid sizes (32^3 to 128x224x128) and view counts (1-8). The headline
is the batched mipmapped density-reconstruction step (every reconstruction's
non-fused render with LINEAR mips and per-sample LoD, L2 loss, exact
gradient, in a few launches over one table; bricks autotuned per grid class),
also replayed from a graph and against running each reconstruction alone.
Beside it: the fused transport step of every reconstruction issued eagerly,
replayed from one graph on one stream, and from one graph with a branch per
reconstruction on its own stream.

``--lit`` instead runs the c2 step under the production light (a spot light
with a 256x196x196 shadow map, LINEAR mips with 4 levels and the gather
sampling gradient, plus an ambient light): advect -> shadow map -> light
factor -> render -> L2 loss -> backward through all of it, with the fused
lit render against the renderer's generic path (the light grid materialised
and the 4-channel raymarch op) on the same shadow ops.

``--chain`` instead measures the global transport chain (workload c3): an
initial density advected through 20 per-frame velocity fields, every frame
rendered to 5 views plus one unseen view and compared with its targets, the
summed loss back-propagated through the whole chain with checkpointing every
4 frames; peak device memory is reported against keeping every frame.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "global-flow-transport_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "voxel-samples/s fwd+bwd (advect+resample+raymarch), HBM GB/s vs plain FP32"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_c2.txt")


def build_inputs(name):
    import numpy as np

    from gft import synthetic as syn
    from gft.camera import camera_params_batch, model_matrices
    w = syn.WORKLOADS[name]
    grid, image = w["grid"], w["image"]
    d0 = syn.gaussian_blobs(grid, w["blobs"], seed=0)
    vel = syn.curl_noise_velocity(grid, seed=1, max_speed=w["max_speed"])
    dt = syn.gaussian_blobs(grid, w["blobs"], seed=2)
    cams, S = syn.arc_cameras(grid, image, w["views"])
    M = model_matrices([syn.grid_transform(grid)])
    V, P, F = camera_params_batch(cams)
    return dict(grid=grid, image=image, S=S, views=w["views"], d0=d0, vel=vel, d_target=dt, M=M, V=V, P=P, F=F,
                step=float(cams[0].depth_step), out_dhw=(S,) + tuple(image))


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(inp, target_np, nvs, ray_samples, boundary="BORDER", warmups=2, runs=5):
    """The FP32 oracle (C++/OpenMP restatement of the reference ops) on the
    same workload: all host threads on one full fwd+bwd step (warm-ups, then
    the median of `runs`), and one thread on a bounded sample of it (the first
    view: the whole advection and one view's render, median of 3 after one
    warm-up). Returns the cpu_baseline record (voxel-samples/s)."""
    import ctypes

    import numpy as np
    from oracle import oracle as orc
    L = orc.lib()
    L.orc_set_threads.argtypes = [ctypes.c_int]
    threads = int(L.orc_num_threads())
    V = int(np.prod(inp["d0"].shape))

    def timed(views, n_warm, n_runs):
        sl = slice(0, views)
        args = (inp["d0"], inp["vel"], target_np[:, sl], inp["M"], inp["V"][sl], inp["P"][sl], inp["F"][sl],
                inp["out_dhw"], inp["step"])
        ts = []
        for k in range(n_warm + n_runs):
            t0 = time.perf_counter()
            orc.step_fwd_bwd(*args, prec=0, boundary_render=boundary)
            if k >= n_warm:
                ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    nviews = int(inp["V"].shape[0])
    t_all = timed(nviews, warmups, runs)
    L.orc_set_threads(1)
    try:
        t_one = timed(1, 1, 3)
    finally:
        L.orc_set_threads(threads)
    nvs_one = 6 * V + 2 * ray_samples // nviews
    return {"value": nvs / t_all, "unit": "voxel-samples/s", "cores": threads, "kind": "port",
            "sample": "one full fwd+bwd step of the same workload with the FP32 oracle, median of %d runs after %d "
                      "warm-ups (%.2f s per step)" % (runs, warmups, t_all),
            "single_thread": {"value": nvs_one / t_one, "cores": 1,
                              "sample": "the advection and the first view's render and backward, median of 3 after "
                                        "one warm-up (%.2f s)" % t_one},
            "cpu_model": _cpu_model()}


def measured_traffic(kernel, workload):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel calibration),
    or None when the summary has no line for it."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        parts = line.split()
        if len(parts) >= 4 and parts[0] == "traffic" and parts[1] == workload and parts[2] == kernel:
            return float(parts[3])
    return None


def copy_kernel():
    import torch
    n = (1 << 30) // 4
    src = torch.ones(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))


def ops_bench(args):
    """One JSON line with the per-op kernels at c2 sizes: HIP-event time per
    call (median of args.steps calls after one warm-up) and the achieved rate
    on the op's compulsory bytes (inputs read once, outputs written once)."""
    import numpy as np
    import torch

    from gft.ops import blending_ops, raymarching_ops, sampling_ops

    inp = build_inputs("c2")
    dev = torch.device("cuda:0")
    M, V, P, F, out_dhw = inp["M"], inp["V"], inp["P"], inp["F"], inp["out_dhw"]
    d = torch.from_numpy(inp["d0"]).to(dev)
    vel = torch.from_numpy(inp["vel"]).to(dev)
    grid4 = torch.cat([d, d, d, d], dim=-1) * inp["step"]  # the renderer's [light rgb, density] grid
    Vx = int(d.numel())
    nv = int(np.asarray(V).reshape(-1, 16).shape[0])
    cells = nv * int(np.prod(out_dhw))
    px = nv * out_dhw[1] * out_dhw[2]
    lut = torch.cat([-vel, torch.zeros_like(vel[..., :1])], dim=-1).contiguous()

    def timeit(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(args.steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b))
        return float(np.median(ts))

    res = {}

    def op(name, fn, alg_bytes, what):
        ms = timeit(fn)
        res[name] = {"ms": ms, "alg_bytes": int(alg_bytes), "GBps": alg_bytes / (ms * 1e-3) / 1e9, "work": what}

    tf = dict(interpolation="LINEAR", boundary="BORDER", coordinate_mode="TRANSFORM_LINDEPTH")
    fr = sampling_ops.sample_grid_transform(grid4, M, V, P, F, out_dhw, **tf)
    op("A1 sample_grid_transform fwd (C=4)", lambda: sampling_ops.sample_grid_transform(grid4, M, V, P, F, out_dhw,
                                                                                      **tf),
       Vx * 16 + cells * 16, "%d frustum cells" % cells)
    op("A2 build_mips (3 levels, C=4)", lambda: sampling_ops.build_mips(grid4, 3), Vx * 16 * (1 + 1.0 / 7),
       "levels 1-3")
    tfm = dict(tf, mipmapping="LINEAR", num_mipmaps=3)
    op("A2 sample_grid_transform fwd, LINEAR mips (C=4)",
       lambda: sampling_ops.sample_grid_transform(grid4, M, V, P, F, out_dhw, **tfm),
       Vx * 16 * (1 + 1.0 / 7) + cells * 16, "%d frustum cells, per-sample LoD" % cells)
    mv = np.stack([V[i] @ M[0] for i in range(len(V))]).astype(np.float32)  # as the op's callers form it
    op("A2 lod_transform", lambda: sampling_ops.lod_transform(tuple(d.shape[1:4]), mv, P, F, out_dhw),
       cells * 16, "%d frustum cells" % cells)
    g_fr = torch.rand_like(fr)
    for gm in ("REFERENCE", "EXACT"):
        op("A3 sample_grid_transform_grad %s (C=4)" % gm,
           lambda gm=gm: sampling_ops.sample_grid_transform_grad(grid4, g_fr, M, V, P, F, grad_mode=gm, **tf),
           cells * 16 + Vx * 16, "%d frustum cells" % cells)
    op("A4 gather gradient (C=4)", lambda: sampling_ops.sample_grid_transform_gather_grad(
        g_fr, M, V, P, F, tuple(d.shape[1:4]), interpolation="LINEAR", boundary="BORDER"),
       cells * 16 + Vx * 16, "%d frustum cells" % cells)
    img = raymarching_ops.raymarch_grid_transform(grid4, M, V, P, F, out_dhw)
    op("A5 raymarch_grid_transform fwd (C=4)", lambda: raymarching_ops.raymarch_grid_transform(grid4, M, V, P, F,
                                                                                            out_dhw),
       Vx * 16 + px * 16, "%d rays x %d samples" % (px, out_dhw[0]))
    g_img = torch.rand_like(img)
    for gm in ("REFERENCE", "EXACT"):
        op("A6 raymarch_grid_transform_grad %s (C=4)" % gm,
           lambda gm=gm: raymarching_ops.raymarch_grid_transform_grad(grid4, img, g_img, M, V, P, F, out_dhw,
                                                                       grad_mode=gm),
           Vx * 16 * 2 + px * 32, "%d rays x %d samples" % (px, out_dhw[0]))
    fr5 = fr.reshape((nv,) + tuple(fr.shape[2:]))
    bl = blending_ops.reduce_grid_blend(fr5, "BEER_LAMBERT")
    op("A7 reduce_grid_blend fwd", lambda: blending_ops.reduce_grid_blend(fr5, "BEER_LAMBERT"), cells * 16 + px * 16,
       "%d cells" % cells)
    g_bl = torch.rand_like(bl)
    op("A7 reduce_grid_blend_grad", lambda: blending_ops.reduce_grid_blend_grad(g_bl, bl, fr5, "BEER_LAMBERT"),
       cells * 32 + px * 32, "%d cells" % cells)
    lk = dict(interpolation="LINEAR", boundary="CLAMP", separate_camera_batch=False, relative_coords=True)
    op("A8 sample_grid_lut fwd (advection LuT)", lambda: sampling_ops.sample_grid_lut(d, lut, **lk),
       Vx * (4 + 16 + 4), "%d voxels" % Vx)
    g_l = torch.rand((1, 1) + tuple(d.shape[1:]), device=dev)
    for gm in ("REFERENCE", "EXACT"):
        op("A9 sample_grid_lut_grad %s" % gm,
           lambda gm=gm: sampling_ops.sample_grid_lut_grad(d, g_l, lut, grad_mode=gm, **lk),
           Vx * (4 + 4 + 16 + 4 + 16), "%d voxels" % Vx)
    print(json.dumps({"metric": "per-op kernel time at c2 sizes", "unit": "ms", "n_gpus": 1, "steps": args.steps,
                      "config": {"workload": "c2 sizes: grid 128x224x128, 5 views 256x456 px, 182 samples"},
                      "ops": res}))


def chain_bench(args):
    """One JSON line for the checkpointed transport chain at c3."""
    import numpy as np
    import torch

    from gft import synthetic as syn
    from gft.camera import camera_params_batch, model_matrices
    from gft.chain import TransportChain
    from gft.pipeline import _Cameras, advect_forward, render_density_forward

    w = syn.WORKLOADS["c3"]
    grid, image, frames, k = w["grid"], w["image"], 20, 4
    cams, S = syn.arc_cameras(grid, image, w["views"], extra_yaws=(w["unseen_yaw"],))
    M = model_matrices([syn.grid_transform(grid)])
    V, P, F = camera_params_batch(cams)
    step = float(cams[0].depth_step)
    out_dhw = (S,) + tuple(image)
    dev = torch.device("cuda:0")
    d0 = torch.from_numpy(syn.gaussian_blobs(grid, w["blobs"], seed=0)).to(dev)
    vels = [torch.from_numpy(syn.curl_noise_velocity(grid, seed=100 + f, max_speed=w["max_speed"])).to(dev)
            for f in range(frames - 1)]
    # targets: a second density (seed 2) advected through the same velocities and rendered
    c = _Cameras(M, V, P, F, out_dhw)
    dt = torch.from_numpy(syn.gaussian_blobs(grid, w["blobs"], seed=2)).to(dev)
    targets = []
    for f in range(frames):
        targets.append(render_density_forward(dt, c, step)[..., :3].contiguous())
        if f < frames - 1:
            dt, _ = advect_forward(dt, vels[f])
    del dt
    torch.cuda.synchronize()

    def run(every, steps):
        chain = TransportChain(grid, M, V, P, F, out_dhw, step, checkpoint_every=every)
        chain.run(d0, vels, targets)
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(steps):
            loss, gd0, gv = chain.run(d0, vels, targets)
            del gd0, gv
        b.record()
        torch.cuda.synchronize()
        peak = torch.cuda.max_memory_allocated() - base
        return a.elapsed_time(b) / steps, peak, float(loss.item())

    ms, peak_k, loss = run(k, args.steps)
    ms_all, peak_all, _ = run(1, 1)
    Vox = int(np.prod(grid))
    px = len(cams) * image[0] * image[1]
    # per frame: render fwd + bwd (2 sum W*H*S); per advection: MacCormack fwd 4V + bwd 2V
    nvs = frames * 2 * px * S + (frames - 1) * 6 * Vox
    print(json.dumps({
        "metric": "voxel-samples/s fwd+bwd of the checkpointed transport chain",
        "value": nvs / (ms * 1e-3), "unit": "voxel-samples/s", "n_gpus": 1, "steps": args.steps,
        "ms_per_chain": ms, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c3: grid %dx%dx%d (x,y,z), %d frames, %d views (5 + unseen yaw 90) %dx%d px, "
                               "%d samples/ray, MacCormack+MC_SMOOTH, EXACT adjoint"
                               % (grid[2], grid[1], grid[0], frames, len(cams), image[1], image[0], S),
                   "checkpoint_every": k, "voxel_samples_per_chain": nvs},
        "peak_bytes_checkpointed": peak_k, "peak_bytes_every_frame": peak_all,
        "ms_per_chain_every_frame": ms_all, "loss": loss}))


def shadow_bench(args):
    """One JSON line for the spot-light shadow map at the production settings
    (256x196x196 shadow map, LINEAR mips 4 levels, gather sampling gradient,
    no cone mask) over the c2 density: the fused shadow op against the op
    sequence it replaces, forward and forward+backward, and the whole
    shadow map (with the reverse resample to the grid)."""
    import numpy as np
    import torch

    from gft import synthetic as syn
    from gft._lib import KernelTimers
    from gft.lighting import SpotLight
    from gft.renderer import Renderer
    from gft.transform import Transform

    grid = syn.WORKLOADS["c2"]["grid"]
    res = [256, 196, 196]
    dev = torch.device("cuda:0")
    gt = syn.grid_transform(grid)
    d = torch.from_numpy(syn.gaussian_blobs(grid, 48, seed=0) * 0.05).to(dev)
    gt.set_data(d)
    # the production light (reconstruct_sequence.py:698-700) in cell units: 300 cells from the
    # grid centre, tilted 40 degrees down, clip 140..460, 50 degree cone
    sl = SpotLight(Transform(translation=[0, 0, 300], parent=Transform(rotation_deg=[-40, 0, 0])), intensity=0.85,
                   cast_shadows=True, shadow_clip=[140, 460], range_scale=0.825, angle_deg=50.0,
                   shadow_resolution=res, cone_mask=False)
    r = Renderer(filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="LINEAR", num_mips=4, mip_bias=0.0,
                 blend_mode="BEER_LAMBERT", sample_gradients=True)
    cam = sl.shadow_cam
    g_frustum = torch.rand([1] + res + [1], device=dev)
    g_grid = torch.rand_like(d)

    def timeit(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(args.steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b))
        return float(np.median(ts))

    def frustum(fused, backward):
        def run():
            x = d.detach().requires_grad_(backward)
            gt.set_data(x)
            t = r.shadow_transmission(gt, cam) if fused else sl._shadow_ops(gt, r)
            if backward:
                t.backward(g_frustum)
        return run

    def whole(fused):
        def run():
            x = d.detach().requires_grad_(True)
            gt.set_data(x)
            sl.render_shadow_map(gt, r, fused=fused).backward(g_grid)
        return run

    out = {}
    for name, fn in (("fused fwd", frustum(True, False)), ("op sequence fwd", frustum(False, False)),
                     ("fused fwd+bwd", frustum(True, True)), ("op sequence fwd+bwd", frustum(False, True)),
                     ("shadow map fused fwd+bwd", whole(True)), ("shadow map op sequence fwd+bwd", whole(False))):
        out[name] = timeit(fn)
    # kernel timers are per thread and autograd runs the backward on its own thread: call the
    # fused forward and the scan backward directly here
    from gft import ops
    M = np.asarray(gt.get_transform_matrix(), dtype=np.float32).T[None]
    V, P, F = r._get_camera_params_batch([cam])
    cams = ops._cams(M, V, P, F)
    with KernelTimers("shadow_fwd,shadow_bwd") as kt:
        for _ in range(args.steps):
            _, saved = ops._shadow_fwd_impl(d, cams, tuple(res), cam.depth_step, ops.FILTER["LINEAR"],
                                            ops.BOUNDARY["BORDER"], ops.MIP["LINEAR"], 4, 0.0,
                                            ops.BLEND["BEER_LAMBERT"], None, save=True)
            ops._shadow_bwd_impl(saved, g_frustum, tuple(res), ops.BOUNDARY["BORDER"], ops.BLEND["BEER_LAMBERT"],
                                 None)
        torch.cuda.synchronize()
    kern = {k: v[1] / max(v[0], 1) for k, v in kt.result.items()}
    cells = int(np.prod(res))
    vox = int(d.numel())
    # compulsory bytes: fwd reads the density and its pyramid once, writes transmittance + prefix sums;
    # bwd reads prefix sums + transmittance gradient, writes the frustum gradient
    fwd_bytes = vox * 4 * (1 + 1.0 / 7) + cells * 8
    bwd_bytes = cells * 12
    gt.set_data(d)
    print(json.dumps({
        "metric": "spot-light shadow map time, fused vs op sequence", "unit": "ms", "n_gpus": 1,
        "steps": args.steps, "higher_is_better": False, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c2 density grid 128x224x128 (x,y,z); shadow map 256x196x196 (D,H,W); "
                               "LINEAR filter, BORDER, LINEAR mips 4 levels, BEER_LAMBERT, gather gradient, "
                               "no cone mask"},
        "ms": out, "speedup_fwd": out["op sequence fwd"] / out["fused fwd"],
        "speedup_fwd_bwd": out["op sequence fwd+bwd"] / out["fused fwd+bwd"],
        "kernels_ms": kern,
        "shadow_fwd": {"voxel_samples_per_s": cells / (kern["shadow_fwd"] * 1e-3),
                       "GBps": fwd_bytes / (kern["shadow_fwd"] * 1e-3) / 1e9, "alg_bytes": int(fwd_bytes)},
        "shadow_bwd": {"GBps": bwd_bytes / (kern["shadow_bwd"] * 1e-3) / 1e9, "alg_bytes": int(bwd_bytes)}}))


def epilogue_bench(args):
    """One JSON line for the loss / optimizer epilogue at c2: render forward,
    AE image loss against targets with the background composited
    (loss_dens_target_raw), render backward, divergence loss of a staggered
    velocity of the grid's shape, and one Adam update of the density and the
    three velocity components. Fused: the loss folded into the render
    backward, the divergence kernel pair and the multi-tensor Adam. Op
    composition: torch autograd for the image loss and the divergence on
    libgft's render kernels, torch.optim.Adam (foreach) for the update."""
    import numpy as np
    import torch

    from gft._lib import KernelTimers
    from gft.grids import VelocityGrid
    from gft.optim import AdamOptimizer, divergence_loss_into
    from gft.pipeline import TransportRenderStep, render_density

    inp = build_inputs("c2")
    dev = torch.device("cuda:0")
    grid = inp["grid"]
    nV, h, w = inp["views"], inp["out_dhw"][1], inp["out_dhw"][2]
    d = torch.from_numpy(inp["d0"]).to(dev)
    vg = VelocityGrid(list(grid), std=0.1, seed=3, device=dev)
    bkg = (0.0, 0.5, 1.0)
    bkg_t = torch.tensor(bkg, device=dev).expand(1, nV, h, w, 3).contiguous()
    target = (torch.rand((1, nV, h, w, 3), device=dev) * 0.5).contiguous()
    scale = 1.0 / (nV * h * w * 3)
    # fused: the fused step's render buffers and folded AE + background loss, on the density itself
    from gft import _lib
    from gft.ops import _np_ptr
    step = TransportRenderStep(grid, inp["M"], inp["V"], inp["P"], inp["F"], inp["out_dhw"], inp["step"],
                               loss="AE", background=bkg, loss_scale=scale)
    params_f = [d.clone(), vg.x.clone(), vg.y.clone(), vg.z.clone()]
    gvel = [torch.empty_like(c) for c in params_f[1:]]
    opt_f = AdamOptimizer(1e-3)
    L, c = step._L, step.cams

    def fused():
        dd, s = params_f[0], _lib.stream_of(params_f[0])
        _lib.check(L.gft_render_density_fwd(s, _lib.ptr(dd), step._shape, _np_ptr(c.M), c.nM, _np_ptr(c.V),
                                            _np_ptr(c.P), _np_ptr(c.F), c.nCam, c.c_out, step.light, step.depth_step,
                                            step.boundary, _lib.ptr(step.render_fwd_ws), step._nfws,
                                            _lib.ptr(step.render_saved), step._nrsaved, _lib.ptr(step.images)))
        _lib.check(L.gft_render_density_bwd_loss(
            s, _lib.ptr(dd), step._shape, _np_ptr(c.M), c.nM, _np_ptr(c.V), _np_ptr(c.P), _np_ptr(c.F), c.nCam,
            c.c_out, step.light, step.depth_step, step.boundary, _lib.ptr(step.images), _lib.ptr(target),
            _lib.ptr(step.background), step.loss_kind, step.loss_scale, _lib.ptr(step.render_saved), step._nrsaved,
            None, 0, _lib.ptr(step.grad_d1), _lib.ptr(step.loss)))
        divergence_loss_into(params_f[1:], (1.0, 1.0, 1.0), 1e-3, grads=gvel)
        opt_f.apply_gradients([(step.grad_d1, params_f[0])] + list(zip(gvel, params_f[1:])))

    params_c = [p.clone().requires_grad_(True) for p in params_f]
    opt_c = torch.optim.Adam(params_c, lr=1e-3)

    def composed():
        opt_c.zero_grad(set_to_none=True)
        img = render_density(params_c[0], inp["M"], inp["V"], inp["P"], inp["F"], inp["out_dhw"], inp["step"])
        comp = img[..., :3] + bkg_t * torch.exp(-img[..., 3:4])
        loss = torch.abs(comp - target).sum() * scale
        v = VelocityGrid(list(grid), x=params_c[1], y=params_c[2], z=params_c[3])
        loss = loss + 1e-3 * (v.divergence((1.0, 1.0, 1.0)) ** 2).sum()
        loss.backward()
        opt_c.step()

    def timeit(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(args.steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b))
        return float(np.median(ts))

    ms_f, ms_c = timeit(fused), timeit(composed)
    with KernelTimers("render_fwd,render_bwd_pixel,render_bwd_gather,divergence,divergence_grad,adam,"
                      "quad_copy") as kt:
        for _ in range(args.steps):
            fused()
        torch.cuda.synchronize()
    kern = {k: v[1] / max(v[0], 1) for k, v in kt.result.items()}
    nparam = sum(p.numel() for p in params_f)
    print(json.dumps({
        "metric": "render + loss + divergence + Adam epilogue time, fused vs op composition", "unit": "ms",
        "n_gpus": 1, "steps": args.steps, "higher_is_better": False, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c2: grid 128x224x128, 5 views 256x456, AE loss with background, divergence "
                               "loss of a staggered velocity, Adam over %d parameters" % nparam},
        "ms": {"fused": ms_f, "op composition": ms_c}, "speedup": ms_c / ms_f, "kernels_ms": kern,
        "adam": {"GBps": nparam * 28 / (kern["adam"] * 1e-3) / 1e9 if "adam" in kern else None,
                 "alg_bytes": nparam * 28}}))


def batched_bench(args):
    """One JSON line for the batched-uneven workload c4 (see the module
    docstring); the replays are checked against the eager results."""
    import numpy as np
    import torch

    from gft import synthetic as syn
    from gft.camera import camera_params_batch, model_matrices
    from gft.pipeline import TransportRenderStep, _Cameras, render_density_forward

    dev = torch.device("cuda:0")
    recs = syn.batched_uneven(32, seed=7)
    steps, inputs = [], []
    for r in recs:
        grid = r["grid"]
        cams, S = syn.arc_cameras(grid, r["image"], r["views"])
        M = model_matrices([syn.grid_transform(grid)])
        V, P, F = camera_params_batch(cams)
        out_dhw = (S,) + tuple(r["image"])
        depth_step = float(cams[0].depth_step)
        d0 = torch.from_numpy(syn.gaussian_blobs(grid, r["blobs"], seed=r["seed"])).to(dev)
        vel = torch.from_numpy(syn.curl_noise_velocity(grid, seed=r["seed"] + 1, max_speed=r["max_speed"])).to(dev)
        dt = torch.from_numpy(syn.gaussian_blobs(grid, r["blobs"], seed=r["seed"] + 2)).to(dev)
        tgt = render_density_forward(dt, _Cameras(M, V, P, F, out_dhw), depth_step)[..., :3].contiguous()
        steps.append(TransportRenderStep(grid, M, V, P, F, out_dhw, depth_step))
        inputs.append((d0, vel, tgt))
    print("batched: %d reconstructions built" % len(steps), file=sys.stderr, flush=True)
    nvs = sum(st.voxel_samples() for st in steps)

    def run_all():
        for st, ins in zip(steps, inputs):
            st.run(*ins)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(args.steps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / args.steps

    ms_eager = timed(run_all)
    eager = [(st.loss.clone(), st.images.clone(), st.grad_density.clone()) for st in steps]

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        run_all()
    torch.cuda.current_stream().wait_stream(side)
    g1 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g1):
        run_all()
    ms_graph = timed(g1.replay)

    branches = [torch.cuda.Stream() for _ in steps]
    for s_, st, ins in zip(branches, steps, inputs):  # warm each branch stream eagerly
        s_.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s_):
            st.run(*ins)
        torch.cuda.current_stream().wait_stream(s_)
    torch.cuda.synchronize()
    gb = torch.cuda.CUDAGraph()
    with torch.cuda.graph(gb):
        cap = torch.cuda.current_stream()
        for s_, st, ins in zip(branches, steps, inputs):
            s_.wait_stream(cap)
            with torch.cuda.stream(s_):
                st.run(*ins)
        for s_ in branches:
            cap.wait_stream(s_)
    ms_branches = timed(gb.replay)
    worst_loss, images_equal, worst_gd = 0.0, True, 0.0
    for st, (lo, img, gd) in zip(steps, eager):
        worst_loss = max(worst_loss, abs(st.loss.item() - lo.item()) / abs(lo.item()))
        images_equal &= bool(torch.equal(st.images, img))
        worst_gd = max(worst_gd, float((st.grad_density - gd).abs().max() / gd.abs().max()))
    best = min(ms_eager, ms_graph, ms_branches)
    sizes = {}
    for r in recs:
        key = "x".join(str(x) for x in r["grid"][::-1])
        sizes[key] = sizes.get(key, 0) + 1
    fused = {"voxel_samples_per_batch": nvs, "ms_per_batch": best, "value": nvs / (best * 1e-3),
             "ms": {"eager, one stream": ms_eager, "one graph, one stream": ms_graph,
                    "one graph, a branch per reconstruction": ms_branches},
             "replay_vs_eager": {"max_rel_loss": worst_loss, "images_equal": images_equal,
                                 "max_norm_grad_density": worst_gd}}
    del steps, inputs, eager, g1, gb
    torch.cuda.empty_cache()

    # the batched mipmapped density-reconstruction step: every reconstruction in one table
    from gft._lib import KernelTimers
    from gft.batch import BatchedRenderStep
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from test_batch_gpu import c4_batch
    brecs, dens, tdens = c4_batch(32, seed=7)
    print("batched: mipmapped step built", file=sys.stderr, flush=True)
    tb = BatchedRenderStep(brecs, num_mipmaps=3)
    tb.bind([torch.from_numpy(d).to(dev) for d in tdens],
            [torch.zeros((r["V"].shape[0],) + tuple(r["out_dhw"][1:]) + (3,), device=dev) for r in brecs])
    tb.run()
    targets = [i[..., :3].contiguous() for i in tb.images]
    b = BatchedRenderStep(brecs, num_mipmaps=3)
    b.bind([torch.from_numpy(d).to(dev) for d in dens], targets)
    ms_default = b.time_ms(args.steps)
    tuned = b.autotune(reps=max(3, args.steps // 2))
    ms_tuned = b.time_ms(args.steps)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        b.run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    gm = torch.cuda.CUDAGraph()
    with torch.cuda.graph(gm):
        b.run()
    ms_mgraph = timed(gm.replay)
    # each reconstruction as a batch of its own, one after another (same kernels and bricks)
    singles = []
    for k, r in enumerate(brecs):
        r1 = dict(r, brick=b.bricks[k])
        s1 = BatchedRenderStep([r1], num_mipmaps=3)
        s1.bind([torch.from_numpy(dens[k]).to(dev)], [targets[k]])
        singles.append(s1)

    def run_singles():
        for s1 in singles:
            s1.run()
    ms_singles = timed(run_singles)
    with KernelTimers("batch_levels,batch_fwd,batch_gather") as kt:
        for _ in range(args.steps):
            b.run()
        torch.cuda.synchronize()
    kern = {k: v[1] / max(v[0], 1) for k, v in kt.result.items()}
    mvs = b.voxel_samples()
    mbest = min(ms_tuned, ms_mgraph)
    print(json.dumps({
        "metric": "voxel-samples/s fwd+bwd of 32 uneven reconstructions (batched mipmapped step)",
        "value": mvs / (mbest * 1e-3), "unit": "voxel-samples/s", "n_gpus": 1, "steps": args.steps,
        "ms_per_batch": mbest, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c4: 32 reconstructions, grids (x,y,z) %s, 1-8 views each, 120x214 to 256x456 px; "
                               "per reconstruction the non-fused ambient render (TRANSFORM_LINDEPTH resample with "
                               "LINEAR mips, LoD 0-3 per sample, BEER_LAMBERT blend), L2 loss, exact density "
                               "gradient" % sizes,
                   "views": [r["views"] for r in recs], "voxel_samples_per_batch": mvs},
        "ms": {"batched, default bricks": ms_default, "batched, autotuned": ms_tuned,
               "batched, autotuned, graph replay": ms_mgraph,
               "each reconstruction alone, one after another": ms_singles},
        "kernel_ms": kern,
        "autotune": {"x".join(str(x) for x in g[::-1]): {"brick": v[0], "ms_per_candidate": v[1]}
                     for g, v in tuned.items()},
        "fused_transport_step_graph": fused}))


def precision_bench(args):
    """One JSON line for the bfloat16-storage step at c5 (see the module
    docstring). Progress goes to stderr: each FP64 oracle step takes tens of
    seconds on the host cores."""
    import numpy as np
    import torch

    from gft.pipeline import TransportRenderStep, _Cameras, render_density_forward
    from oracle import oracle as orc

    def say(msg):
        print("precision: " + msg, file=sys.stderr, flush=True)

    inp = build_inputs("c5")
    say("inputs built")
    dev = torch.device("cuda:0")
    d32 = torch.from_numpy(inp["d0"]).to(dev)
    v32 = torch.from_numpy(inp["vel"]).to(dev)
    d16, v16 = d32.to(torch.bfloat16), v32.to(torch.bfloat16)
    cams = _Cameras(inp["M"], inp["V"], inp["P"], inp["F"], inp["out_dhw"])
    target = render_density_forward(torch.from_numpy(inp["d_target"]).to(dev), cams, inp["step"])[..., :3]
    target = target.contiguous()
    tgt64 = target.cpu().numpy().astype(np.float64)
    rounded = (d16.float().cpu().numpy().astype(np.float64), v16.float().cpu().numpy().astype(np.float64))
    exact = (inp["d0"].astype(np.float64), inp["vel"].astype(np.float64))
    cam_args = (inp["grid"], inp["M"], inp["V"], inp["P"], inp["F"], inp["out_dhw"], inp["step"])

    def timed(step, d, v):
        for _ in range(args.warmup):
            step.run(d, v, target)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        for _ in range(args.steps):
            step.run(d, v, target)
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / args.steps

    def nerr(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))

    def frac(a, b, tol=1e-3):
        return float(np.mean(np.abs(a - b) > tol * np.abs(b).max()))

    res = {}
    # the BORDER render gradient reduced two ways on the same render (the density sampled at float32):
    # the deterministic voxel-block gather and the reverse walk with float32 atomics; and the headline
    # variant whose render samples the advected density stored in bfloat16
    variants = (("bf16_render_gather", "BORDER", "gather", "bfloat16", "deterministic segmented gather"),
                ("fp32_render_gather", "BORDER", "gather", "float32", "deterministic segmented gather"),
                ("fp32_render_scatter", "BORDER", "scatter", "float32", "float32-atomic scatter (reverse walk)"))
    oracle_cache = {}
    for key, boundary, reduction, rstore, desc in variants:
        step16 = TransportRenderStep(*cam_args, boundary=boundary, storage="bfloat16", reduction=reduction,
                                     render_storage=rstore)
        step32 = TransportRenderStep(*cam_args, boundary=boundary, reduction=reduction)
        ms16, ms32 = timed(step16, d16, v16), timed(step32, d32, v32)
        loss, gd, gv = step16.run(d16, v16, target)
        torch.cuda.synchronize()
        got = dict(loss=float(loss.item()), img=step16.images.cpu().numpy(), gd=gd.cpu().numpy(),
                   gv=gv.cpu().numpy())
        say("%s timed: %.3f ms (bf16 storage), %.3f ms (fp32)" % (key, ms16, ms32))
        errs = {}
        for name, (d64, v64) in (("vs_fp64_on_bf16_values", rounded), ("vs_fp64_on_unrounded_inputs", exact)):
            # on the bf16 values the oracle stores the advected density as the step does; on the
            # unrounded inputs it is the FP64 step itself
            rs = rstore if name == "vs_fp64_on_bf16_values" else "float32"
            ck = (name, rs)
            if ck not in oracle_cache:
                oracle_cache[ck] = orc.step_fwd_bwd(d64, v64, tgt64, inp["M"], inp["V"], inp["P"], inp["F"],
                                                    inp["out_dhw"], inp["step"], boundary_render=boundary,
                                                    render_storage=rs)
                say("oracle %s %s done" % ck)
            r = oracle_cache[ck]
            errs[name] = {"images": nerr(got["img"], r["img"]), "loss": abs(got["loss"] - r["loss"]) / abs(r["loss"]),
                          "grad_density": nerr(got["gd"], r["gd"]), "grad_velocity": nerr(got["gv"], r["gv"]),
                          "grad_density_frac_beyond_1e-3": frac(got["gd"], r["gd"]),
                          "grad_velocity_frac_beyond_1e-3": frac(got["gv"], r["gv"])}
        res[key] = {"boundary": boundary, "reduction": desc, "render_storage": rstore, "ms_bf16_storage": ms16,
                    "ms_fp32_storage": ms32, "errors": errs}
        del step16, step32
        torch.cuda.empty_cache()
    nvs = TransportRenderStep(*cam_args, storage="bfloat16").voxel_samples()
    ms = res["bf16_render_gather"]["ms_bf16_storage"]
    print(json.dumps({
        "metric": "voxel-samples/s fwd+bwd, density and velocity stored in bfloat16", "value": nvs / (ms * 1e-3),
        "unit": "voxel-samples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "dtype": "bf16 storage, f32 arithmetic", "data": "synthetic",
        "config": {"workload": "c5: grid 200x356x200 (x,y,z), 8 views 540x960 px, %d samples/ray, MacCormack+"
                               "MC_SMOOTH, L2 loss" % inp["S"], "voxel_samples_per_step": nvs},
        "variants": res}))


def lit_bench(args):
    """One JSON line for the production-lit c2 step (see the module docstring)."""
    import numpy as np
    import torch

    from gft import synthetic as syn
    from gft._lib import KernelTimers
    from gft.lighting import Light, SpotLight
    from gft.pipeline import _Cameras, advect, render_lit_forward, render_lit_saved_bytes
    from gft.renderer import Renderer
    from gft.transform import Transform

    inp = build_inputs("c2")
    dev = torch.device("cuda:0")
    grid = inp["grid"]
    cams, _ = syn.arc_cameras(grid, syn.WORKLOADS["c2"]["image"], syn.WORKLOADS["c2"]["views"])
    d0 = torch.from_numpy(inp["d0"]).to(dev)
    vel = torch.from_numpy(inp["vel"]).to(dev)
    # the production light (reconstruct_sequence.py:698-704) in cell units: 300 cells from the grid
    # centre, tilted 40 degrees down, clip 140..460, 50 degree cone; range_scale 0.825 per reference
    # world unit, with the light ~2 units (300 cells here) away, is 0.0055 per cell
    spot = SpotLight(Transform(translation=[0, 0, 300], parent=Transform(rotation_deg=[-40, 0, 0])), intensity=0.85,
                     cast_shadows=True, shadow_clip=[140, 460], range_scale=0.825 / 150.0, angle_deg=50.0,
                     shadow_resolution=[256, 196, 196], cone_mask=False)
    lights = [spot, Light(intensity=0.64)]

    def renderer(fused_lit):
        return Renderer(filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="LINEAR", num_mips=4, mip_bias=0.0,
                        blend_mode="BEER_LAMBERT", sample_gradients=True, grad_mode="EXACT", fused=True,
                        fused_lit=fused_lit)

    gt = syn.grid_transform(grid)
    gt.set_data(torch.from_numpy(inp["d_target"]).to(dev))
    with torch.no_grad():
        target = torch.cat(renderer(True).render_density(gt, lights, cams, cut_alpha=True), 0)[None].contiguous()

    def step_fn(r):
        def run():
            d = d0.detach().requires_grad_(True)
            v = vel.detach().requires_grad_(True)
            d1 = advect(d, v, 1.0, 2, "MC_SMOOTH")
            gt.set_data(d1)
            img = torch.cat(r.render_density(gt, lights, cams, cut_alpha=True), 0)[None]
            loss = torch.mean((img - target) ** 2)
            loss.backward()
            return loss, d.grad, v.grad
        return run

    def timeit(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(args.steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            torch.cuda.synchronize()
            ts.append(a.elapsed_time(b))
        return float(np.median(ts))

    res = {}
    for name, fused_lit in (("fused_lit", True), ("generic", False)):
        fn = step_fn(renderer(fused_lit))
        res[name] = {"ms_per_step": timeit(fn)}
        loss, gd, gv = fn()
        torch.cuda.synchronize()
        res[name]["loss"] = float(loss)
        res[name]["gd"], res[name]["gv"] = gd.clone(), gv.clone()
    gdiff = float((res["fused_lit"]["gd"] - res["generic"]["gd"]).abs().max() / res["generic"]["gd"].abs().max())
    vdiff = float((res["fused_lit"]["gv"] - res["generic"]["gv"]).abs().max() / res["generic"]["gv"].abs().max())
    for r_ in res.values():
        del r_["gd"], r_["gv"]
    # the lit render kernels alone, called directly (kernel timers are per thread and autograd runs
    # the backward on its own thread)
    import ctypes
    from gft import _lib
    from gft.ops import _np_ptr
    r = renderer(True)
    gt.set_data(d0)
    lam = (spot.light_factor(gt, r) * 0.85).detach().contiguous()
    rc = _Cameras(inp["M"], inp["V"], inp["P"], inp["F"], inp["out_dhw"])
    saved = torch.empty((render_lit_saved_bytes(d0.shape, rc),), dtype=torch.uint8, device=dev)
    gd, gl, loss = torch.empty_like(d0), torch.empty_like(d0), torch.zeros((1,), device=dev)
    lc = (ctypes.c_float * 3)(1.0, 1.0, 1.0)
    tgt = target[0].contiguous()
    with KernelTimers("*") as kt:
        for _ in range(args.steps):
            img = render_lit_forward(d0, lam, rc, inp["step"], (1.0, 1.0, 1.0), 0.64, saved=saved)
            _lib.check(_lib.lib().gft_render_density_lit_bwd_loss(
                _lib.stream_of(d0), _lib.ptr(d0), _lib.ptr(lam), _lib.i64_array(d0.shape), _np_ptr(rc.M), rc.nM,
                _np_ptr(rc.V), _np_ptr(rc.P), _np_ptr(rc.F), rc.nCam, rc.c_out, lc, 0.64, float(inp["step"]), 0,
                _lib.ptr(img), _lib.ptr(tgt), None, 0, 1.0 / tgt.numel(), _lib.ptr(saved), saved.numel(),
                _lib.ptr(gd), _lib.ptr(gl), _lib.ptr(loss)))
        torch.cuda.synchronize()
    kern = {k: v[1] / max(v[0], 1) for k, v in kt.result.items()}
    print(json.dumps({
        "metric": "ms per fwd+bwd step under the production spot + ambient light (lower is better)",
        "value": res["fused_lit"]["ms_per_step"], "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "higher_is_better": False, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c2 grid 128x224x128, 5 views 256x456, MacCormack+MC_SMOOTH, spot light with a "
                               "256x196x196 shadow map (LINEAR mips, 4 levels, gather gradient) + ambient 0.64, "
                               "L2 loss, exact render gradient"},
        "fused_lit": res["fused_lit"], "generic": res["generic"],
        "speedup": res["generic"]["ms_per_step"] / res["fused_lit"]["ms_per_step"],
        "grad_density_diff": gdiff, "grad_velocity_diff": vdiff, "lit_kernels_ms": kern}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--boundary", default="BORDER", choices=["BORDER", "CLAMP", "WRAP"],
                    help="render boundary of the fused step (the reference renderer's default is BORDER)")
    ap.add_argument("--reduction", default="gather", choices=["gather", "scatter"],
                    help="BORDER render backward: deterministic gather over saved per-sample values, or the "
                         "recompute walk with float atomics")
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-kernel", action="store_true")
    ap.add_argument("--chain", action="store_true")
    ap.add_argument("--ops", action="store_true")
    ap.add_argument("--shadow", action="store_true")
    ap.add_argument("--epilogue", action="store_true")
    ap.add_argument("--precision", action="store_true")
    ap.add_argument("--batched", action="store_true")
    ap.add_argument("--lit", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: libgft has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.copy_kernel:
        copy_kernel()
        return
    if args.chain:
        chain_bench(args)
        return
    if args.ops:
        ops_bench(args)
        return
    if args.shadow:
        shadow_bench(args)
        return
    if args.epilogue:
        epilogue_bench(args)
        return
    if args.precision:
        precision_bench(args)
        return
    if args.batched:
        batched_bench(args)
        return
    if args.lit:
        lit_bench(args)
        return

    import numpy as np
    import torch

    from gft._lib import KernelTimers
    from gft.pipeline import TransportRenderStep, _Cameras, render_density_forward

    inp = build_inputs(args.workload)
    dev = torch.device("cuda:0")
    d0 = torch.from_numpy(inp["d0"]).to(dev)
    vel = torch.from_numpy(inp["vel"]).to(dev)
    step = TransportRenderStep(inp["grid"], inp["M"], inp["V"], inp["P"], inp["F"], inp["out_dhw"], inp["step"],
                               boundary=args.boundary, reduction=args.reduction)
    # targets: render of the second density through the same kernels
    cams = _Cameras(inp["M"], inp["V"], inp["P"], inp["F"], inp["out_dhw"])
    tgt_img = render_density_forward(torch.from_numpy(inp["d_target"]).to(dev), cams, inp["step"])
    target = tgt_img[..., :3].contiguous()
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        step.run(d0, vel, target)
    torch.cuda.synchronize()

    # timed region: exactly K fused steps, inputs resident in HBM
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        step.run(d0, vel, target)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = start.elapsed_time(stop) / args.steps

    # second pass over K steps with HIP events around every kernel launch
    with KernelTimers("*") as kt:
        for _ in range(args.steps):
            step.run(d0, vel, target)
        torch.cuda.synchronize()
    kbytes = step.kernel_algorithmic_bytes()
    kernels = {}
    for name, (n, tot) in kt.result.items():
        avg = tot / max(n, 1)
        b = kbytes.get(name, 0)
        kernels[name] = {"launches_per_step": n / args.steps, "avg_ms": avg, "alg_bytes": b,
                         "GBps": b / (avg * 1e-3) / 1e9 if avg > 0 else 0.0,
                         "traffic": measured_traffic(name, args.workload)}
    dominant = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches_per_step"])
    dk = kernels[dominant]
    traffic = measured_traffic(dominant, args.workload)

    nvs = step.voxel_samples()
    value = nvs / (ms_per_step * 1e-3)
    b_alg = sum(step.algorithmic_bytes().values())

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian blobs, curl-noise velocity, arc cameras)",
        "config": {
            "workload": "%s: fused fwd+bwd advect(MacCormack+MC_SMOOTH)->raymarch->L2, grid %dx%dx%d (x,y,z), "
                        "%d views %dx%d px, %d samples/ray" % (args.workload, inp["grid"][2], inp["grid"][1],
                                                               inp["grid"][0], inp["views"], inp["image"][1],
                                                               inp["image"][0], inp["S"]),
            "voxel_samples_per_step": nvs,
            "grad_mode": "EXACT",
            "render_boundary": args.boundary,
            "render_reduction": args.reduction,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": dk["GBps"],
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": dk["GBps"] / HBM_PEAK_GBPS,
            "traffic": traffic,
            "algorithmic_bytes": dk["alg_bytes"],
            "avg_ms": dk["avg_ms"],
        },
        "kernels": kernels,
        "step_algorithmic_bytes": b_alg,
        "step_hbm_alg_GBps": b_alg / (ms_per_step * 1e-3) / 1e9,
        "loss": float(step.loss.item()),
        "wall_s_timed": wall,
    }

    if not args.no_plain and args.boundary == "BORDER":  # the plain composition renders with BORDER
        from gft.plain import PlainTransportRenderStep
        plain = PlainTransportRenderStep(inp["grid"], inp["M"], inp["V"], inp["P"], inp["F"], inp["out_dhw"],
                                         inp["step"])
        plain.run(d0, vel, target)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(args.plain_steps):
            plain.run(d0, vel, target)
        b.record()
        torch.cuda.synchronize()
        ms_plain = a.elapsed_time(b) / args.plain_steps
        b_plain = plain.plain_bytes()
        result["plain_fp32"] = {"ms_per_step": ms_plain, "value": nvs / (ms_plain * 1e-3), "steps": args.plain_steps,
                                "plain_bytes": b_plain, "hbm_plain_GBps": b_plain / (ms_plain * 1e-3) / 1e9,
                                "hbm_alg_GBps": b_alg / (ms_plain * 1e-3) / 1e9,
                                "peak_allocated_GiB": torch.cuda.max_memory_allocated() / 2 ** 30}
        result["vs_plain"] = ms_plain / ms_per_step

    if not args.no_cpu_baseline:
        tgt_np = target.cpu().numpy().astype(np.float64)
        ray = step.n_pixels * step.cams.out_dhw[0]
        result["cpu_baseline"] = cpu_baseline(inp, tgt_np, nvs, ray, args.boundary)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
