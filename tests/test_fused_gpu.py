"""Parity of the fused HIP transport/render path with the oracle.

Advection (MacCormack + MC_SMOOTH, semi-Lagrangian, NEGATIVE) forward and
exact adjoint, ambient density render forward/backward, the L2 loss, and the
whole advect -> render -> loss step. Tolerances (normalised max error vs the
FP64 oracle): images 1e-5, advected grids 1e-5 outside clamp-switch voxels,
gradients 1e-3.
"""
import numpy as np
import pytest
import torch

from gft import synthetic as syn
from helpers import norm_err, scene
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _fields(grid, seed=0, speed=1.5, blobs=6):
    d = syn.gaussian_blobs(grid, blobs, seed=seed)
    v = syn.curl_noise_velocity(grid, seed=seed + 1, max_speed=speed, periods=(8, 4, 2))
    return d, v


@pytest.mark.parametrize("order,clamp", [(2, "MC_SMOOTH"), (2, "NONE"), (1, "NONE"), (1, "NEGATIVE"),
                                         (2, "NEGATIVE")])
def test_advect_forward(cuda, order, clamp):
    from gft.pipeline import advect
    grid = (14, 22, 18)
    d, v = _fields(grid, seed=2)
    if clamp == "NEGATIVE":
        d = d - 0.3
    got = advect(torch.from_numpy(d).cuda(), torch.from_numpy(v).cuda(), 1.0, order, clamp).cpu().numpy()
    ref, parts = orc.advect_fwd(d, v, order, clamp, prec=0, return_parts=True)
    diff = np.abs(got - ref) > 1e-5 * np.abs(ref).max()
    # a clamp decision may flip between FP32 implementations on exact ties
    assert diff.mean() < 2e-3
    ref64 = orc.advect_fwd(d, v, order, clamp, prec=1)
    assert norm_err(got[~diff], ref64[~diff]) < 1e-5


@pytest.mark.parametrize("order,clamp", [(2, "MC_SMOOTH"), (1, "NONE"), (2, "NONE")])
def test_advect_backward(cuda, order, clamp):
    from gft.pipeline import advect
    grid = (14, 22, 18)
    d, v = _fields(grid, seed=4)
    rng = np.random.Generator(np.random.PCG64(9))
    g = rng.uniform(-1, 1, size=d.shape).astype(np.float32)
    td = torch.from_numpy(d).cuda().requires_grad_(True)
    tv = torch.from_numpy(v).cuda().requires_grad_(True)
    out = advect(td, tv, 1.0, order, clamp)
    out.backward(torch.from_numpy(g).cuda())
    gd_ref, gv_ref = orc.advect_bwd(d, v, g, order, clamp, grad_mode="EXACT", prec=1)
    assert norm_err(td.grad.cpu().numpy(), gd_ref) < 1e-3
    assert norm_err(tv.grad.cpu().numpy(), gv_ref) < 1e-3


@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP"])
def test_render_density_forward_backward(cuda, boundary):
    """The autograd render op (with CLAMP both directions build the CLAMP
    corner-quad copy in their own workspace)."""
    from gft.pipeline import render_density
    sc = scene(grid=(14, 22, 18), image=(20, 13), views=4, seed=6)
    d = sc["data"]
    td = torch.from_numpy(d).cuda().requires_grad_(True)
    light = (0.9, 1.0, 1.1)
    img = render_density(td, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], light, boundary=boundary)
    ref = orc.render_density_fwd(d, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], light,
                                 boundary=boundary)
    assert norm_err(img.detach().cpu().numpy(), ref) < 1e-5
    rng = np.random.Generator(np.random.PCG64(3))
    g = rng.uniform(-1, 1, size=ref.shape).astype(np.float32)
    img.backward(torch.from_numpy(g).cuda())
    gref = orc.render_density_bwd(d, img.detach().cpu().numpy(), g, sc["M"], sc["V"], sc["P"], sc["F"],
                                  sc["out_dhw"], sc["step"], light, boundary=boundary)
    assert norm_err(td.grad.cpu().numpy(), gref) < 1e-4


@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP", "WRAP"])
def test_fused_step_matches_oracle(cuda, boundary):
    from gft.pipeline import TransportRenderStep
    grid = (16, 26, 18)
    sc = scene(grid=grid, image=(24, 14), views=3, seed=8)
    d, v = _fields(grid, seed=10)
    tgt_d = syn.gaussian_blobs(grid, 6, seed=2)
    tgt = orc.render_density_fwd(tgt_d, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"],
                                 boundary=boundary)[..., :3]
    step = TransportRenderStep(grid, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], boundary=boundary)
    td, tv = torch.from_numpy(d).cuda(), torch.from_numpy(v).cuda()
    loss, gd, gv = step.run(td, tv, torch.from_numpy(tgt.astype(np.float32)).cuda())
    torch.cuda.synchronize()
    r = orc.step_fwd_bwd(d, v, tgt, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"],
                         boundary_render=boundary)
    assert abs(loss.item() - r["loss"]) <= 1e-5 * abs(r["loss"])
    assert norm_err(step.images.cpu().numpy(), r["img"]) < 1e-5
    assert norm_err(gd.cpu().numpy(), r["gd"]) < 1e-3
    assert norm_err(gv.cpu().numpy(), r["gv"]) < 1e-3


def test_fused_step_repeatable_and_graph_capturable(cuda):
    from gft.pipeline import TransportRenderStep
    grid = (16, 26, 18)
    sc = scene(grid=grid, image=(24, 14), views=3, seed=8)
    d, v = _fields(grid, seed=12)
    tgt = torch.zeros((1, 3) + sc["out_dhw"][1:] + (3,), device="cuda")
    step = TransportRenderStep(grid, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"])
    td, tv = torch.from_numpy(d).cuda(), torch.from_numpy(v).cuda()
    step.run(td, tv, tgt)
    torch.cuda.synchronize()
    eager = (step.loss.clone(), step.grad_density.clone(), step.grad_velocity.clone(), step.images.clone())
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step.run(td, tv, tgt)
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step.run(td, tv, tgt)
    for _ in range(3):
        step.grad_density.fill_(123.0)
        step.images.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(step.images, eager[3])
        assert abs(step.loss.item() - eager[0].item()) <= 1e-6 * abs(eager[0].item())
        assert norm_err(step.grad_density.cpu().numpy(), eager[1].cpu().numpy()) < 1e-5
        assert norm_err(step.grad_velocity.cpu().numpy(), eager[2].cpu().numpy()) < 1e-5


def test_border_backward_is_deterministic(cuda):
    from gft.pipeline import TransportRenderStep
    grid = (16, 26, 18)
    sc = scene(grid=grid, image=(24, 14), views=3, seed=8)
    d, v = _fields(grid, seed=14)
    tgt = torch.zeros((1, 3) + sc["out_dhw"][1:] + (3,), device="cuda")
    step = TransportRenderStep(grid, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], boundary="BORDER")
    td, tv = torch.from_numpy(d).cuda(), torch.from_numpy(v).cuda()
    step.forward(td, tv)
    step.backward(td, tv, tgt)
    first = step.grad_d1.clone()
    for _ in range(3):
        # the backward consumes the forward's saved samples: run the pair again
        step.forward(td, tv)
        step.backward(td, tv, tgt)
        assert torch.equal(step.grad_d1, first)


def _render_case(d, sc, light, boundary="BORDER", seed=5):
    """GPU render forward/backward through the autograd op vs the oracle."""
    from gft.pipeline import render_density
    td = torch.from_numpy(d).cuda().requires_grad_(True)
    img = render_density(td, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], light, boundary)
    ref = orc.render_density_fwd(d, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], light,
                                 boundary=boundary)
    rng = np.random.Generator(np.random.PCG64(seed))
    g = rng.uniform(-1, 1, size=ref.shape).astype(np.float32)
    img.backward(torch.from_numpy(g).cuda())
    gref = orc.render_density_bwd(d, img.detach().cpu().numpy(), g, sc["M"], sc["V"], sc["P"], sc["F"],
                                  sc["out_dhw"], sc["step"], light, boundary=boundary)
    return img.detach().cpu().numpy(), ref, td.grad.cpu().numpy(), gref


@pytest.mark.parametrize("views,yaw", [(18, (-60.0, 60.0)), (7, (-180.0, 154.0))])
def test_render_density_many_views_and_full_circle(cuda, views, yaw):
    """18 views cross the 16-view launch chunk (the second chunk adds to the
    first's gradient); a full circle of cameras sees the grid from behind and
    from the sides (negative and near-zero ray step components)."""
    sc = scene(grid=(13, 21, 17), image=(19, 12), views=views, seed=9, yaw_range=yaw)
    img, ref, gd, gref = _render_case(sc["data"], sc, (1.0, 0.8, 1.2))
    assert norm_err(img, ref) < 1e-5
    assert norm_err(gd, gref) < 1e-4


@pytest.mark.parametrize("yaw", [0.0, 90.0, 33.0])
def test_render_density_orthographic_axis_aligned_rays(cuda, yaw):
    """Orthographic cameras with rays along a grid axis (yaw 0 and 90): two
    step components are exactly zero, so the gather's ray/box cut runs on
    infinite reciprocals."""
    from gft.camera import camera_params_batch, model_matrices
    from helpers import ortho_camera
    grid = (10, 12, 9)
    cam = ortho_camera(grid, (12, 9), 11, yaw=yaw)
    V, P, F = camera_params_batch([cam])
    sc = dict(M=model_matrices([syn.grid_transform(grid)]), V=V, P=P, F=F, out_dhw=(11, 12, 9), step=cam.depth_step)
    d = syn.gaussian_blobs(grid, 4, seed=3).astype(np.float32)
    img, ref, gd, gref = _render_case(d, sc, (1.0, 1.0, 1.0), seed=8)
    assert np.abs(ref).max() > 0.05
    assert norm_err(img, ref) < 1e-5
    assert norm_err(gd, gref) < 1e-4


def test_render_density_saved_buffer_is_checked(cuda):
    """The BORDER backward needs the forward's saved buffer: a missing or short
    buffer is a workspace error, not a fault."""
    import ctypes
    from gft import _lib
    from gft.ops import BOUNDARY, _np_ptr
    from gft.pipeline import _Cameras, render_density_forward, render_saved_bytes
    sc = scene(grid=(8, 10, 9), image=(10, 7), views=2, seed=4)
    cams = _Cameras(sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"])
    d = torch.from_numpy(sc["data"]).cuda()
    n = render_saved_bytes(d.shape, cams)
    assert n > 0 and render_saved_bytes(d.shape, cams, "CLAMP") == 0
    saved = torch.empty((n,), dtype=torch.uint8, device="cuda")
    img = render_density_forward(d, cams, sc["step"], saved=saved)
    L = _lib.lib()
    lt = (ctypes.c_float * 3)(1.0, 1.0, 1.0)
    gd = torch.empty_like(d)
    g = torch.ones_like(img)
    args = lambda sp, sn: (_lib.stream_of(d), _lib.ptr(d), _lib.i64_array(d.shape), _np_ptr(cams.M), cams.nM,
                           _np_ptr(cams.V), _np_ptr(cams.P), _np_ptr(cams.F), cams.nCam, cams.c_out, lt,
                           float(sc["step"]), BOUNDARY["BORDER"], _lib.ptr(img), _lib.ptr(g), sp, sn, None, 0,
                           _lib.ptr(gd))
    assert L.gft_render_density_bwd(*args(None, 0)) == _lib.GFT_ERR_WORKSPACE
    assert L.gft_render_density_bwd(*args(_lib.ptr(saved), n - 1)) == _lib.GFT_ERR_WORKSPACE
    assert L.gft_render_density_bwd(*args(_lib.ptr(saved), n)) == 0
    torch.cuda.synchronize()
    assert torch.isfinite(gd).all()


@pytest.mark.parametrize("branches", [False, True])
def test_uneven_reconstructions_in_one_graph(cuda, branches):
    """Independent reconstructions of uneven grid sizes and view counts (the
    batched-uneven workload at test size) captured together in one graph, in
    sequence on the capture stream or each as a branch on its own stream
    (forked from and joined to the capture stream): every replay reproduces
    each reconstruction's eager loss, images and gradients, and each matches
    the FP64 oracle."""
    from gft.pipeline import TransportRenderStep
    cases = [((16, 26, 18), 3, (24, 14)), ((12, 20, 14), 1, (18, 11)), ((20, 16, 24), 4, (16, 20)),
             ((10, 10, 10), 2, (12, 12))]
    steps, inputs, eager = [], [], []
    for k, (grid, views, image) in enumerate(cases):
        sc = scene(grid=grid, image=image, views=views, seed=30 + k)
        d, v = _fields(grid, seed=40 + k)
        tgt = orc.render_density_fwd(syn.gaussian_blobs(grid, 4, seed=50 + k), sc["M"], sc["V"], sc["P"], sc["F"],
                                     sc["out_dhw"], sc["step"])[..., :3]
        st = TransportRenderStep(grid, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"])
        ins = (torch.from_numpy(d).cuda(), torch.from_numpy(v).cuda(), torch.from_numpy(tgt.astype(np.float32)).cuda())
        st.run(*ins)
        torch.cuda.synchronize()
        r = orc.step_fwd_bwd(d, v, tgt, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"])
        assert abs(st.loss.item() - r["loss"]) <= 1e-5 * abs(r["loss"])
        assert norm_err(st.grad_density.cpu().numpy(), r["gd"]) < 1e-3
        steps.append(st)
        inputs.append(ins)
        eager.append((st.loss.clone(), st.images.clone(), st.grad_density.clone(), st.grad_velocity.clone()))
    streams = [torch.cuda.Stream() for _ in steps] if branches else [torch.cuda.Stream()] * len(steps)
    for s, st, ins in zip(streams, steps, inputs):  # warm-up on the streams the capture uses
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            st.run(*ins)
        torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        if branches:
            cap = torch.cuda.current_stream()
            for s, st, ins in zip(streams, steps, inputs):
                s.wait_stream(cap)
                with torch.cuda.stream(s):
                    st.run(*ins)
            for s in streams:
                cap.wait_stream(s)
        else:
            for st, ins in zip(steps, inputs):
                st.run(*ins)
    for _ in range(2):
        for st in steps:
            st.grad_density.fill_(7.0)
            st.images.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for st, (lo, img, gd, gv) in zip(steps, eager):
            assert torch.equal(st.images, img)
            assert abs(st.loss.item() - lo.item()) <= 1e-6 * abs(lo.item())
            assert norm_err(st.grad_density.cpu().numpy(), gd.cpu().numpy()) < 1e-5
            assert norm_err(st.grad_velocity.cpu().numpy(), gv.cpu().numpy()) < 1e-5


@pytest.mark.parametrize("loss", ["SE", "AE"])
def test_border_scatter_reduction_matches_gather_and_oracle(cuda, loss):
    """The BORDER render backward reduced two ways: the deterministic
    voxel-block gather over the forward's saved per-sample values, and the
    reverse walk that recomputes the transmittance and scatters with float
    atomics (nothing saved). Both against the FP64 oracle (loss 1e-5,
    gradients 1e-3) and against each other (1e-5); the gather repeats bit for
    bit."""
    from gft.pipeline import TransportRenderStep
    grid = (18, 26, 22)
    sc = scene(grid=grid, image=(24, 18), views=3, seed=90)
    d, v = _fields(grid, seed=91)
    tgt = orc.render_density_fwd(syn.gaussian_blobs(grid, 5, seed=92), sc["M"], sc["V"], sc["P"], sc["F"],
                                 sc["out_dhw"], sc["step"])[..., :3]
    ins = (torch.from_numpy(d).cuda(), torch.from_numpy(v).cuda(), torch.from_numpy(tgt.astype(np.float32)).cuda())
    out = {}
    for red in ("gather", "scatter"):
        st = TransportRenderStep(grid, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], loss=loss,
                                 reduction=red)
        st.run(*ins)
        torch.cuda.synchronize()
        out[red] = (st.loss.item(), st.grad_d1.cpu().numpy(), st.grad_density.cpu().numpy())
    r = orc.step_fwd_bwd(d, v, tgt, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], loss_kind=loss)
    for red in out:
        assert abs(out[red][0] - r["loss"]) <= 1e-5 * abs(r["loss"]), red
        assert norm_err(out[red][1], r["gd1"]) < 1e-3, red
        assert norm_err(out[red][2], r["gd"]) < 1e-3, red
    assert norm_err(out["scatter"][1], out["gather"][1]) < 1e-5
