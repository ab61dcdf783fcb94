"""The headline workloads at their full sizes.

c2 (128x224x128, 5 views of 256x456, 182 samples per ray): the fused
fwd+bwd step against the FP64 oracle (loss and images 1e-5, gradients 1e-3,
the bar of the small-size tests; the oracle takes a few seconds on the host
cores), the BORDER render backward bit-for-bit repeatable (the advection
adjoint accumulates with atomics, so its output is repeatable to rounding),
and the adjoint identity <A d, g> = <d, A^T g> of the semi-Lagrangian
advection (linear in d).

The velocity gradient is discontinuous where a trace position crosses a cell
face (the trilinear weights' derivative jumps there), so FP32 and FP64
positions on either side of a face give O(1) different gradients for that
cell; it is held to at most 1e-4 of the cells beyond 1e-3.

MC_SMOOTH makes the step's gradient discontinuous in the density: a cell
whose corrected value sits at a near-tie with the min/max of its corners
takes the first-order branch in one arithmetic and the corrected one in the
other, and the gradients of the voxels it reads differ by O(1) there. At c2
the FP32 restatement of the reference's arithmetic and the FP64 oracle
disagree on 20 such decisions, which moves 2070 voxels of the density
gradient beyond 1e-3; the GPU path moves 83 (2.3e-5 of the grid). So the
MacCormack step without clamping is held to 1e-3 everywhere, and with
MC_SMOOTH at most 1e-4 of the voxels may exceed it.

c5 (200x356x200, 8 views of 540x960, 283 samples per ray), FP32 storage:
the fused step repeatable as above and its images equal to the per-op
raymarch kernel's (an independent kernel of the same semantics) within 1e-5.
With the density and velocity stored in bfloat16: the step against the FP64
oracle on the bfloat16 values. Loss, images and the render backward to the
c2 bars; without clamping the density gradient within 1e-3 everywhere. With
MC_SMOOTH at most 1e-3 of the voxels may exceed 1e-3: where the c5 density
underflows to exactly 0 the corrected value ties its corners' extrema
exactly (3.0e5 voxels with float32 inputs, 4.3e5 with bfloat16 ones, none at
c2), so the branch there is decided by rounding; the FP32 restatement of the
reference's arithmetic moves 2.0e-4 of the voxels beyond 1e-3, the GPU 3.4e-4.
"""
import numpy as np
import pytest
import torch

from gft import synthetic as syn
from helpers import norm_err
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _workload(name):
    from gft.camera import camera_params_batch, model_matrices
    w = syn.WORKLOADS[name]
    grid, image = w["grid"], w["image"]
    cams, S = syn.arc_cameras(grid, image, w["views"])
    M = model_matrices([syn.grid_transform(grid)])
    V, P, F = camera_params_batch(cams)
    d0 = syn.gaussian_blobs(grid, w["blobs"], seed=0)
    vel = syn.curl_noise_velocity(grid, seed=1, max_speed=w["max_speed"])
    return dict(grid=grid, M=M, V=V, P=P, F=F, out_dhw=(S,) + tuple(image), step=float(cams[0].depth_step), d0=d0,
                vel=vel, d_target=syn.gaussian_blobs(grid, w["blobs"], seed=2))


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.fixture(scope="module")
def c2():
    return _workload("c2")


def _target(w):
    from gft.pipeline import _Cameras, render_density_forward
    cams = _Cameras(w["M"], w["V"], w["P"], w["F"], w["out_dhw"])
    return render_density_forward(_t(w["d_target"]), cams, w["step"])[..., :3].contiguous()


def _frac_beyond(got, ref, tol=1e-3):
    scale = np.abs(ref).max()
    return float(np.mean(np.abs(got - ref) > tol * scale))


@pytest.mark.parametrize("clamp", ["NONE", "MC_SMOOTH"])
def test_c2_fused_step_matches_fp64_oracle(cuda, c2, clamp):
    from gft.pipeline import TransportRenderStep
    w = c2
    tgt = _target(w)
    step = TransportRenderStep(w["grid"], w["M"], w["V"], w["P"], w["F"], w["out_dhw"], w["step"], clamp=clamp)
    loss, gd, gv = step.run(_t(w["d0"]), _t(w["vel"]), tgt)
    torch.cuda.synchronize()
    r = orc.step_fwd_bwd(w["d0"], w["vel"], tgt.cpu().numpy().astype(np.float64), w["M"], w["V"], w["P"], w["F"],
                         w["out_dhw"], w["step"], clamp=clamp)
    assert abs(loss.item() - r["loss"]) <= 1e-5 * abs(r["loss"])
    assert norm_err(step.images.cpu().numpy(), r["img"]) < 1e-5
    assert norm_err(step.grad_d1.cpu().numpy(), r["gd1"]) < 1e-3  # the render backward alone
    gd, gv = gd.cpu().numpy(), gv.cpu().numpy()
    if clamp == "NONE":
        assert norm_err(gd, r["gd"]) < 1e-3
    else:
        assert _frac_beyond(gd, r["gd"]) <= 1e-4
    assert _frac_beyond(gv, r["gv"]) <= 1e-4


def _repeat(step, d0, vel, tgt):
    runs = []
    for _ in range(2):
        loss, gd, gv = step.run(d0, vel, tgt)
        runs.append((loss.clone(), step.grad_d1.clone(), gd.clone(), gv.clone()))
    torch.cuda.synchronize()
    (l0, r0, gd0, gv0), (l1, r1, gd1, gv1) = runs
    assert torch.equal(l0, l1) and torch.equal(r0, r1)  # loss and render backward: fixed order
    assert norm_err(gd1.cpu().numpy(), gd0.cpu().numpy()) < 1e-6  # advection adjoint: atomics
    assert norm_err(gv1.cpu().numpy(), gv0.cpu().numpy()) < 1e-6
    return runs[0]


def test_c2_border_backward_is_bitwise_repeatable(cuda, c2):
    from gft.pipeline import TransportRenderStep
    w = c2
    step = TransportRenderStep(w["grid"], w["M"], w["V"], w["P"], w["F"], w["out_dhw"], w["step"])
    _repeat(step, _t(w["d0"]), _t(w["vel"]), _target(w))


def test_c2_semi_lagrangian_adjoint_identity(cuda, c2):
    """Order-1 advection without clamping is linear in the advected field: its
    exact adjoint must satisfy <A d, g> = <d, A^T g> (host float64 sums)."""
    from gft.pipeline import advect
    w = c2
    rng = np.random.Generator(np.random.PCG64(11))
    d = rng.uniform(-1, 1, size=(1,) + tuple(w["grid"]) + (1,)).astype(np.float32)
    g = rng.uniform(-1, 1, size=d.shape).astype(np.float32)
    td = _t(d).requires_grad_(True)
    out = advect(td, _t(w["vel"]), 1.0, 1, "NONE")
    out.backward(_t(g))
    lhs = float(np.dot(out.detach().cpu().numpy().astype(np.float64).ravel(), g.astype(np.float64).ravel()))
    rhs = float(np.dot(d.astype(np.float64).ravel(), td.grad.cpu().numpy().astype(np.float64).ravel()))
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + 1.0)


@pytest.fixture(scope="module")
def c5():
    return _workload("c5")


def test_c5_fp32_step_repeatable_and_images_match_raymarch_op(cuda, c5):
    from gft.ops import raymarching_ops
    from gft.pipeline import TransportRenderStep
    w = c5
    tgt = _target(w)
    step = TransportRenderStep(w["grid"], w["M"], w["V"], w["P"], w["F"], w["out_dhw"], w["step"])
    d0, vel = _t(w["d0"]), _t(w["vel"])
    loss, _, gd, gv = _repeat(step, d0, vel, tgt)
    img = step.images.clone()
    assert torch.isfinite(gd).all() and torch.isfinite(gv).all() and loss.item() > 0
    # the images against the per-op raymarch of the materialised [light, density] * step grid of d1
    grid4 = torch.cat([step.d1, step.d1, step.d1, step.d1], dim=-1) * w["step"]
    ref = raymarching_ops.raymarch_grid_transform(grid4, w["M"], w["V"], w["P"], w["F"], w["out_dhw"])
    assert norm_err(img.cpu().numpy(), ref.cpu().numpy()) < 1e-5


@pytest.mark.parametrize("clamp", ["NONE", "MC_SMOOTH"])
def test_c5_bf16_storage_step_matches_fp64_oracle(cuda, c5, clamp):
    from gft.pipeline import TransportRenderStep
    w = c5
    d0 = _t(w["d0"]).to(torch.bfloat16)
    vel = _t(w["vel"]).to(torch.bfloat16)
    tgt = _target(w)
    step = TransportRenderStep(w["grid"], w["M"], w["V"], w["P"], w["F"], w["out_dhw"], w["step"], clamp=clamp,
                               storage="bfloat16")
    loss, gd, gv = step.run(d0, vel, tgt)
    torch.cuda.synchronize()
    r = orc.step_fwd_bwd(d0.float().cpu().numpy().astype(np.float64), vel.float().cpu().numpy().astype(np.float64),
                         tgt.cpu().numpy().astype(np.float64), w["M"], w["V"], w["P"], w["F"], w["out_dhw"], w["step"],
                         clamp=clamp, render_storage="bfloat16")
    assert abs(loss.item() - r["loss"]) <= 1e-5 * abs(r["loss"])
    # the render samples the advected density at bfloat16: the GPU's rounding and the oracle's agree
    # but where FP32 and FP64 values straddle a rounding tie; the images against the oracle render of
    # the GPU's own rounded density
    d1r = step.d1.to(torch.bfloat16).float().cpu().numpy()
    assert np.mean(d1r != r["d1r"]) <= 1e-3
    ref_img = orc.render_density_fwd(d1r, w["M"], w["V"], w["P"], w["F"], w["out_dhw"], w["step"])
    assert norm_err(step.images.cpu().numpy(), ref_img) < 1e-5
    assert norm_err(step.grad_d1.cpu().numpy(), r["gd1"]) < 1e-3
    gd, gv = gd.cpu().numpy(), gv.cpu().numpy()
    if clamp == "NONE":
        assert norm_err(gd, r["gd"]) < 1e-3
    else:
        assert _frac_beyond(gd, r["gd"]) <= 1e-3  # exact MC_SMOOTH ties where the density is 0
    assert _frac_beyond(gv, r["gv"]) <= 1e-4  # face crossings, as at c2
