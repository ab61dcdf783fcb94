"""Density and velocity stored in bfloat16, sampled and accumulated in
float32 (the precision variant of the transport step).

The kernels widen every bfloat16 value to float32 on load, so a float32
tensor holding the same (bfloat16-representable) values must give the same
advected grid bit for bit. The BORDER render of the step samples the advected
density stored in bfloat16 too: its images equal the float32 render of the
rounded density bit for bit. CLAMP renders float32, so there the images and
the loss equal the float32-storage step, the gradients to rounding (the
advection adjoint and the CLAMP render backward accumulate with atomics).
Against the FP64 oracle on the bfloat16 values (with the advected density
rounded to bfloat16 for the BORDER render): advected grids 1e-5 outside
clamp-switch voxels, images 1e-5, gradients 1e-3 (normalised max error).
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


def _bf16(a):
    """(bfloat16 CUDA tensor, its exact values as float64)"""
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda().to(torch.bfloat16)
    return t, t.float().cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("order,clamp", [(2, "MC_SMOOTH"), (2, "NONE"), (1, "NONE")])
def test_advect_bf16_storage(cuda, order, clamp):
    from gft.pipeline import advect, advect_backward, advect_forward
    grid = (14, 22, 18)
    d, v = _fields(grid, seed=2)
    td, d64 = _bf16(d)
    tv, v64 = _bf16(v)
    out, saved = advect_forward(td, tv, 1.0, order, clamp)
    out32, saved32 = advect_forward(td.float(), tv.float(), 1.0, order, clamp)
    assert out.dtype == torch.float32 and torch.equal(out, out32)
    got = out.cpu().numpy()
    ref = orc.advect_fwd(d64, v64, order, clamp, prec=0)
    diff = np.abs(got - ref) > 1e-5 * np.abs(ref).max()
    assert diff.mean() < 2e-3  # clamp decisions may flip between FP32 implementations on exact ties
    assert norm_err(got[~diff], orc.advect_fwd(d64, v64, order, clamp, prec=1)[~diff]) < 1e-5
    rng = np.random.Generator(np.random.PCG64(9))
    g = torch.from_numpy(rng.uniform(-1, 1, size=d.shape).astype(np.float32)).cuda()
    gd, gv = advect_backward(td, tv, saved, g, 1.0, order, clamp)
    gd32, gv32 = advect_backward(td.float(), tv.float(), saved32, g, 1.0, order, clamp)
    assert gd.dtype == torch.float32 and gv.dtype == torch.float32
    assert norm_err(gd.cpu().numpy(), gd32.cpu().numpy()) < 1e-6
    assert norm_err(gv.cpu().numpy(), gv32.cpu().numpy()) < 1e-6
    gd_ref, gv_ref = orc.advect_bwd(d64, v64, g.cpu().numpy(), order, clamp, grad_mode="EXACT", prec=1)
    assert norm_err(gd.cpu().numpy(), gd_ref) < 1e-3
    assert norm_err(gv.cpu().numpy(), gv_ref) < 1e-3
    # autograd: float32 result, gradients cast to the inputs' bfloat16
    a = td.clone().requires_grad_(True)
    b = tv.clone().requires_grad_(True)
    advect(a, b, 1.0, order, clamp).backward(g)
    assert a.grad.dtype == torch.bfloat16 and b.grad.dtype == torch.bfloat16
    assert norm_err(a.grad.float().cpu().numpy(), gd_ref) < 1e-2  # bfloat16 rounding of the gradient


def test_advect_bf16_rejects_mixed_storage(cuda):
    from gft import _lib
    from gft.pipeline import advect
    d, v = _fields((6, 7, 8), seed=3)
    td, _ = _bf16(d)
    with pytest.raises(_lib.InvalidArgumentError):
        advect(td, torch.from_numpy(v).cuda())


@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP"])
def test_fused_step_bf16_storage(cuda, boundary):
    from gft import _lib
    from gft.pipeline import TransportRenderStep
    grid = (16, 26, 18)
    sc = scene(grid=grid, image=(24, 14), views=3, seed=8)
    d, v = _fields(grid, seed=10)
    td, d64 = _bf16(d)
    tv, v64 = _bf16(v)
    tgt = orc.render_density_fwd(syn.gaussian_blobs(grid, 5, seed=3), sc["M"], sc["V"], sc["P"], sc["F"],
                                 sc["out_dhw"], sc["step"], boundary=boundary)[..., :3].astype(np.float32)
    ttgt = torch.from_numpy(tgt).cuda()
    args = (grid, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"])
    step = TransportRenderStep(*args, boundary=boundary, storage="bfloat16")
    step32 = TransportRenderStep(*args, boundary=boundary)
    lo, gd, gv = step.run(td, tv, ttgt)
    lo32, gd32, gv32 = step32.run(td.float(), tv.float(), ttgt)
    torch.cuda.synchronize()
    assert torch.equal(step.d1, step32.d1)
    render_storage = "float32"
    if boundary == "BORDER":  # the render samples the advected density at bfloat16
        from gft.pipeline import _Cameras, render_density_forward
        cams = _Cameras(sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"])
        d1r = step.d1.to(torch.bfloat16).float()
        assert not torch.equal(d1r, step.d1)
        assert torch.equal(step.images, render_density_forward(d1r, cams, sc["step"]))
        render_storage = "bfloat16"
        r = orc.step_fwd_bwd(d64, v64, tgt.astype(np.float64), sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"],
                             sc["step"], boundary_render=boundary, render_storage="bfloat16")
        # the GPU's and the oracle's bfloat16 roundings of the advected density agree but at ties
        assert np.mean(d1r.cpu().numpy() != r["d1r"]) <= 1e-3
        ref_img = orc.render_density_fwd(d1r.cpu().numpy(), sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"],
                                         sc["step"])
    else:  # CLAMP renders float32: the float32-storage step's images and loss, the backward to rounding
        assert torch.equal(step.images, step32.images)
        assert abs(lo.item() - lo32.item()) <= 1e-6 * abs(lo32.item())
        assert norm_err(gd.cpu().numpy(), gd32.cpu().numpy()) < 1e-6
        assert norm_err(gv.cpu().numpy(), gv32.cpu().numpy()) < 1e-6
        r = orc.step_fwd_bwd(d64, v64, tgt.astype(np.float64), sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"],
                             sc["step"], boundary_render=boundary)
        ref_img = r["img"]
    assert abs(lo.item() - r["loss"]) <= 1e-5 * abs(r["loss"])
    assert norm_err(step.images.cpu().numpy(), ref_img) < 1e-5
    assert norm_err(gd.cpu().numpy(), r["gd"]) < 1e-3
    assert norm_err(gv.cpu().numpy(), r["gv"]) < 1e-3
    with pytest.raises(_lib.InvalidArgumentError):
        step.run(td.float(), tv.float(), ttgt)
    with pytest.raises(_lib.InvalidArgumentError):
        TransportRenderStep(*args, storage="float16")
