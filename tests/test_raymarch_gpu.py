"""Parity of the HIP raymarch (RaymarchGridTransform fwd/bwd) with the oracle.

Tolerances: forward images within 1e-5 normalised max error of the FP64
oracle; gradients within 1e-4 (EXACT and REFERENCE) of the FP64 oracle.
"""
import numpy as np
import pytest
import torch

from helpers import norm_err, scene
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _ops():
    from gft.ops import raymarching_ops
    return raymarching_ops


def _run_fwd(sc, data, **kw):
    ops = _ops()
    t = torch.from_numpy(data).cuda()
    out = ops.raymarch_grid_transform(t, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], **kw)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("channels", [1, 2, 4])
@pytest.mark.parametrize("blend", ["BEER_LAMBERT", "ALPHA", "ALPHA_ADDITIVE", "ADDITIVE"])
@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP", "WRAP"])
def test_forward_linear(cuda, channels, blend, boundary):
    sc = scene(channels=channels, seed=3)
    data = sc["data"] * sc["step"] * (0.2 if blend.startswith("ALPHA") else 1.0)
    got = _run_fwd(sc, data, interpolation="LINEAR", boundary=boundary, blending_mode=blend)
    ref = orc.raymarch_fwd(data, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], "LINEAR", boundary, blend)
    assert got.shape == ref.shape
    assert norm_err(got, ref) < 1e-5


@pytest.mark.parametrize("filt", ["NEAREST", "MIN", "MAX"])
def test_forward_other_filters(cuda, filt):
    sc = scene(channels=4, seed=5)
    data = sc["data"] * sc["step"]
    got = _run_fwd(sc, data, interpolation=filt, boundary="CLAMP", blending_mode="ADDITIVE")
    ref = orc.raymarch_fwd(data, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], filt, "CLAMP", "ADDITIVE")
    assert norm_err(got, ref) < 1e-5


def test_forward_keep_dims_batch(cuda):
    sc = scene(channels=2, seed=7, views=2)
    data = np.concatenate([sc["data"], 0.5 * sc["data"]], axis=0) * sc["step"]
    M = np.concatenate([sc["M"], sc["M"]], axis=0)
    ops = _ops()
    out = ops.raymarch_grid_transform(torch.from_numpy(data).cuda(), M, sc["V"], sc["P"], sc["F"], sc["out_dhw"],
                                      keep_dims=True).cpu().numpy()
    assert out.shape == (2, 2, 1) + sc["out_dhw"][1:] + (2,)
    ref = orc.raymarch_fwd(data, M, sc["V"], sc["P"], sc["F"], sc["out_dhw"])
    assert norm_err(out[:, :, 0], ref) < 1e-5


def test_forward_camera_per_batch(cuda):
    sc = scene(channels=1, seed=8, views=2)
    data = np.concatenate([sc["data"], 2.0 * sc["data"]], axis=0) * sc["step"]
    M = np.concatenate([sc["M"], sc["M"]], axis=0)
    ops = _ops()
    out = ops.raymarch_grid_transform(torch.from_numpy(data).cuda(), M, sc["V"], sc["P"], sc["F"], sc["out_dhw"],
                                      keep_dims=True, separate_camera_batch=False).cpu().numpy()
    ref = orc.raymarch_fwd(data, M, sc["V"], sc["P"], sc["F"], sc["out_dhw"], separate_camera_batch=False)
    assert out.shape[1] == 1
    assert norm_err(out[:, :, 0], ref) < 1e-5


@pytest.mark.parametrize("grad_mode", ["EXACT", "REFERENCE"])
@pytest.mark.parametrize("channels", [1, 2, 4])
@pytest.mark.parametrize("blend", ["BEER_LAMBERT", "ALPHA", "ALPHA_ADDITIVE", "ADDITIVE"])
@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP"])
def test_backward(cuda, grad_mode, channels, blend, boundary):
    sc = scene(channels=channels, seed=11)
    data = sc["data"] * sc["step"] * (0.2 if blend.startswith("ALPHA") else 1.0)
    ref_out = orc.raymarch_fwd(data, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], "LINEAR", boundary, blend)
    rng = np.random.Generator(np.random.PCG64(4))
    og = rng.uniform(-1, 1, size=ref_out.shape).astype(np.float32)
    ops = _ops()
    t = torch.from_numpy(data).cuda()
    out = ops.raymarch_grid_transform(t, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], boundary=boundary,
                                      blending_mode=blend)
    g = ops.raymarch_grid_transform_grad(t, out, torch.from_numpy(og).cuda(), sc["M"], sc["V"], sc["P"], sc["F"],
                                         sc["out_dhw"], boundary=boundary, blending_mode=blend,
                                         grad_mode=grad_mode).cpu().numpy()
    ref = orc.raymarch_bwd(data, out.cpu().numpy(), og, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], boundary,
                           blend, True, grad_mode)
    assert norm_err(g, ref) < 1e-4


def test_autograd_matches_grad_op(cuda):
    sc = scene(channels=4, seed=13)
    data = torch.from_numpy(sc["data"] * sc["step"]).cuda().requires_grad_(True)
    ops = _ops()
    out = ops.raymarch_grid_transform(data, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], grad_mode="EXACT")
    w = torch.randn_like(out)
    (out * w).sum().backward()
    g2 = ops.raymarch_grid_transform_grad(data.detach(), out.detach(), w, sc["M"], sc["V"], sc["P"], sc["F"],
                                          sc["out_dhw"], grad_mode="EXACT")
    assert norm_err(data.grad.cpu().numpy(), g2.cpu().numpy()) < 1e-5


def test_invalid_arguments(cuda):
    from gft._lib import InvalidArgumentError, UnimplementedError
    sc = scene(channels=1)
    ops = _ops()
    t = torch.from_numpy(sc["data"]).cuda()
    with pytest.raises(InvalidArgumentError):
        ops.raymarch_grid_transform(t, sc["M"], sc["V"], sc["P"][:2], sc["F"], sc["out_dhw"])
    with pytest.raises(InvalidArgumentError):
        ops.raymarch_grid_transform(t, sc["M"], sc["V"], sc["P"], sc["F"], (0, 4, 4))
    with pytest.raises(UnimplementedError):
        ops.raymarch_grid_transform(torch.zeros((1, 4, 4, 4, 3), device="cuda"), sc["M"], sc["V"], sc["P"], sc["F"],
                                    sc["out_dhw"])
    with pytest.raises(InvalidArgumentError):
        ops.raymarch_grid_transform(t, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], blending_mode="FOO")


def _ortho_scene(yaw, channels=4):
    from gft import synthetic as syn
    from gft.camera import camera_params_batch, model_matrices
    from helpers import ortho_camera
    grid = (9, 11, 8)
    cam = ortho_camera(grid, (11, 8), 10, yaw=yaw)
    V, P, F = camera_params_batch([cam])
    d = syn.gaussian_blobs(grid, 4, seed=6)
    rng = np.random.Generator(np.random.PCG64(17))
    d = (d * rng.uniform(0.3, 1.0, size=(1, 1, 1, 1, channels))).astype(np.float32)
    return dict(data=d, M=model_matrices([syn.grid_transform(grid)]), V=V, P=P, F=F, out_dhw=(10, 11, 8),
                step=cam.depth_step)


@pytest.mark.parametrize("yaw", [0.0, 90.0, 180.0])
@pytest.mark.parametrize("grad_mode", ["EXACT", "REFERENCE"])
def test_orthographic_axis_aligned(cuda, yaw, grad_mode):
    """Rays along a grid axis (two step components exactly zero) through the
    op's forward and both gradient modes."""
    sc = _ortho_scene(yaw)
    data = sc["data"] * sc["step"]
    ops = _ops()
    t = torch.from_numpy(data).cuda()
    out = ops.raymarch_grid_transform(t, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"])
    ref_out = orc.raymarch_fwd(data, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"])
    assert np.abs(ref_out).max() > 0.05
    assert norm_err(out.cpu().numpy(), ref_out) < 1e-5
    og = np.random.Generator(np.random.PCG64(2)).uniform(-1, 1, size=ref_out.shape).astype(np.float32)
    g = ops.raymarch_grid_transform_grad(t, out, torch.from_numpy(og).cuda(), sc["M"], sc["V"], sc["P"], sc["F"],
                                         sc["out_dhw"], grad_mode=grad_mode).cpu().numpy()
    ref = orc.raymarch_bwd(data, out.cpu().numpy(), og, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], "BORDER",
                           "BEER_LAMBERT", True, grad_mode)
    assert norm_err(g, ref) < 1e-4
