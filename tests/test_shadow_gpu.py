"""The fused spot-light shadow transmittance (lighting.py:106-141) against the
op sequence it replaces (frustum resample of the step-scaled density,
reduce_grid_blend keep_dims, one-cell shift, exp / clip, cone mask) on the
same HIP ops, and against the FP64 oracle composition.

Tolerances: the prefix sums are bit-identical to reduce_grid_blend over the
sampled frustum (same taps, same summation order); the transmittance is
within 2 ulp of the op sequence (exp of the same argument); density
gradients within 1e-5 normalised max error (the REFERENCE and EXACT
resample gradients accumulate with atomics, in no fixed order); the oracle
within 1e-5 (forward) and 1e-3 (gradient).
"""
import numpy as np
import pytest
import torch

from gft import synthetic as syn
from helpers import norm_err, scene
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GRID = (16, 20, 18)


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _setup(res, seed=60, cone=True):
    from gft.lighting import SpotLight
    from gft.transform import Transform
    sc = scene(grid=GRID, image=(8, 6), views=1, seed=seed)
    gt = syn.grid_transform(GRID)
    gt.set_data(_t(sc["data"] * 0.1))
    t = Transform(translation=[9.0, 36.0, 8.0], rotation_deg=[-90.0, 0.0, 0.0])
    sl = SpotLight(t, intensity=1.0, range_scale=0.01, angle_deg=60, cast_shadows=True, shadow_resolution=list(res),
                   shadow_clip=[12, 45], cone_mask=cone)
    return sc, gt, sl


def _renderer(boundary, blend, filt, mips, grad_mode="REFERENCE", gather=False):
    from gft.renderer import Renderer
    mm, nm = mips
    return Renderer(filter_mode=filt, boundary_mode=boundary, mipmapping=mm, num_mips=nm, mip_bias=0.3,
                    blend_mode=blend, grad_mode=grad_mode, sample_gradients=gather, fast_gradient_mip_bias_add=0.25)


CASES = [(b, bl, f, m) for b in ("BORDER", "CLAMP") for bl in ("BEER_LAMBERT", "ALPHA")
         for f, m in (("LINEAR", ("NONE", 0)), ("NEAREST", ("NONE", 0)), ("LINEAR", ("LINEAR", 2)))]


@pytest.mark.parametrize("boundary,blend,filt,mips", CASES)
@pytest.mark.parametrize("res", [(70, 13, 11), (24, 16, 16)])
def test_shadow_fused_matches_op_sequence(cuda, boundary, blend, filt, mips, res):
    from gft import ops
    sc, gt, sl = _setup(res, cone=res[0] != 70)
    r = _renderer(boundary, blend, filt, mips)
    mask = sl._shadow_mask(cuda) if sl.cone_mask else None
    fused = r.shadow_transmission(gt, sl.shadow_cam, mask)
    assert fused is not None and tuple(fused.shape) == (1,) + tuple(res) + (1,)
    seq = sl._shadow_ops(gt, r)
    if mask is not None:
        seq = seq * mask
    f, s = fused.cpu().numpy(), seq.cpu().numpy()
    assert np.isfinite(f).all()
    lit = f[f > 0]
    assert lit.min() < 0.5 and lit.max() > 0.9  # shadowed and lit cells
    assert np.abs(f - s).max() <= 2.5e-7
    # prefix sums: bit-identical to reduce_grid_blend over the sampled frustum
    M = np.asarray(gt.get_transform_matrix(), dtype=np.float32).T[None]
    V, P, F = r._get_camera_params_batch([sl.shadow_cam])
    cams = ops._cams(M, V, P, F)
    fm, bm, cm, mm = ops._check_transform_attrs(filt, boundary, "TRANSFORM_LINDEPTH", 0.5, mips[0], mips[1])
    step = sl.shadow_cam.depth_step
    _, saved = ops._shadow_fwd_impl(gt.data, cams, tuple(res), step, fm, bm, mm, mips[1], 0.3, ops.BLEND[blend],
                                    mask, save=True)
    scaled = gt.data * torch.tensor([step], dtype=torch.float32, device=cuda)
    fr = ops.sampling_ops.sample_grid_transform(scaled, M, V, P, F, list(res), interpolation=filt, boundary=boundary,
                                                mipmapping=mips[0], num_mipmaps=mips[1], mip_bias=0.3)
    pre = ops.blending_ops.reduce_grid_blend(fr[0], blend, keep_dims=True)
    assert torch.equal(saved, pre.reshape(saved.shape))


@pytest.mark.parametrize("boundary,blend", [("BORDER", "BEER_LAMBERT"), ("CLAMP", "ALPHA"), ("CLAMP", "BEER_LAMBERT"),
                                            ("BORDER", "ALPHA")])
@pytest.mark.parametrize("grad_mode,gather,mips", [("REFERENCE", False, ("NONE", 0)), ("EXACT", False, ("NONE", 0)),
                                                   ("EXACT", False, ("LINEAR", 2)), ("REFERENCE", True, ("LINEAR", 2))])
def test_shadow_fused_gradient_matches_op_sequence(cuda, boundary, blend, grad_mode, gather, mips):
    sc, gt, sl = _setup((70, 13, 11), seed=61)
    r = _renderer(boundary, blend, "LINEAR", mips, grad_mode=grad_mode, gather=gather)
    rng = np.random.Generator(np.random.PCG64(7))
    w = _t(rng.uniform(-1, 1, size=(1,) + GRID + (1,)))
    grads = []
    for fused in (True, False):
        x = gt.data.detach().clone().requires_grad_(True)
        gt.set_data(x)
        light = sl.render_shadow_map(gt, r, fused=fused)
        (light * w).sum().backward()
        grads.append(x.grad.cpu().numpy())
    assert np.abs(grads[1]).max() > 0
    assert norm_err(grads[0], grads[1]) < 1e-5


def test_shadow_fused_against_oracle_gradient(cuda):
    """EXACT gradient of sum(w * T_frustum) through the fused op against the
    FP64 oracle's resample adjoint of the analytic scan gradient."""
    from gft import ops
    sc, gt, sl = _setup((40, 12, 10), seed=62)
    res = sl.shadow_resolution
    r = _renderer("BORDER", "BEER_LAMBERT", "LINEAR", ("NONE", 0), grad_mode="EXACT")
    mask = sl._shadow_mask(cuda)
    x = gt.data.detach().clone().requires_grad_(True)
    gt.set_data(x)
    tr = r.shadow_transmission(gt, sl.shadow_cam, mask)
    rng = np.random.Generator(np.random.PCG64(8))
    w = rng.uniform(-1, 1, size=tuple(tr.shape))
    (tr * _t(w)).sum().backward()
    # oracle: frustum samples, prefix sums, shifted transmittance and its adjoint
    from gft.camera import camera_params_batch
    Vs, Ps, Fs = camera_params_batch([sl.shadow_cam])
    M = np.asarray(gt.get_transform_matrix(), dtype=np.float32).T[None]
    step = sl.shadow_cam.depth_step
    d = x.detach().cpu().numpy().astype(np.float64)
    fr = orc.sample_transform_fwd(d * step, M, Vs, Ps, Fs, res, "LINEAR", "BORDER")[0, 0, ..., 0]
    pre = np.cumsum(fr, axis=0)
    m = mask.cpu().numpy().reshape(res[1], res[2]).astype(np.float64)
    T = np.concatenate([np.ones_like(pre[:1]), np.exp(-pre[:-1])], axis=0) * m
    assert norm_err(tr.detach().cpu().numpy()[0, ..., 0], T) < 1e-5
    gT = w[0, ..., 0]
    u = np.zeros_like(pre)
    u[:-1] = -gT[1:] * m * np.exp(-pre[:-1])
    gs = np.cumsum(u[::-1], axis=0)[::-1]
    gref = orc.sample_transform_bwd(d, (gs * step)[None, None, ..., None], M, Vs, Ps, Fs, "LINEAR", "BORDER",
                                    grad_mode="EXACT")
    assert np.abs(gref).max() > 0
    assert norm_err(x.grad.cpu().numpy(), gref) < 1e-3


def test_shadow_fused_errors(cuda):
    from gft import _lib, ops
    sc, gt, sl = _setup((8, 8, 8))
    M = np.asarray(gt.get_transform_matrix(), dtype=np.float32).T[None]
    from gft.camera import camera_params_batch
    V, P, F = camera_params_batch([sl.shadow_cam])
    with pytest.raises(_lib.InvalidArgumentError):
        ops.lighting_ops.shadow_transmission(gt.data, M, V, P, F, [8, 8, 8], 1.0, blending_mode="ADDITIVE")
    with pytest.raises(_lib.InvalidArgumentError):
        ops.lighting_ops.shadow_transmission(gt.data, M, V, P, F, [8, 8, 8], 1.0, boundary="WRAP")
    with pytest.raises(_lib.InvalidArgumentError):
        ops.lighting_ops.shadow_transmission(gt.data.expand(1, -1, -1, -1, 2).contiguous(), M, V, P, F, [8, 8, 8],
                                             1.0)
    with pytest.raises(_lib.InvalidArgumentError):
        ops.lighting_ops.shadow_transmission(gt.data, M, V, P, F, [8, 8, 8], 1.0, cone_mask=torch.ones(5, device=cuda))
    # the renderer falls back to the op sequence outside the fused op's settings
    from gft.renderer import Renderer
    assert Renderer(filter_mode="MAX", mipmapping="NONE").shadow_transmission(gt, sl.shadow_cam) is None
    assert Renderer(mipmapping="NONE", fused_shadow=False).shadow_transmission(gt, sl.shadow_cam) is None
