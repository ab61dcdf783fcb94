"""The plain per-op path (the benchmark's baseline) composed from the HIP
ops matches the same composition of the oracle's ops: MacCormack through
four LuT passes, materialised TRANSFORM_LINDEPTH frusta, depth blend, L2
loss, and the reference (count-normalised) gradients through autograd.
Tolerances: loss and images 1e-5, gradients 1e-3 (normalised max error vs
the FP64 oracle)."""
import numpy as np
import pytest
import torch

from gft import synthetic as syn
from helpers import norm_err, scene
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _oracle_plain(d0, v, tgt, sc, light):
    step = sc["step"]
    d1 = orc.advect_fwd(d0, v, 2, "MC_SMOOTH")
    grid = np.concatenate([d1 * np.asarray(light), d1], -1) * step
    fr = orc.sample_transform_fwd(grid, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], "LINEAR", "BORDER",
                                  coordinate_mode="TRANSFORM_LINDEPTH")
    img = orc.reduce_blend_fwd(fr[0], "BEER_LAMBERT", False)
    diff = img[..., :3] - tgt[0]
    loss = float(np.mean(diff ** 2))
    gimg = np.zeros_like(img)
    gimg[..., :3] = 2 * diff / diff.size
    gfr = orc.reduce_blend_bwd(fr[0], img, gimg, "BEER_LAMBERT", False, "REFERENCE")
    ggrid = orc.sample_transform_bwd(grid, gfr[None], sc["M"], sc["V"], sc["P"], sc["F"],
                                     coordinate_mode="TRANSFORM_LINDEPTH", grad_mode="REFERENCE")
    gd1 = step * (ggrid[..., :3] @ np.asarray(light) + ggrid[..., 3])[..., None]
    gd, gv = orc.advect_bwd(d0, v, gd1, 2, "MC_SMOOTH", grad_mode="REFERENCE")
    return loss, img, gd, gv


def test_plain_step_matches_oracle_composition(cuda):
    from gft.plain import PlainTransportRenderStep
    grid = (12, 18, 14)
    sc = scene(grid=grid, image=(16, 11), views=3, seed=21)
    d0 = syn.gaussian_blobs(grid, 6, seed=22)
    v = syn.curl_noise_velocity(grid, seed=23, max_speed=1.5, periods=(8, 4, 2))
    rng = np.random.Generator(np.random.PCG64(24))
    tgt = rng.uniform(0, 0.3, size=(1, 3, 16, 11, 3)).astype(np.float32)
    light = (0.9, 1.0, 1.1)
    step = PlainTransportRenderStep(grid, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], light)
    loss, gd, gv = step.run(torch.from_numpy(d0).cuda(), torch.from_numpy(v).cuda(), torch.from_numpy(tgt).cuda())
    rloss, rimg, rgd, rgv = _oracle_plain(d0.astype(np.float64), v.astype(np.float64), tgt.astype(np.float64), sc,
                                          light)
    assert abs(loss.item() - rloss) <= 1e-5 * rloss
    assert norm_err(step.images[0].cpu().numpy(), rimg) < 1e-5
    assert norm_err(gd.cpu().numpy(), rgd) < 1e-3
    assert norm_err(gv.cpu().numpy(), rgv) < 1e-3
