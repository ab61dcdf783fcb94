"""Parity of the HIP ReduceGridBlend / ReduceGridBlendGrad with the oracle
(all blend modes, C in {1,2,4}, keep_dims, REFERENCE and EXACT gradients).
Tolerance: 1e-5 normalised max error forward, 1e-5 gradients (no atomics,
same sequential order as the oracle)."""
import numpy as np
import pytest
import torch

from helpers import norm_err
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

BLENDS = ["BEER_LAMBERT", "ALPHA", "ALPHA_ADDITIVE", "ADDITIVE"]


def _data(channels, seed, alpha):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.uniform(0.0, 0.3 if alpha else 1.0, size=(2, 9, 13, 17, channels)).astype(np.float32)
    return x


@pytest.mark.parametrize("blend", BLENDS)
@pytest.mark.parametrize("channels", [1, 2, 4])
@pytest.mark.parametrize("keep_dims", [False, True])
def test_reduce_blend_forward(cuda, blend, channels, keep_dims):
    from gft.ops import blending_ops
    x = _data(channels, 1, blend.startswith("ALPHA"))
    got = blending_ops.reduce_grid_blend(torch.from_numpy(x).cuda(), blend, keep_dims).cpu().numpy()
    ref = orc.reduce_blend_fwd(x, blend, keep_dims)
    assert got.shape == ref.shape
    assert norm_err(got, ref) < 1e-5


@pytest.mark.parametrize("blend", BLENDS)
@pytest.mark.parametrize("channels", [1, 2, 4])
@pytest.mark.parametrize("keep_dims", [False, True])
@pytest.mark.parametrize("grad_mode", ["REFERENCE", "EXACT"])
def test_reduce_blend_backward(cuda, blend, channels, keep_dims, grad_mode):
    from gft._lib import UnimplementedError
    from gft.ops import blending_ops
    x = _data(channels, 2, blend.startswith("ALPHA"))
    out = orc.reduce_blend_fwd(x, blend, keep_dims)
    rng = np.random.Generator(np.random.PCG64(3))
    og = rng.uniform(-1, 1, size=out.shape).astype(np.float32)
    args = (torch.from_numpy(og).cuda(), torch.from_numpy(out.astype(np.float32)).cuda(), torch.from_numpy(x).cuda(),
            blend, keep_dims, grad_mode)
    if grad_mode == "REFERENCE" and blend == "ALPHA_ADDITIVE":
        with pytest.raises(UnimplementedError):
            blending_ops.reduce_grid_blend_grad(*args)
        return
    got = blending_ops.reduce_grid_blend_grad(*args).cpu().numpy()
    ref = orc.reduce_blend_bwd(x, out.astype(np.float32), og, blend, keep_dims, grad_mode)
    assert norm_err(got, ref) < 1e-5


def test_reduce_blend_autograd(cuda):
    from gft.ops import blending_ops
    x = _data(4, 4, False)
    t = torch.from_numpy(x).cuda().requires_grad_(True)
    y = blending_ops.reduce_grid_blend(t, "BEER_LAMBERT", False, grad_mode="EXACT")
    rng = np.random.Generator(np.random.PCG64(5))
    og = rng.uniform(-1, 1, size=tuple(y.shape)).astype(np.float32)
    y.backward(torch.from_numpy(og).cuda())
    ref = orc.reduce_blend_bwd(x, y.detach().cpu().numpy(), og, "BEER_LAMBERT", False, "EXACT")
    assert norm_err(t.grad.cpu().numpy(), ref) < 1e-5
