"""Host-side behaviour of the point/spot light falloff factor (CPU tensors).

The factor 1 / (1 + (|p - l| s)^2) per cell (lighting.py:54-73) depends on
the grid transform, the grid shape, the light's position and its range scale,
not on the density, so a light keeps it between steps that render the same
grid. These tests check that a repeated call returns the kept factor, and that
moving the light, changing its range scale or rendering another grid gives
the factor a fresh light computes for those inputs.
"""
import numpy as np
import torch

from gft import synthetic as syn
from gft.lighting import PointLight
from gft.transform import Transform


def _grid(shape, value=0.0):
    gt = syn.grid_transform(shape)
    gt.set_data(torch.full((1,) + tuple(shape) + (1,), value))
    return gt


def _fresh(translation, scale, gt):
    return PointLight(Transform(translation=translation), intensity=0.5, range_scale=scale).falloff_factor(gt)


def test_falloff_is_kept_for_the_same_inputs():
    gt = _grid((6, 8, 7))
    pl = PointLight(Transform(translation=[0, 0, 30]), intensity=0.5, range_scale=0.05)
    a = pl.falloff_factor(gt)
    gt.set_data(torch.full((1, 6, 8, 7, 1), 2.0))  # another density on the same grid
    b = pl.falloff_factor(gt)
    assert b is a
    assert torch.equal(a, _fresh([0, 0, 30], 0.05, gt))
    assert a.shape == (1, 6, 8, 7, 1) and a.dtype == torch.float32
    assert float(a.min()) > 0.0 and float(a.max()) <= 1.0


def test_falloff_follows_the_light_and_the_grid():
    gt = _grid((6, 8, 7))
    pl = PointLight(Transform(translation=[0, 0, 30]), intensity=0.5, range_scale=0.05)
    a = pl.falloff_factor(gt)
    pl.t = Transform(translation=[0, 5, 30])
    moved = pl.falloff_factor(gt)
    assert moved is not a
    assert torch.equal(moved, _fresh([0, 5, 30], 0.05, gt))
    assert float((moved - a).abs().max()) > 1e-3
    pl.s = 0.1
    rescaled = pl.falloff_factor(gt)
    assert torch.equal(rescaled, _fresh([0, 5, 30], 0.1, gt))
    other = _grid((5, 9, 4))
    g2 = pl.falloff_factor(other)
    assert g2.shape == (1, 5, 9, 4, 1)
    assert torch.equal(g2, _fresh([0, 5, 30], 0.1, other))
    # back on the first grid: recomputed, and equal to the earlier result
    assert np.array_equal(pl.falloff_factor(gt).numpy(), rescaled.numpy())
