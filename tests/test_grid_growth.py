"""Grid growth rescaling (renderer.py:653-692, data_structures.py:236-261,
551-590, 1178-1183): the aligned resample's sample positions against their
closed forms on a linear field, through the oracle's TRANSFORM resample on
the CPU (pins the matrices), and the HIP path against the same closed forms
and the oracle on the GPU.

A linear field is reproduced exactly by trilinear interpolation wherever the
sample position lies inside the grid's cell-centre box, and CLAMP sampling
outside it returns the field at the clamped position; the expected value of
output cell q along an axis with n_in -> n_out cells is f at
  BORDER          (q+0.5) n_in/n_out - 0.5
  CENTER          q (n_in-1)/(n_out-1)
  STAGGER_INPUT   (q+0.5) (n_in-1)/n_out
  STAGGER_OUTPUT  (q+0.5) (n_in+1)/n_out - 1
clamped to [0, n_in-1]. Tolerance 1e-5 of the field's range (FP32
positions).
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as orc

COEF = (0.3, 0.7, 1.3, -0.4)  # f = c0 + c1 x + c2 y + c3 z in cell-centred index coordinates


def _field(dhw):
    D, H, W = dhw
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    return (COEF[0] + COEF[1] * x + COEF[2] * y + COEF[3] * z)[None, ..., None].astype(np.float64)


def _pos(align, n_in, n_out):
    q = np.arange(n_out, dtype=np.float64)
    a = align.upper()
    if a == "BORDER":
        p = (q + 0.5) * n_in / n_out - 0.5
    elif a == "CENTER":
        p = q * (n_in - 1) / (n_out - 1)
    elif a == "STAGGER_INPUT":
        p = (q + 0.5) * (n_in - 1) / n_out
    else:
        p = (q + 0.5) * (n_in + 1) / n_out - 1
    return np.clip(p, 0, n_in - 1)


def _expected(in_dhw, out_dhw, ax, ay, az):
    pz = _pos(az, in_dhw[0], out_dhw[0])
    py = _pos(ay, in_dhw[1], out_dhw[1])
    px = _pos(ax, in_dhw[2], out_dhw[2])
    z, y, x = np.meshgrid(pz, py, px, indexing="ij")
    return (COEF[0] + COEF[1] * x + COEF[2] * y + COEF[3] * z)[None, ..., None]


ALIGNS = ["BORDER", "CENTER", "STAGGER_INPUT", "STAGGER_OUTPUT"]
SHAPES = [((5, 7, 6), (9, 12, 11)), ((8, 9, 10), (4, 5, 6))]


def _params(in_dhw, out_dhw, ax, ay, az):
    from gft.renderer import Renderer
    tin, tout = Renderer.aligned_transforms(in_dhw, out_dhw, ax, ay, az)
    return Renderer._sample_transform_params([tin], [tout])


@pytest.mark.parametrize("in_dhw,out_dhw", SHAPES)
@pytest.mark.parametrize("align", ALIGNS)
def test_aligned_resample_closed_form_oracle(in_dhw, out_dhw, align):
    """Per-axis alignment on one axis at a time (the others BORDER), all
    three axes together, through the FP64 oracle's TRANSFORM resample."""
    f = _field(in_dhw)
    rng = (f.max() - f.min())
    for axes in ((align, "BORDER", "BORDER"), ("BORDER", align, "BORDER"), ("BORDER", "BORDER", align),
                 (align, align, align)):
        M, V, P, F = _params(in_dhw, out_dhw, *axes)
        got = orc.sample_transform_fwd(f, M, V, P, F, out_dhw, "LINEAR", "CLAMP", coordinate_mode="TRANSFORM")[:, 0]
        want = _expected(in_dhw, out_dhw, *axes)
        assert np.abs(got - want).max() <= 1e-5 * rng, (axes, np.abs(got - want).max())


def test_aligned_transforms_reject_unknown_alignment():
    from gft.renderer import Renderer
    with pytest.raises(ValueError):
        Renderer.aligned_transforms((4, 4, 4), (8, 8, 8), align_x="MIDDLE")


def _t(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("in_dhw,out_dhw", SHAPES)
@pytest.mark.parametrize("ax,ay,az", list(itertools.product(["BORDER", "CENTER"], ["STAGGER_INPUT", "BORDER"],
                                                            ["STAGGER_OUTPUT", "CENTER"])))
def test_aligned_resample_gpu(cuda, in_dhw, out_dhw, ax, ay, az):
    from gft.renderer import Renderer
    f = _field(in_dhw)
    r = Renderer(filter_mode="LINEAR", boundary_mode="CLAMP", mipmapping="NONE", blend_mode="ADDITIVE")
    got = r.resample_grid3D_aligned(_t(f), out_dhw, ax, ay, az).cpu().numpy()
    want = _expected(in_dhw, out_dhw, ax, ay, az)
    assert got.shape == want.shape
    assert np.abs(got - want).max() <= 1e-5 * (f.max() - f.min())


@pytest.mark.gpu
@pytest.mark.parametrize("boundary,mips", [("BORDER", ("NONE", 0)), ("CLAMP", ("LINEAR", 2)), ("BORDER", ("LINEAR", 3))])
def test_aligned_resample_gpu_vs_oracle(cuda, boundary, mips):
    """Random two-channel batch, shrink by 2 with mips (the LoD path) and
    BORDER zeros, against the oracle's resample with the same matrices."""
    from gft.renderer import Renderer
    rng = np.random.Generator(np.random.PCG64(3))
    data = rng.uniform(0, 1, size=(2, 12, 16, 14, 2))
    out_dhw = (6, 8, 7)
    r = Renderer(filter_mode="LINEAR", boundary_mode=boundary, mipmapping=mips[0], num_mips=mips[1], mip_bias=0.0,
                 blend_mode="ADDITIVE")
    got = r.resample_grid3D_aligned(_t(data), out_dhw, "CENTER", "BORDER", "STAGGER_OUTPUT").cpu().numpy()
    from gft.renderer import Renderer as R
    tin, tout = R.aligned_transforms(data.shape[1:4], out_dhw, "CENTER", "BORDER", "STAGGER_OUTPUT")
    M, V, P, F = R._sample_transform_params([tin, tin], [tout])
    ref = orc.sample_transform_fwd(data, M, V, P, F, out_dhw, "LINEAR", boundary, mipmapping=mips[0],
                                   num_mipmaps=mips[1], coordinate_mode="TRANSFORM")[:, 0]
    assert np.abs(ref).max() > 0.1
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.gpu
def test_velocity_and_density_scaled(cuda):
    """VelocityGrid.scaled: each component centre-aligned along its own axis,
    border-aligned along the others, magnitudes scaled by new/old size; a
    linear field per component reproduces the closed form. DensityGrid and
    State follow."""
    from gft.grids import DensityGrid, State, VelocityGrid
    c_old, c_new = [6, 8, 7], [11, 15, 13]
    shapes_old = VelocityGrid.component_shapes(c_old)
    shapes_new = VelocityGrid.component_shapes(c_new)
    comps = [_t(_field(s)) for s in shapes_old]
    v = VelocityGrid(c_old, x=comps[0], y=comps[1], z=comps[2], boundary="CLAMP")
    x, y, z = v.scaled(c_new)
    factors = [n / o for o, n in zip(c_old, c_new)]  # z, y, x
    aligns = [("CENTER", "BORDER", "BORDER"), ("BORDER", "CENTER", "BORDER"), ("BORDER", "BORDER", "CENTER")]
    for got, so, sn, al, fac in zip((x, y, z), shapes_old, shapes_new, aligns, factors[::-1]):
        want = _expected(so, sn, *al) * fac
        assert list(got.shape[1:4]) == sn
        assert np.abs(got.cpu().numpy() - want).max() <= 1e-5 * np.abs(want).max()
    x2, _, _ = v.scaled(c_new, scale_magnitude=False)
    assert np.allclose(x2.cpu().numpy() * factors[2], x.cpu().numpy(), rtol=1e-6, atol=1e-6)
    same = v.scaled(c_old)
    assert all(np.array_equal(a.cpu().numpy(), b.cpu().numpy()) for a, b in zip(same, comps))
    d = DensityGrid([6, 8, 7], d=_t(_field([6, 8, 7])))
    dn = d.copy_scaled([12, 16, 14])
    assert dn.shape == [12, 16, 14]
    assert np.abs(dn.d.cpu().numpy() - _expected([6, 8, 7], [12, 16, 14], "BORDER", "BORDER", "BORDER")).max() < 1e-4
    st = State(d, v, 0)
    st.rescale([12, 16, 14], c_new)
    assert st.density.shape == [12, 16, 14] and st.velocity.centered_shape == c_new
