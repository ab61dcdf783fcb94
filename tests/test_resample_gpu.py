"""Parity of the HIP resampling ops with the oracle: SampleGridTransform
(four coordinate modes, mips), its REFERENCE/EXACT gradients and the
renderer's gather gradient, SampleGridLut (all filters and boundaries) and
its gradients, the mip pyramid and LodTransform.

Tolerances (normalised max error): LINEAR forward 1e-5 of the FP64 oracle;
NEAREST/MIN/MAX and NEAREST mip selection may flip at exact ties between FP32
implementations, so at most 0.2 % of cells may differ from the FP32 oracle and
the rest match within 1e-5; gradients 1e-4 of the FP64 oracle.
"""
import numpy as np
import pytest
import torch

from helpers import norm_err, scene
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

REVERSE = {"TRANSFORM_LINDEPTH": False, "TRANSFORM": False, "TRANSFORM_LINDEPTH_REVERSE": True,
           "TRANSFORM_REVERSE": True}


def _ops():
    from gft.ops import sampling_ops
    return sampling_ops


def _setup(channels=1, seed=0, views=3, frustum_input=False):
    sc = scene(grid=(10, 18, 14), image=(12, 9), views=views, seed=seed, channels=channels)
    vol = sc["data"]
    vol_dhw = vol.shape[1:4]
    fr_dhw = tuple(sc["out_dhw"])
    rng = np.random.Generator(np.random.PCG64(seed + 100))
    if frustum_input:
        data = rng.uniform(0.0, 1.0, size=(1,) + fr_dhw + (channels,)).astype(np.float32)
        return sc, data, vol_dhw
    return sc, vol, fr_dhw


def _close_with_ties(got, ref32, ref64, tol=1e-5, frac=2e-3):
    scale = max(np.abs(ref64).max(), 1e-30)
    bad = np.abs(got - ref32) > tol * scale
    assert bad.mean() <= frac, bad.mean()
    assert np.abs(got[~bad] - ref64[~bad]).max() <= 5 * tol * scale


@pytest.mark.parametrize("coord", list(REVERSE))
@pytest.mark.parametrize("channels", [1, 2, 4])
@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP"])
def test_transform_forward_linear(cuda, coord, channels, boundary):
    sc, data, out_dhw = _setup(channels, seed=1, frustum_input=REVERSE[coord])
    got = _ops().sample_grid_transform(torch.from_numpy(data).cuda(), sc["M"], sc["V"], sc["P"], sc["F"], out_dhw,
                                       interpolation="LINEAR", boundary=boundary, coordinate_mode=coord)
    ref = orc.sample_transform_fwd(data, sc["M"], sc["V"], sc["P"], sc["F"], out_dhw, "LINEAR", boundary,
                                   coordinate_mode=coord)
    assert tuple(got.shape) == ref.shape
    assert norm_err(got.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("coord", list(REVERSE))
@pytest.mark.parametrize("mip", [("NONE", 0), ("NEAREST", 2), ("LINEAR", 2), ("LINEAR", 5)])
def test_transform_forward_filters_and_mips(cuda, coord, mip):
    sc, data, out_dhw = _setup(2, seed=2, frustum_input=REVERSE[coord])
    for filt in ("NEAREST", "LINEAR"):
        kw = dict(interpolation=filt, boundary="BORDER", mipmapping=mip[0], num_mipmaps=mip[1], mip_bias=0.3,
                  coordinate_mode=coord)
        got = _ops().sample_grid_transform(torch.from_numpy(data).cuda(), sc["M"], sc["V"], sc["P"], sc["F"],
                                           out_dhw, **kw).cpu().numpy()
        ref64 = orc.sample_transform_fwd(data, sc["M"], sc["V"], sc["P"], sc["F"], out_dhw, prec=1, **kw)
        if filt == "LINEAR" and mip[0] != "NEAREST":
            assert norm_err(got, ref64) < 1e-5
        else:
            ref32 = orc.sample_transform_fwd(data, sc["M"], sc["V"], sc["P"], sc["F"], out_dhw, prec=0, **kw)
            _close_with_ties(got, ref32, ref64)


def test_transform_forward_camera_per_batch(cuda):
    sc, data, out_dhw = _setup(1, seed=3, views=2)
    data2 = np.concatenate([data, 0.5 * data[:, ::-1].copy()], axis=0)
    M = np.concatenate([sc["M"], sc["M"]], axis=0)
    got = _ops().sample_grid_transform(torch.from_numpy(data2).cuda(), M, sc["V"], sc["P"], sc["F"], out_dhw,
                                       separate_camera_batch=False).cpu().numpy()
    ref = orc.sample_transform_fwd(data2, M, sc["V"], sc["P"], sc["F"], out_dhw, separate_camera_batch=False)
    assert got.shape == ref.shape == (2, 1) + tuple(out_dhw) + (1,)
    assert norm_err(got, ref) < 1e-5


@pytest.mark.parametrize("grad_mode", ["REFERENCE", "EXACT"])
@pytest.mark.parametrize("coord", ["TRANSFORM_LINDEPTH", "TRANSFORM_LINDEPTH_REVERSE", "TRANSFORM"])
@pytest.mark.parametrize("channels", [1, 4])
@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP"])
def test_transform_backward(cuda, grad_mode, coord, channels, boundary):
    sc, data, out_dhw = _setup(channels, seed=4, frustum_input=REVERSE[coord])
    rng = np.random.Generator(np.random.PCG64(5))
    og = rng.uniform(-1, 1, size=(1, 3) + tuple(out_dhw) + (channels,)).astype(np.float32)
    kw = dict(interpolation="LINEAR", boundary=boundary, coordinate_mode=coord)
    got = _ops().sample_grid_transform_grad(torch.from_numpy(data).cuda(), torch.from_numpy(og).cuda(), sc["M"],
                                            sc["V"], sc["P"], sc["F"], grad_mode=grad_mode, **kw).cpu().numpy()
    ref = orc.sample_transform_bwd(data, og, sc["M"], sc["V"], sc["P"], sc["F"], grad_mode=grad_mode, **kw)
    assert norm_err(got, ref) < 1e-4


@pytest.mark.parametrize("mip", [("NEAREST", 2), ("LINEAR", 3)])
@pytest.mark.parametrize("filt", ["NEAREST", "LINEAR"])
def test_transform_backward_exact_mips(cuda, mip, filt):
    sc, data, out_dhw = _setup(2, seed=6)
    rng = np.random.Generator(np.random.PCG64(7))
    og = rng.uniform(-1, 1, size=(1, 3) + tuple(out_dhw) + (2,)).astype(np.float32)
    kw = dict(interpolation=filt, boundary="BORDER", mipmapping=mip[0], num_mipmaps=mip[1], mip_bias=0.2,
              coordinate_mode="TRANSFORM_LINDEPTH")
    got = _ops().sample_grid_transform_grad(torch.from_numpy(data).cuda(), torch.from_numpy(og).cuda(), sc["M"],
                                            sc["V"], sc["P"], sc["F"], grad_mode="EXACT", **kw).cpu().numpy()
    ref = orc.sample_transform_bwd(data, og, sc["M"], sc["V"], sc["P"], sc["F"], grad_mode="EXACT", prec=0, **kw)
    assert norm_err(got, ref) < 1e-4


def test_transform_exact_adjoint_identity(cuda):
    """<A x, y> == <x, A^T y> for the EXACT gradient (linear in the data)."""
    sc, data, out_dhw = _setup(4, seed=8)
    rng = np.random.Generator(np.random.PCG64(9))
    y = rng.uniform(-1, 1, size=(1, 3) + tuple(out_dhw) + (4,)).astype(np.float32)
    ops = _ops()
    kw = dict(mipmapping="LINEAR", num_mipmaps=2, coordinate_mode="TRANSFORM_LINDEPTH")
    x = torch.from_numpy(data).cuda().requires_grad_(True)
    ax = ops.sample_grid_transform(x, sc["M"], sc["V"], sc["P"], sc["F"], out_dhw, grad_mode="EXACT", **kw)
    ax.backward(torch.from_numpy(y).cuda())
    lhs = float((ax.detach().double().cpu() * torch.from_numpy(y).double()).sum())
    rhs = float((x.detach().double().cpu() * x.grad.double().cpu()).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


@pytest.mark.parametrize("coord", ["TRANSFORM_LINDEPTH", "TRANSFORM"])
@pytest.mark.parametrize("mip", [("NONE", 0), ("LINEAR", 2)])
def test_gather_gradient(cuda, coord, mip):
    sc, data, out_dhw = _setup(1, seed=10)
    rng = np.random.Generator(np.random.PCG64(11))
    og = rng.uniform(-1, 1, size=(1, 3) + tuple(out_dhw) + (1,)).astype(np.float32)
    kw = dict(interpolation="LINEAR", boundary="BORDER", mipmapping=mip[0], num_mipmaps=mip[1], mip_bias=0.5,
              coordinate_mode=coord)
    got = _ops().sample_grid_transform_gather_grad(torch.from_numpy(og).cuda(), sc["M"], sc["V"], sc["P"], sc["F"],
                                                   data.shape[1:4], **kw).cpu().numpy()
    ref = orc.sample_transform_gather_grad(og, sc["M"], sc["V"], sc["P"], sc["F"], data.shape[1:4], **kw)
    assert got.shape == data.shape
    assert norm_err(got, ref) < 1e-5
    # through autograd with grad_mode GATHER
    x = torch.from_numpy(data).cuda().requires_grad_(True)
    y = _ops().sample_grid_transform(x, sc["M"], sc["V"], sc["P"], sc["F"], out_dhw, grad_mode="GATHER",
                                     gradient_mip_bias_add=0.25, **dict(kw, mip_bias=0.25))
    y.backward(torch.from_numpy(og).cuda())
    assert norm_err(x.grad.cpu().numpy(), ref) < 1e-5


def _lut(shape_dhw, views, seed, spread=3.0, normalized=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    lut = rng.uniform(-spread, spread, size=(views,) + tuple(shape_dhw) + (4,)).astype(np.float32)
    if normalized:
        lut[..., :3] = rng.uniform(-0.1, 1.1, size=lut[..., :3].shape)
    lut[..., 3] = rng.uniform(0.0, 2.5, size=lut[..., 3].shape)
    return lut


@pytest.mark.parametrize("filt", ["NEAREST", "LINEAR", "MIN", "MAX"])
@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP", "WRAP"])
@pytest.mark.parametrize("channels", [1, 2, 4])
def test_lut_forward(cuda, filt, boundary, channels):
    sc, data, _ = _setup(channels, seed=12)
    lut = _lut((7, 9, 11), 2, seed=13)
    kw = dict(interpolation=filt, boundary=boundary)
    got = _ops().sample_grid_lut(torch.from_numpy(data).cuda(), torch.from_numpy(lut).cuda(), **kw).cpu().numpy()
    ref64 = orc.sample_lut_fwd(data, lut, prec=1, **kw)
    assert got.shape == ref64.shape
    if filt == "LINEAR":
        assert norm_err(got, ref64) < 1e-5
    else:
        _close_with_ties(got, orc.sample_lut_fwd(data, lut, prec=0, **kw), ref64)


@pytest.mark.parametrize("mode", ["absolute", "normalized", "mips_nearest", "mips_linear", "per_batch"])
def test_lut_forward_modes(cuda, mode):
    sc, data, _ = _setup(2, seed=14)
    kw = dict(interpolation="LINEAR", boundary="CLAMP")
    data_in = data
    if mode == "absolute":
        lut = _lut((6, 8, 5), 2, seed=15)
        lut[..., :3] = np.abs(lut[..., :3]) * 4.0
        kw.update(relative_coords=False)
    elif mode == "normalized":
        lut = _lut((6, 8, 5), 2, seed=15, normalized=True)
        kw.update(relative_coords=False, normalized_coords=True, cell_center_offset=0.5)
    elif mode.startswith("mips"):
        lut = _lut((6, 8, 5), 2, seed=15)
        kw.update(mipmapping=mode.split("_")[1].upper(), num_mipmaps=2, mip_bias=0.1, cell_center_offset=0.5)
    else:
        lut = _lut((6, 8, 5), 2, seed=15)
        data_in = np.concatenate([data, 2.0 * data], axis=0)
        kw.update(separate_camera_batch=False)
    got = _ops().sample_grid_lut(torch.from_numpy(data_in).cuda(), torch.from_numpy(lut).cuda(), **kw).cpu().numpy()
    ref64 = orc.sample_lut_fwd(data_in, lut, prec=1, **kw)
    if mode == "mips_nearest":
        _close_with_ties(got, orc.sample_lut_fwd(data_in, lut, prec=0, **kw), ref64)
    else:
        assert norm_err(got, ref64) < 1e-5


@pytest.mark.parametrize("grad_mode", ["REFERENCE", "EXACT"])
@pytest.mark.parametrize("filt", ["NEAREST", "LINEAR", "MAX"])
@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP", "WRAP"])
def test_lut_backward(cuda, grad_mode, filt, boundary):
    sc, data, _ = _setup(2, seed=16)
    lut = _lut((7, 9, 11), 2, seed=17)
    rng = np.random.Generator(np.random.PCG64(18))
    og = rng.uniform(-1, 1, size=(1, 2, 7, 9, 11, 2)).astype(np.float32)
    kw = dict(interpolation=filt, boundary=boundary, grad_mode=grad_mode)
    gin, glut = _ops().sample_grid_lut_grad(torch.from_numpy(data).cuda(), torch.from_numpy(og).cuda(),
                                            torch.from_numpy(lut).cuda(), **kw)
    p = 0 if filt != "LINEAR" else 1
    rin, rlut = orc.sample_lut_bwd(data, lut, og, prec=p, **kw)
    assert norm_err(gin.cpu().numpy(), rin) < 1e-4
    assert norm_err(glut.cpu().numpy(), rlut) < 1e-4


def test_lut_backward_normalized_and_batches(cuda):
    sc, data, _ = _setup(1, seed=19)
    data2 = np.concatenate([data, 0.3 * data], axis=0)
    lut = _lut((5, 6, 7), 3, seed=20, normalized=True)
    rng = np.random.Generator(np.random.PCG64(21))
    og = rng.uniform(-1, 1, size=(2, 3, 5, 6, 7, 1)).astype(np.float32)
    for gm in ("REFERENCE", "EXACT"):
        kw = dict(interpolation="LINEAR", boundary="BORDER", relative_coords=False, normalized_coords=True,
                  cell_center_offset=0.5, grad_mode=gm)
        gin, glut = _ops().sample_grid_lut_grad(torch.from_numpy(data2).cuda(), torch.from_numpy(og).cuda(),
                                                torch.from_numpy(lut).cuda(), **kw)
        rin, rlut = orc.sample_lut_bwd(data2, lut, og, **kw)
        assert norm_err(gin.cpu().numpy(), rin) < 1e-4
        assert norm_err(glut.cpu().numpy(), rlut) < 1e-4


def test_lut_autograd(cuda):
    sc, data, _ = _setup(1, seed=22)
    lut = _lut((6, 7, 8), 1, seed=23)
    x = torch.from_numpy(data).cuda().requires_grad_(True)
    lt = torch.from_numpy(lut).cuda().requires_grad_(True)
    y = _ops().sample_grid_lut(x, lt, boundary="CLAMP", grad_mode="EXACT")
    rng = np.random.Generator(np.random.PCG64(24))
    og = rng.uniform(-1, 1, size=tuple(y.shape)).astype(np.float32)
    y.backward(torch.from_numpy(og).cuda())
    rin, rlut = orc.sample_lut_bwd(data, lut, og, boundary="CLAMP", grad_mode="EXACT")
    assert norm_err(x.grad.cpu().numpy(), rin) < 1e-4
    assert norm_err(lt.grad.cpu().numpy(), rlut) < 1e-4


def test_build_mips(cuda):
    rng = np.random.Generator(np.random.PCG64(25))
    data = rng.uniform(0, 1, size=(2, 13, 18, 22, 4)).astype(np.float32)
    levels = _ops().build_mips(torch.from_numpy(data).cuda(), 5)
    assert len(levels) == 3  # 13 >> 4 == 0
    for b in range(2):
        ref = orc.build_mips(data[b], 3)
        for lv, r in zip(levels, ref):
            assert norm_err(lv[b].cpu().numpy(), r) < 1e-6


@pytest.mark.parametrize("inverse", [False, True])
def test_lod_transform(cuda, inverse):
    sc, data, out_dhw = _setup(1, seed=26)
    MV = np.stack([sc["V"][i] @ sc["M"][0] for i in range(len(sc["V"]))]).astype(np.float32)
    in_dhw = data.shape[1:4] if not inverse else out_dhw
    o_dhw = out_dhw if not inverse else data.shape[1:4]
    got = _ops().lod_transform(in_dhw, MV, sc["P"], sc["F"], o_dhw, inverse_transform=inverse).cpu().numpy()
    ref = orc.lod_transform(in_dhw, MV, sc["P"], sc["F"], o_dhw, inverse_transform=inverse)
    assert got.shape == ref.shape
    assert norm_err(got[..., :3], ref[..., :3]) < 1e-5
    assert np.abs(got[..., 3] - ref[..., 3]).max() < 1e-3


def test_invalid_arguments(cuda):
    from gft._lib import InvalidArgumentError, UnimplementedError
    sc, data, out_dhw = _setup(1, seed=27)
    x = torch.from_numpy(data).cuda()
    ops = _ops()
    with pytest.raises(InvalidArgumentError):
        ops.sample_grid_transform(x, sc["M"], sc["V"], sc["P"], sc["F"], out_dhw, boundary="WRAP")
    with pytest.raises(InvalidArgumentError):
        ops.sample_grid_transform(x, sc["M"], sc["V"], sc["P"], sc["F"], out_dhw, cell_center_offset=0.0)
    with pytest.raises(InvalidArgumentError):
        ops.sample_grid_transform(x, sc["M"], sc["V"], sc["P"], sc["F"], out_dhw, mipmapping="LINEAR", num_mipmaps=0)
    with pytest.raises(UnimplementedError):
        ops.sample_grid_transform(torch.zeros((1, 4, 4, 4, 3), device="cuda"), sc["M"], sc["V"], sc["P"], sc["F"],
                                  out_dhw)
    lut = torch.zeros((1, 3, 3, 3, 4), device="cuda")
    with pytest.raises(InvalidArgumentError):
        ops.sample_grid_lut(x, lut, normalized_coords=True, cell_center_offset=0.0)
    with pytest.raises(InvalidArgumentError):
        ops.sample_grid_lut(x, torch.zeros((1, 3, 3, 3, 3), device="cuda"))


@pytest.mark.parametrize("yaw", [0.0, 90.0])
@pytest.mark.parametrize("coord", ["TRANSFORM", "TRANSFORM_LINDEPTH"])
def test_transform_orthographic_axis_aligned(cuda, yaw, coord):
    """Frustum rows along a grid axis (orthographic camera at yaw 0 / 90):
    LINEAR forward and EXACT backward (the REFERENCE count is tie-sensitive at
    exact grid points and is covered by the generic cases)."""
    from gft import synthetic as syn
    from gft.camera import camera_params_batch, model_matrices
    from helpers import ortho_camera
    grid = (9, 11, 8)
    cam = ortho_camera(grid, (11, 8), 10, yaw=yaw)
    V, P, F = camera_params_batch([cam])
    M = model_matrices([syn.grid_transform(grid)])
    data = syn.gaussian_blobs(grid, 4, seed=6).astype(np.float32)
    kw = dict(interpolation="LINEAR", boundary="BORDER", coordinate_mode=coord)
    t = torch.from_numpy(data).cuda()
    got = _ops().sample_grid_transform(t, M, V, P, F, (10, 11, 8), **kw).cpu().numpy()
    ref = orc.sample_transform_fwd(data, M, V, P, F, (10, 11, 8), **kw)
    assert np.abs(ref).max() > 0.05
    assert norm_err(got, ref) < 1e-5
    og = np.random.Generator(np.random.PCG64(3)).uniform(-1, 1, size=ref.shape).astype(np.float32)
    g = _ops().sample_grid_transform_grad(t, torch.from_numpy(og).cuda(), M, V, P, F, grad_mode="EXACT",
                                          **kw).cpu().numpy()
    gref = orc.sample_transform_bwd(data, og, M, V, P, F, grad_mode="EXACT", **kw)
    assert norm_err(g, gref) < 1e-4


@pytest.mark.parametrize("grad_mode,filt", [("EXACT", "NEAREST"), ("EXACT", "LINEAR"), ("REFERENCE", "LINEAR")])
@pytest.mark.parametrize("views,sep", [(11, True), (2, False)])
def test_transform_backward_gather_form(cuda, grad_mode, filt, views, sep):
    """BORDER without mips runs the gradient in gather form: eleven views
    (EXACT adds over two launches, REFERENCE takes the counted scatter because
    its counts span all views of a batch), a camera per batch, NEAREST taps
    (against the FP32 oracle: index rounding at exact halves)."""
    sc = scene(grid=(10, 18, 14), image=(12, 9), views=views, seed=12, channels=2)
    data, M = sc["data"], sc["M"]
    if not sep:
        data = np.concatenate([data, 0.7 * data[:, ::-1]], axis=0).astype(np.float32)
        M = np.concatenate([M, M], axis=0)
    rng = np.random.Generator(np.random.PCG64(13))
    og = rng.uniform(-1, 1, size=(data.shape[0], views if sep else 1) + tuple(sc["out_dhw"]) + (2,)).astype(np.float32)
    kw = dict(interpolation=filt, boundary="BORDER", coordinate_mode="TRANSFORM_LINDEPTH", separate_camera_batch=sep)
    got = _ops().sample_grid_transform_grad(torch.from_numpy(data).cuda(), torch.from_numpy(og).cuda(), M, sc["V"],
                                            sc["P"], sc["F"], grad_mode=grad_mode, **kw).cpu().numpy()
    ref = orc.sample_transform_bwd(data, og, M, sc["V"], sc["P"], sc["F"], grad_mode=grad_mode,
                                   prec=0 if filt == "NEAREST" else 1, **kw)
    assert np.abs(ref).max() > 0
    assert norm_err(got, ref) < 1e-4


@pytest.mark.parametrize("grad_mode", ["REFERENCE", "EXACT"])
def test_transform_backward_camera_inside_grid(cuda, grad_mode):
    """A camera inside the volume: voxel blocks behind it have corner images
    behind the camera, and the gather walks every frustum cell for them."""
    from gft.camera import Camera, camera_params_batch, model_matrices
    from gft import synthetic as syn
    from gft.transform import GridTransform
    grid = (12, 14, 16)
    rng = np.random.Generator(np.random.PCG64(14))
    data = rng.uniform(0, 1, size=(1,) + grid + (1,)).astype(np.float32)
    cam = Camera(GridTransform([10, 9, 11], translation=[0.5, -0.5, 1.0]), nearFar=[0.5, 9.0], fov=70.0, aspect=11 / 9)
    V, P, F = camera_params_batch([cam])
    M = model_matrices([syn.grid_transform(grid)])
    og = rng.uniform(-1, 1, size=(1, 1, 10, 9, 11, 1)).astype(np.float32)
    kw = dict(interpolation="LINEAR", boundary="BORDER", coordinate_mode="TRANSFORM_LINDEPTH")
    got = _ops().sample_grid_transform_grad(torch.from_numpy(data).cuda(), torch.from_numpy(og).cuda(), M, V, P, F,
                                            grad_mode=grad_mode, **kw).cpu().numpy()
    ref = orc.sample_transform_bwd(data, og, M, V, P, F, grad_mode=grad_mode, **kw)
    assert np.abs(ref).max() > 0 and (np.abs(ref) > 0).mean() > 0.05
    assert norm_err(got, ref) < 1e-4


def test_transform_backward_gather_adjoint_identity(cuda):
    """<A x, y> == <x, A^T y> for the gather-form EXACT gradient."""
    sc, data, out_dhw = _setup(4, seed=15, views=4)
    rng = np.random.Generator(np.random.PCG64(16))
    y = rng.uniform(-1, 1, size=(1, 4) + tuple(out_dhw) + (4,)).astype(np.float32)
    x = torch.from_numpy(data).cuda().requires_grad_(True)
    ax = _ops().sample_grid_transform(x, sc["M"], sc["V"], sc["P"], sc["F"], out_dhw, grad_mode="EXACT",
                                      boundary="BORDER", coordinate_mode="TRANSFORM")
    ax.backward(torch.from_numpy(y).cuda())
    lhs = float((ax.detach().double().cpu() * torch.from_numpy(y).double()).sum())
    rhs = float((x.detach().double().cpu() * x.grad.double().cpu()).sum())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


@pytest.mark.parametrize("grad_mode", ["REFERENCE", "EXACT"])
@pytest.mark.parametrize("filt", ["NEAREST", "LINEAR", "MAX"])
@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP", "WRAP"])
@pytest.mark.parametrize("spread", [1.5, 8.0])
def test_lut_backward_one_channel_tiled(cuda, grad_mode, filt, boundary, spread):
    """One-channel relative lookups take the tiled gradient (LDS region of the
    tile plus a 3-cell margin; taps beyond it, as with spread 8, go to global
    atomics); EXACT MAX keeps the untiled kernel."""
    sc, data, _ = _setup(1, seed=22)
    lut = _lut((9, 13, 40), 2, seed=23, spread=spread)
    rng = np.random.Generator(np.random.PCG64(24))
    og = rng.uniform(-1, 1, size=(1, 2, 9, 13, 40, 1)).astype(np.float32)
    kw = dict(interpolation=filt, boundary=boundary, grad_mode=grad_mode)
    gin, glut = _ops().sample_grid_lut_grad(torch.from_numpy(data).cuda(), torch.from_numpy(og).cuda(),
                                            torch.from_numpy(lut).cuda(), **kw)
    p = 0 if filt != "LINEAR" else 1
    rin, rlut = orc.sample_lut_bwd(data, lut, og, prec=p, **kw)
    assert np.abs(rin).max() > 0
    assert norm_err(gin.cpu().numpy(), rin) < 1e-4
    assert norm_err(glut.cpu().numpy(), rlut) < 1e-4
