"""The fused spot-lit density render (a light with a per-cell factor plus an
ambient light of the same colour, BORDER).

The op (gft_render_density_lit_fwd / _bwd through render_density_lit)
against the FP64 oracle composition render_lit_fwd / render_lit_bwd (the
pinned raymarch of the materialised [color d (lambda + a), d] * step grid and
its exact gradient): images within 1e-5, gradients w.r.t. the density and
the light factor within 1e-4 (normalised max error). Also more than 16 views
(two launches whose gathers accumulate) and the folded image loss.

Renderer.render_density with [SpotLight (shadows, LINEAR mips, gather
sampling gradient), Light (ambient)]: the fused lit path against the same
renderer's generic path (the light grid materialised by the same shadow and
falloff ops, then the 4-channel raymarch op with its exact gradient): images
within 1e-5, the density gradient through the render AND the shadow map
within 1e-4.
"""
import numpy as np
import pytest
import torch

from gft import synthetic as syn
from helpers import norm_err, scene
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


@pytest.mark.parametrize("views,color,ambient", [(3, (1.0, 1.0, 1.0), 0.64), (2, (0.9, 0.5, 0.2), 0.0),
                                                 (18, (1.0, 0.8, 0.6), 0.3)])
def test_lit_render_matches_oracle(cuda, views, color, ambient):
    from gft.pipeline import render_density_lit
    grid = (14, 20, 16)
    sc = scene(grid=grid, image=(16, 12), views=views, seed=70 + views, yaw_range=(-170.0, 170.0))
    rng = np.random.Generator(np.random.PCG64(views))
    lam = rng.uniform(0.0, 1.5, size=sc["data"].shape).astype(np.float32)
    d = _t(sc["data"]).requires_grad_(True)
    lt = _t(lam).requires_grad_(True)
    img = render_density_lit(d, lt, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], color, ambient)
    w = torch.linspace(-1, 1, img.numel(), device="cuda").reshape(img.shape)
    (img * w).sum().backward()
    torch.cuda.synchronize()
    args = (sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"])
    ref = orc.render_lit_fwd(sc["data"], lam, *args, color=color, ambient=ambient)
    assert norm_err(img.detach().cpu().numpy(), ref) < 1e-5
    gd, gl = orc.render_lit_bwd(sc["data"], lam, ref, w.cpu().numpy(), *args, color=color, ambient=ambient)
    assert norm_err(d.grad.cpu().numpy(), gd) < 1e-4
    assert norm_err(lt.grad.cpu().numpy(), gl) < 1e-4


def test_lit_render_folded_loss(cuda):
    """gft_render_density_lit_bwd_loss (AE loss with a background) against the
    oracle's image loss chained through the lit render gradient."""
    import ctypes
    from gft import _lib
    from gft.ops import _np_ptr
    from gft.pipeline import _Cameras, render_lit_forward, render_lit_saved_bytes
    grid = (12, 18, 14)
    sc = scene(grid=grid, image=(14, 10), views=3, seed=77)
    rng = np.random.Generator(np.random.PCG64(77))
    lam = rng.uniform(0.2, 1.2, size=sc["data"].shape).astype(np.float32)
    color, ambient = (1.0, 0.9, 0.7), 0.5
    cams = _Cameras(sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"])
    d, lt = _t(sc["data"]), _t(lam)
    saved = torch.empty((render_lit_saved_bytes(d.shape, cams),), dtype=torch.uint8, device="cuda")
    img = render_lit_forward(d, lt, cams, sc["step"], color, ambient, saved=saved)
    args = (sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"])
    ref = orc.render_lit_fwd(sc["data"], lam, *args, color=color, ambient=ambient)
    tgt = (ref[..., :3] + rng.uniform(-0.05, 0.05, size=ref[..., :3].shape)).astype(np.float32)
    bkg = rng.uniform(0, 1, size=tgt.shape).astype(np.float32)
    scale = 0.37
    gd, gl, loss = torch.empty_like(d), torch.empty_like(lt), torch.zeros((1,), device="cuda")
    tt, tb = _t(tgt), _t(bkg)  # held for the call
    lc = (ctypes.c_float * 3)(*color)
    _lib.check(_lib.lib().gft_render_density_lit_bwd_loss(
        _lib.stream_of(d), _lib.ptr(d), _lib.ptr(lt), _lib.i64_array(d.shape), _np_ptr(cams.M), cams.nM,
        _np_ptr(cams.V), _np_ptr(cams.P), _np_ptr(cams.F), cams.nCam, cams.c_out, lc, ambient, float(sc["step"]), 0,
        _lib.ptr(img), _lib.ptr(tt), _lib.ptr(tb), 1, scale, _lib.ptr(saved), saved.numel(), _lib.ptr(gd),
        _lib.ptr(gl), _lib.ptr(loss)))
    torch.cuda.synchronize()
    lref, gimg = orc.image_loss_and_grad(ref, tgt, "AE", bkg, scale)
    assert abs(loss.item() - lref) <= 1e-5 * abs(lref)
    gdr, glr = orc.render_lit_bwd(sc["data"], lam, ref, gimg, *args, color=color, ambient=ambient)
    assert norm_err(gd.cpu().numpy(), gdr) < 1e-4
    assert norm_err(gl.cpu().numpy(), glr) < 1e-4


def _lit_scene():
    from gft.lighting import Light, SpotLight
    from gft.transform import Transform
    grid = (20, 28, 24)
    sc = scene(grid=grid, image=(22, 16), views=4, seed=81)
    t = Transform(translation=[0, 0, 60], parent=Transform(rotation_deg=[-40, 0, 0]))
    spot = SpotLight(t, intensity=0.85, cast_shadows=True, shadow_clip=[25, 95], range_scale=0.02, angle_deg=50.0,
                     shadow_resolution=[48, 36, 36], cone_mask=False)
    return grid, sc, [spot, Light(intensity=0.64)]


@pytest.mark.parametrize("cone", [False, True])
def test_renderer_spot_and_ambient_fused_equals_generic(cuda, cone):
    from gft.renderer import Renderer
    grid, sc, lights = _lit_scene()
    lights[0].cone_mask = cone
    cams = sc["cams"]
    res = {}
    for fused_lit in (True, False):
        r = Renderer(filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="LINEAR", num_mips=3, mip_bias=0.0,
                     blend_mode="BEER_LAMBERT", sample_gradients=True, grad_mode="EXACT", fused=True,
                     fused_lit=fused_lit)
        gt = syn.grid_transform(grid)
        d = _t(sc["data"] * 0.3).requires_grad_(True)
        gt.set_data(d)
        imgs = r.render_density(gt, lights, cams, cut_alpha=False)
        img = torch.cat(imgs, 0)
        w = torch.linspace(-1, 1, img.numel(), device="cuda").reshape(img.shape)
        (img * w).sum().backward()
        torch.cuda.synchronize()
        res[fused_lit] = (img.detach().cpu().numpy(), d.grad.cpu().numpy())
    assert norm_err(res[True][0], res[False][0]) < 1e-5
    assert np.abs(res[True][1]).max() > 0
    assert norm_err(res[True][1], res[False][1]) < 1e-4


def test_renderer_lit_images_match_oracle_from_its_light_factor(cuda):
    """The fused lit images against the FP64 oracle render of the same
    density and light factor (the shadow map itself is checked against the
    oracle in the shadow tests)."""
    from gft.renderer import Renderer
    grid, sc, lights = _lit_scene()
    r = Renderer(filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="LINEAR", num_mips=3, mip_bias=0.0,
                 blend_mode="BEER_LAMBERT", sample_gradients=True, fused=True)
    gt = syn.grid_transform(grid)
    gt.set_data(_t(sc["data"] * 0.3))
    imgs = torch.cat(r.render_density(gt, lights, sc["cams"], cut_alpha=False), 0)
    lam = (lights[0].light_factor(gt, r) * 0.85).cpu().numpy()
    torch.cuda.synchronize()
    ref = orc.render_lit_fwd(sc["data"] * 0.3, lam, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"],
                             ambient=0.64)
    assert norm_err(imgs.cpu().numpy()[None], ref) < 1e-5


def test_lit_render_rejects_clamp_and_shape_mismatch(cuda):
    import ctypes
    from gft import _lib
    from gft.ops import _np_ptr
    from gft.pipeline import _Cameras, render_density_lit
    sc = scene(grid=(8, 10, 9), image=(10, 7), views=2, seed=4)
    d = _t(sc["data"])
    with pytest.raises(_lib.InvalidArgumentError):
        render_density_lit(d, _t(np.ones((1, 8, 10, 8, 1))), sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], 1.0)
    cams = _Cameras(sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"])
    out = torch.empty((1, 2, 10, 7, 4), device="cuda")
    lc = (ctypes.c_float * 3)(1.0, 1.0, 1.0)
    rc = _lib.lib().gft_render_density_lit_fwd(_lib.stream_of(d), _lib.ptr(d), _lib.ptr(d), _lib.i64_array(d.shape),
                                               _np_ptr(cams.M), cams.nM, _np_ptr(cams.V), _np_ptr(cams.P),
                                               _np_ptr(cams.F), cams.nCam, cams.c_out, lc, 0.5, 1.0, 1, None, 0,
                                               None, 0, _lib.ptr(out))
    assert rc == _lib.GFT_ERR_UNIMPLEMENTED


@pytest.mark.parametrize("lit", [False, True])
def test_render_backward_twice_on_a_retained_graph(cuda, lit):
    """The render backward rewrites the per-pixel records of its saved buffer
    in place (the per-sample values are only read); the autograd function
    keeps a copy of that part only, so a retained graph's second backward
    restores it and gives the same gradients bit for bit, for the ambient
    render (gft_render_density_bwd) and the lit one."""
    from gft.pipeline import render_density, render_density_lit
    sc = scene(grid=(14, 20, 16), image=(16, 12), views=3, seed=91, yaw_range=(-170.0, 170.0))
    args = (sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"])
    d = _t(sc["data"]).requires_grad_(True)
    lt = _t(np.random.Generator(np.random.PCG64(91)).uniform(0.0, 1.5, size=sc["data"].shape)).requires_grad_(True)
    img = render_density_lit(d, lt, *args, (1.0, 0.8, 0.6), 0.3) if lit else render_density(d, *args)
    w = torch.linspace(-1, 1, img.numel(), device="cuda").reshape(img.shape)
    loss = (img * w).sum()
    loss.backward(retain_graph=True)
    g1 = d.grad.clone()
    l1 = lt.grad.clone() if lit else None
    d.grad = None
    lt.grad = None
    loss.backward()
    torch.cuda.synchronize()
    assert torch.equal(g1, d.grad)
    assert float(g1.abs().max()) > 0.0
    if lit:
        assert torch.equal(l1, lt.grad)
