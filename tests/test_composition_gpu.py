"""Parity of the Python compositions on the HIP ops with the oracle:
density/staggered-velocity grids and their warps (A10), the Renderer subset
with ambient, point and spot lights (A11), and the checkpointed transport
chain (A12).

Tolerances (normalised max error vs the FP64 oracle): forward fields and
images 1e-5 (warps with MC_SMOOTH may flip a clamp decision at exact ties,
at most 0.2 % of cells), gradients 1e-3.
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


def _stag(grid, seed, speed=1.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    D, H, W = grid
    shapes = [(1, D, H, W + 1, 1), (1, D, H + 1, W, 1), (1, D + 1, H, W, 1)]
    # smooth-ish components: blobs scaled to the speed
    comps = []
    for i, s in enumerate(shapes):
        base = syn.gaussian_blobs(s[1:4], 4, seed=seed + i)[..., :1]
        comps.append((speed * (2 * base / max(base.max(), 1e-6) - 1) * rng.uniform(0.5, 1.0)).astype(np.float32))
    return comps


def _close_ties(got, ref, tol=1e-5, frac=2e-3):
    scale = max(np.abs(ref).max(), 1e-30)
    bad = np.abs(got - ref) > tol * scale
    assert bad.mean() <= frac, bad.mean()


@pytest.mark.parametrize("grad_mode", ["EXACT", "REFERENCE"])
@pytest.mark.parametrize("order,clamp", [(1, "NONE"), (2, "MC_SMOOTH")])
def test_velocity_self_advection(cuda, grad_mode, order, clamp):
    from gft.grids import VelocityGrid
    grid = (10, 14, 12)
    comps = _stag(grid, 3)
    vg = VelocityGrid(grid, x=_t(comps[0]), y=_t(comps[1]), z=_t(comps[2]), grad_mode=grad_mode)
    for c in vg.variables():
        c.requires_grad_(True)
    out = vg.warped(order=order, dt=1.0, clamp=clamp)
    ref = orc.velocity_self_advect(comps, order, clamp)
    for o, r in zip(out, ref):
        _close_ties(o.detach().cpu().numpy(), r)
    if grad_mode == "EXACT":
        rng = np.random.Generator(np.random.PCG64(4))
        ws = [rng.uniform(-1, 1, size=r.shape) for r in ref]
        torch.autograd.backward(list(out), [_t(w) for w in ws])
        gref = orc.velocity_self_advect_bwd(comps, ws, order, clamp)
        for c, g in zip(vg.variables(), gref):
            assert norm_err(c.grad.cpu().numpy(), g) < 1e-3


@pytest.mark.parametrize("grad_mode", ["EXACT", "REFERENCE"])
def test_density_warp_by_staggered_velocity(cuda, grad_mode):
    from gft.grids import DensityGrid, VelocityGrid
    grid = (10, 14, 12)
    comps = _stag(grid, 7, speed=1.4)
    d = syn.gaussian_blobs(grid, 5, seed=8)
    vg = VelocityGrid(grid, x=_t(comps[0]), y=_t(comps[1]), z=_t(comps[2]), grad_mode=grad_mode)
    for c in vg.variables():
        c.requires_grad_(True)
    dens = DensityGrid(grid, d=_t(d).requires_grad_(True))
    out = dens.warped(vg, order=2, dt=1.0, clamp="MC_SMOOTH")
    vc = np.concatenate([0.5 * (comps[0][:, :, :, 1:] + comps[0][:, :, :, :-1]),
                             0.5 * (comps[1][:, :, 1:] + comps[1][:, :, :-1]),
                             0.5 * (comps[2][:, 1:] + comps[2][:, :-1])], axis=-1)
    ref = orc.advect_fwd(d, vc, 2, "MC_SMOOTH")
    _close_ties(out.detach().cpu().numpy(), ref)
    rng = np.random.Generator(np.random.PCG64(9))
    w = rng.uniform(-1, 1, size=ref.shape)
    out.backward(_t(w))
    gd, gv = orc.advect_bwd(d, vc, w, 2, "MC_SMOOTH", grad_mode=grad_mode)
    assert norm_err(dens.d.grad.cpu().numpy(), gd) < 1e-3
    # centring adjoint: each face receives half of each adjacent cell's gradient
    gx = np.zeros_like(comps[0], dtype=np.float64)
    gx[:, :, :, 1:] += 0.5 * gv[..., 0:1]
    gx[:, :, :, :-1] += 0.5 * gv[..., 0:1]
    assert norm_err(vg.x.grad.cpu().numpy(), gx) < 1e-3


def test_sequence_advection_modes(cuda):
    from gft.grids import DensityGrid, Sequence, State, VelocityGrid
    grid = (8, 12, 10)
    d0 = syn.gaussian_blobs(grid, 4, seed=11) - 0.05
    vels = [_stag(grid, 20 + i, speed=1.2) for i in range(3)]
    for mode in ("NONE", "LOCAL", "GLOBAL", "NEGATIVE"):
        seq = Sequence()
        for i in range(3):
            vg = VelocityGrid(grid, x=_t(vels[i][0]), y=_t(vels[i][1]), z=_t(vels[i][2]))
            seq.append(State(DensityGrid(grid, d=_t(d0) if i == 0 else None), vg, i))
        seq.densities_advect_fwd(dt=1.0, order=2, clamp=mode)
        ref = d0.astype(np.float64)
        refs = [ref]
        for i in range(1, 3):
            vc = np.concatenate([0.5 * (vels[i - 1][0][:, :, :, 1:] + vels[i - 1][0][:, :, :, :-1]),
                                 0.5 * (vels[i - 1][1][:, :, 1:] + vels[i - 1][1][:, :, :-1]),
                                 0.5 * (vels[i - 1][2][:, 1:] + vels[i - 1][2][:, :-1])], axis=-1)
            nxt = orc.advect_fwd(refs[-1], vc, 2, "NEGATIVE" if mode == "NEGATIVE" else "NONE")
            if mode == "LOCAL":
                nxt = np.maximum(nxt, 0)
            refs.append(nxt)
        if mode == "GLOBAL":
            refs = [refs[0]] + [np.maximum(r, 0) for r in refs[1:]]
        for i in range(1, 3):
            assert norm_err(seq[i].density.d.cpu().numpy(), refs[i]) < 1e-5, mode


def _render_setup(seed=30):
    from gft.camera import Camera
    from gft.transform import GridTransform
    grid, image = (12, 16, 14), (14, 10)
    sc = scene(grid=grid, image=image, views=3, seed=seed)
    gt = syn.grid_transform(grid)
    gt.set_data(_t(sc["data"]))
    return sc, gt, sc["cams"]


@pytest.mark.parametrize("fused", [True, False])
def test_renderer_ambient(cuda, fused):
    from gft.lighting import Light
    from gft.renderer import Renderer
    sc, gt, cams = _render_setup()
    r = Renderer(filter_mode="LINEAR", boundary_mode="BORDER", mipmapping="NONE", blend_mode="BEER_LAMBERT",
                 fused=fused)
    lights = [Light([0.9, 1.0, 1.1], 0.8), Light([1, 1, 1], 0.3)]
    bkg = [np.full((3,), 0.2 * (i + 1), dtype=np.float32) for i in range(3)]
    imgs = r.render_density(gt, lights, cams, cut_alpha=True, background=bkg)
    light = (np.array([0.9, 1.0, 1.1]) * 0.8 + 0.3)
    d = sc["data"].astype(np.float64)
    if fused:
        ref = orc.render_density_fwd(d, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], light)[0]
    else:
        grid4 = np.concatenate([d * light, d], -1) * sc["step"]
        fr = orc.sample_transform_fwd(grid4, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], "LINEAR", "BORDER")
        ref = orc.reduce_blend_fwd(fr[0], "BEER_LAMBERT")
    for i in range(3):
        expect = ref[i, ..., :3] + bkg[i] * np.exp(-ref[i, ..., 3:])
        assert norm_err(imgs[i][0].cpu().numpy(), expect) < 1e-5


def test_renderer_point_light_fused_raymarch_and_gather_gradient(cuda):
    from gft.lighting import PointLight
    from gft.renderer import Renderer
    from gft.transform import Transform
    sc, gt, cams = _render_setup(31)
    pl = PointLight(Transform(translation=[3.0, 10.0, -4.0]), color=[1.0, 0.8, 0.6], intensity=2.0, range_scale=0.05)
    r = Renderer(boundary_mode="CLAMP", mipmapping="NONE", blend_mode="BEER_LAMBERT", fused=True, grad_mode="EXACT")
    imgs = r.render_density(gt, [pl], cams, cut_alpha=False)
    light = pl._falloff(gt).cpu().numpy().astype(np.float64)
    d = sc["data"].astype(np.float64)
    grid4 = np.concatenate([d * light, d], -1) * sc["step"]
    ref = orc.raymarch_fwd(grid4, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], "LINEAR", "CLAMP")[0]
    for i in range(3):
        assert norm_err(imgs[i][0].cpu().numpy(), ref[i]) < 1e-5
    # sample_camera with the renderer's gather gradient
    rs = Renderer(boundary_mode="BORDER", mipmapping="LINEAR", num_mips=2, mip_bias=0.2, sample_gradients=True,
                  fast_gradient_mip_bias_add=0.5)
    x = _t(sc["data"]).requires_grad_(True)
    fr = rs.sample_camera(x, gt, cams)
    rng = np.random.Generator(np.random.PCG64(5))
    dy = rng.uniform(-1, 1, size=tuple(fr.shape))
    fr.backward(_t(dy))
    gref = orc.sample_transform_gather_grad(dy[None], sc["M"], sc["V"], sc["P"], sc["F"], sc["data"].shape[1:4],
                                            "LINEAR", "BORDER", "LINEAR", 2, 0.7, "TRANSFORM_LINDEPTH")
    assert norm_err(x.grad.cpu().numpy(), gref) < 1e-5


def test_spot_light_shadow_map(cuda):
    from gft.lighting import SpotLight
    from gft.renderer import Renderer
    from gft.transform import Transform
    sc, gt, cams = _render_setup(32)
    t = Transform(translation=[7.0, 30.0, 7.0], rotation_deg=[-90.0, 0.0, 0.0])
    sl = SpotLight(t, intensity=1.5, range_scale=0.01, angle_deg=60, cast_shadows=True,
                   shadow_resolution=[24, 20, 20], shadow_clip=[10, 40], cone_mask=True)
    r = Renderer(boundary_mode="BORDER", mipmapping="NONE", blend_mode="BEER_LAMBERT")
    trans = sl.render_shadow_map(gt, r)[0].cpu().numpy()
    # oracle composition: frustum resample (step-scaled), prefix sum, shift, exp, cone mask, reverse resample
    from gft.camera import camera_params_batch
    Vs, Ps, Fs = camera_params_batch([sl.shadow_cam])
    d = sc["data"].astype(np.float64) * sl.shadow_cam.depth_step
    fr = orc.sample_transform_fwd(d, sc["M"], Vs, Ps, Fs, sl.shadow_resolution, "LINEAR", "BORDER")[0]
    pre = orc.reduce_blend_fwd(fr, "BEER_LAMBERT", keep_dims=True)
    pre = np.concatenate([np.zeros_like(pre[:, :1]), pre[:, :-1]], axis=1)
    tr = np.exp(-pre)
    a = np.linspace(-1, 1, 20)
    mask = (np.sqrt(a[:, None] ** 2 + a[None, :] ** 2) <= 1).astype(np.float64)
    tr = tr * mask[None, None, :, :, None]
    ref = orc.sample_transform_fwd(tr, sc["M"], Vs, Ps, Fs, sc["data"].shape[1:4], "LINEAR", "BORDER",
                                   coordinate_mode="TRANSFORM_LINDEPTH_REVERSE")[0, 0]
    assert np.abs(ref).max() > 0.1
    assert norm_err(trans, ref) < 1e-5


def _chain_inputs(T, seed=40):
    grid = (10, 16, 12)
    sc = scene(grid=grid, image=(12, 9), views=2, seed=seed)
    d0 = syn.gaussian_blobs(grid, 5, seed=seed + 1)
    vels = [syn.curl_noise_velocity(grid, seed=seed + 2 + f, max_speed=1.5, periods=(8, 4, 2)) for f in range(T - 1)]
    rng = np.random.Generator(np.random.PCG64(seed + 50))
    tg = [rng.uniform(0, 0.3, size=(1, 2) + tuple(sc["out_dhw"][1:]) + (3,)).astype(np.float32) for _ in range(T)]
    return grid, sc, d0, vels, tg


@pytest.mark.parametrize("k", [1, 2, 5])
@pytest.mark.parametrize("grad_mode,decay", [("EXACT", 0.0), ("REFERENCE", 0.9), ("REFERENCE", 0.0)])
def test_transport_chain(cuda, k, grad_mode, decay):
    from gft.chain import TransportChain
    T = 5
    grid, sc, d0, vels, tg = _chain_inputs(T)
    ch = TransportChain(grid, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], checkpoint_every=k,
                        grad_mode=grad_mode, decay=decay)
    loss, gd0, gv = ch.run(_t(d0), [_t(v) for v in vels], [_t(t) for t in tg])
    r = orc.chain_fwd_bwd(d0, vels, tg, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"],
                          grad_mode=grad_mode, decay=decay)
    assert abs(loss.item() - r["loss"]) <= 1e-5 * r["loss"]
    assert norm_err(gd0.cpu().numpy(), r["gd0"]) < 1e-3
    for f in range(T - 1):
        assert norm_err(gv[f].cpu().numpy(), r["gv"][f]) < 1e-3


def test_transport_chain_checkpointing_saves_memory(cuda):
    from gft.chain import TransportChain
    T = 8
    grid = (32, 48, 40)
    sc = scene(grid=grid, image=(24, 20), views=2, seed=44)
    d0 = _t(syn.gaussian_blobs(grid, 8, seed=45))
    vels = [_t(syn.curl_noise_velocity(grid, seed=46 + f, max_speed=1.5)) for f in range(T - 1)]
    tg = [torch.zeros((1, 2, 24, 20, 3), device="cuda") for _ in range(T)]
    peaks, grads = {}, {}
    for k in (T, 2):
        ch = TransportChain(grid, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], checkpoint_every=k)
        torch.cuda.synchronize()
        base = torch.cuda.memory_allocated()
        torch.cuda.reset_peak_memory_stats()
        loss, gd0, gv = ch.run(d0, vels, tg)
        torch.cuda.synchronize()
        peaks[k] = torch.cuda.max_memory_allocated() - base
        grads[k] = gd0.clone()
        del ch, loss, gv
    assert torch.allclose(grads[2], grads[T], rtol=0, atol=1e-6 * float(grads[T].abs().max()))
    # the velocity gradients (outputs, (T-1) x 12 B per voxel) dominate both peaks; the
    # recomputed densities and MacCormack states make the difference
    assert peaks[2] < peaks[T], peaks
