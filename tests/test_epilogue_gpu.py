"""Loss and optimizer epilogue on the GPU against the FP64 oracle: the image
losses folded into the fused render backward (SE / AE, with and without a
composited background), the divergence loss of a staggered velocity, and the
multi-tensor Adam update of tf.train.AdamOptimizer.

Tolerances: losses 1e-5 relative; gradients 1e-3 (fused step) and 1e-5
(divergence) normalised max error; Adam parameters 1e-6 relative. The AE
targets sit 0.05-0.2 away from the rendered images so no error changes sign
between FP32 and FP64.
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


def _fields(grid, seed):
    d = syn.gaussian_blobs(grid, 5, seed=seed)
    v = syn.curl_noise_velocity(grid, seed=seed + 1, max_speed=1.5)
    return d, v


@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP"])
@pytest.mark.parametrize("loss", ["SE", "AE"])
@pytest.mark.parametrize("bkg", [None, (0.0, 0.5, 1.0)])
def test_fused_step_image_losses(cuda, boundary, loss, bkg):
    from gft.pipeline import TransportRenderStep
    grid = (16, 26, 18)
    sc = scene(grid=grid, image=(24, 14), views=3, seed=8)
    d, v = _fields(grid, seed=10)
    nV, h, w = 3, sc["out_dhw"][1], sc["out_dhw"][2]
    background = None if bkg is None else np.broadcast_to(np.asarray(bkg, dtype=np.float64), (1, nV, h, w, 3))
    # targets: the oracle's own composite image shifted by +-(0.05..0.2) per element
    r0 = orc.step_fwd_bwd(d, v, np.zeros((1, nV, h, w, 3)), sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"],
                          sc["step"], boundary_render=boundary, loss_kind="SE", background=background)
    comp = r0["img"][..., :3] + (0 if background is None else background * np.exp(-r0["img"][..., 3:4]))
    rng = np.random.Generator(np.random.PCG64(4))
    tgt = (comp + rng.choice([-1.0, 1.0], size=comp.shape) * rng.uniform(0.05, 0.2, size=comp.shape))
    tgt = tgt.astype(np.float32)
    scale = 0.5 / (nV * h * w)
    step = TransportRenderStep(grid, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], boundary=boundary,
                               loss=loss, background=bkg, loss_scale=scale)
    lo, gd, gv = step.run(_t(d), _t(v), _t(tgt))
    torch.cuda.synchronize()
    r = orc.step_fwd_bwd(d, v, tgt.astype(np.float64), sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"],
                         boundary_render=boundary, loss_kind=loss, background=background, loss_scale=scale)
    assert abs(lo.item() - r["loss"]) <= 1e-5 * abs(r["loss"])
    assert norm_err(gd.cpu().numpy(), r["gd"]) < 1e-3
    assert norm_err(gv.cpu().numpy(), r["gv"]) < 1e-3


def _stag(shape, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    D, H, W = shape
    return [rng.uniform(-1, 1, size=s) for s in ((1, D, H, W + 1, 1), (1, D, H + 1, W, 1), (1, D + 1, H, W, 1))]


@pytest.mark.parametrize("shape", [(5, 7, 6), (33, 20, 70)])
def test_divergence_loss(cuda, shape):
    from gft.grids import VelocityGrid
    from gft.optim import divergence_loss, divergence_loss_into
    comps = _stag(shape, 3)
    ws = (0.5, 2.0, 1.5)
    loss_ref, _, g_ref = orc.divergence_loss_and_grad(*comps, ws, scale=0.3)
    tc = [_t(c) for c in comps]
    grads = [torch.empty_like(c) for c in tc]
    lo = divergence_loss_into(tc, ws, 0.3, grads=grads)
    assert abs(lo.item() - loss_ref) <= 1e-5 * loss_ref
    for g, gr in zip(grads, g_ref):
        assert norm_err(g.cpu().numpy(), gr) < 1e-5
    # accumulate adds onto what is there
    base = [torch.full_like(c, 0.25) for c in tc]
    divergence_loss_into(tc, ws, 0.3, grads=base, accumulate=True)
    for g, gr in zip(base, g_ref):
        assert norm_err(g.cpu().numpy() - 0.25, gr) < 1e-5
    # autograd form on a VelocityGrid against the torch composition of VelocityGrid.divergence
    vg = VelocityGrid(list(shape), x=tc[0].clone().requires_grad_(True), y=tc[1].clone().requires_grad_(True),
                      z=tc[2].clone().requires_grad_(True))
    (2.0 * divergence_loss(vg, ws, 0.3)).backward()
    xs = [c.clone().requires_grad_(True) for c in tc]
    vt = VelocityGrid(list(shape), x=xs[0], y=xs[1], z=xs[2])
    (2.0 * 0.3 * (vt.divergence(ws) ** 2).sum()).backward()
    for a, b in zip((vg.x, vg.y, vg.z), xs):
        assert norm_err(a.grad.cpu().numpy(), b.grad.cpu().numpy()) < 1e-5


def test_divergence_loss_without_gradients_and_errors(cuda):
    from gft import _lib
    from gft.optim import divergence_loss_into
    comps = [_t(c) for c in _stag((4, 5, 6), 9)]
    loss_ref = orc.divergence_loss_and_grad(*[c.cpu().numpy() for c in comps], (1, 1, 1), 1.0)[0]
    assert abs(divergence_loss_into(comps).item() - loss_ref) <= 1e-5 * loss_ref
    with pytest.raises(_lib.InvalidArgumentError):
        divergence_loss_into((comps[0], comps[2], comps[1]))


def test_adam_multi_tensor(cuda):
    """Nine tensors (two launches), ragged sizes and a misaligned view (the
    scalar path), five steps against the oracle's float64 update."""
    from gft.optim import AdamOptimizer
    rng = np.random.Generator(np.random.PCG64(12))
    sizes = [1, 3, 4, 5, 1023, 4096, 77, 10, 6]
    params = [rng.uniform(-1, 1, size=n) for n in sizes]
    big = torch.zeros(sizes[-1] + 1, device=cuda)
    tp = [_t(p) for p in params[:-1]]
    big[1:] = _t(params[-1])
    tp.append(big[1:])  # contiguous, 4-byte but not 16-byte aligned
    assert tp[-1].data_ptr() % 16 != 0
    opt = AdamOptimizer(learning_rate=0.01, beta1=0.8, beta2=0.99, epsilon=1e-4)
    ref = [(p, np.zeros_like(p), np.zeros_like(p)) for p in params]
    for t in range(1, 6):
        grads = [rng.uniform(-1, 1, size=n) for n in sizes]
        opt.apply_gradients([(_t(g), p) for g, p in zip(grads, tp)])
        ref = [orc.adam_tf(p, g, m, v, 0.01, 0.8, 0.99, 1e-4, t) for (p, m, v), g in zip(ref, grads)]
    torch.cuda.synchronize()
    for p, (pr, mr, vr) in zip(tp, ref):
        assert np.allclose(p.cpu().numpy(), pr, rtol=1e-5, atol=1e-6)
        m = opt.get_slot(p, "m")
        assert np.allclose(m.cpu().numpy(), mr, rtol=1e-5, atol=1e-7)
    assert big[0].item() == 0.0  # the element before the view is untouched


def test_optimisation_step_in_one_graph(cuda):
    """The fused step and the Adam update captured in one graph with the
    device step counter: replays continue the eager sequence (same
    parameters as an eager run of the same number of updates with the host
    counter)."""
    from gft.optim import AdamOptimizer
    from gft.pipeline import TransportRenderStep
    grid = (16, 26, 18)
    sc = scene(grid=grid, image=(24, 14), views=3, seed=8)
    d, v = _fields(grid, seed=12)
    tgt = torch.from_numpy(orc.render_density_fwd(syn.gaussian_blobs(grid, 5, seed=3), sc["M"], sc["V"], sc["P"],
                                                  sc["F"], sc["out_dhw"], sc["step"])[..., :3].astype(np.float32)).cuda()

    def make():
        st = TransportRenderStep(grid, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"], loss="AE")
        return st, _t(d), _t(v)

    def iterate(st, d0, vel, opt):
        st.run(d0, vel, tgt)
        opt.apply_gradients([(st.grad_density, d0), (st.grad_velocity, vel)])

    st, d0, vel = make()
    opt = AdamOptimizer(0.01, device_step=True)
    iterate(st, d0, vel, opt)  # eager first update (allocates the moments and the counter)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        iterate(st, d0, vel, opt)
    for _ in range(3):
        g.replay()
    torch.cuda.synchronize()
    assert opt.device_iterations == 4

    rs, rd0, rvel = make()
    ropt = AdamOptimizer(0.01)
    for _ in range(4):
        iterate(rs, rd0, rvel, ropt)
    torch.cuda.synchronize()
    assert norm_err(d0.cpu().numpy(), rd0.cpu().numpy()) < 1e-5
    assert norm_err(vel.cpu().numpy(), rvel.cpu().numpy()) < 1e-5
    assert not torch.equal(rd0, _t(d))  # the parameters moved
