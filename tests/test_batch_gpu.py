"""The batched density-reconstruction step (gft_batch_render_*): many
reconstructions of uneven grid sizes and view counts, each rendered with the
non-fused mipmapped resample (TRANSFORM_LINDEPTH, LINEAR mips, per-sample
LoD) and the BEER_LAMBERT blend, L2 loss and exact density gradient: prep
passes (pyramid levels and corner quads), one persistent forward launch and
the gathers (coarse levels, then levels 0 and 1 with the mip adjoint and the
losses).

Against the FP64 oracle composition mip_render_fwd_bwd (the pinned resample
with mips, the pinned blend and the resample's exact adjoint): images within
1e-5, gradients within 1e-4 (normalised max error), losses equal to the FP64
loss of the GPU's own images and within what the image tolerance allows of
the oracle's. Images and losses do not depend on the brick (the gather's
voxel block); gradients agree across bricks to 1e-6 (summation order), and a
graph replay repeats the eager run bit for bit (no atomics in global
memory). The seeded c4 draw of 32 reconstructions: replay equals eager, the
reconstructions up to 64x112x64 vs the oracle, and every reconstruction at
full size vs an independent GPU path (the per-op resample kernel with its
exact adjoint, a float64 torch blend and loss).
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


def _check_loss(loss, img, tgt, ref_img, ref_loss, img_tol=1e-5):
    """The loss of one reconstruction: equal (to float summation error) to the
    mean squared rgb error of the GPU's own images, computed here in FP64;
    and within what the image tolerance allows of the oracle's loss (an
    image error of at most t moves each squared error by at most 2|e| t +
    t^2)."""
    rgb = img[..., :3].astype(np.float64)
    n = rgb.size
    own = float(((rgb - tgt) ** 2).sum()) / n
    assert abs(loss - own) <= 1e-5 * own + 1e-12
    t = img_tol * np.abs(ref_img).max()
    e = np.abs(ref_img[..., :3] - tgt)
    assert abs(loss - ref_loss) <= (2.0 * t * e.sum() + n * t * t) / n * 1.01 + 1e-6 * ref_loss


def _small_batch():
    """Four reconstructions; the third renders a 40x36x30 grid into 9x8-pixel
    images, so its far samples select coarser mip levels (LoD > 1)."""
    specs = [((16, 26, 18), 3, (24, 14), 0), ((12, 20, 14), 1, (18, 11), 1), ((30, 36, 40), 2, (9, 8), 2),
             ((10, 10, 10), 4, (12, 12), 0)]
    recons, dens, tgts, meta = [], [], [], []
    rng = np.random.Generator(np.random.PCG64(5))
    for k, (grid, views, image, brick) in enumerate(specs):
        sc = scene(grid=grid, image=image, views=views, seed=60 + k)
        recons.append(dict(grid=grid, M=sc["M"], V=sc["V"], P=sc["P"], F=sc["F"], out_dhw=sc["out_dhw"],
                           depth_step=sc["step"], light=(1.0, 0.8, 0.6), brick=brick))
        dens.append(sc["data"] * 2.0)
        tgts.append(rng.uniform(0, 0.5, size=(views,) + tuple(image) + (3,)).astype(np.float32))
        meta.append(sc)
    return recons, dens, tgts, meta


@pytest.mark.parametrize("mips", [3, 0])
def test_batched_render_matches_oracle(cuda, mips):
    from gft.batch import BatchedRenderStep
    recons, dens, tgts, meta = _small_batch()
    b = BatchedRenderStep(recons, num_mipmaps=mips)
    b.bind([_t(d) for d in dens], [_t(t) for t in tgts])
    b.run()
    torch.cuda.synchronize()
    for k, (r, d, t, sc) in enumerate(zip(recons, dens, tgts, meta)):
        ref = orc.mip_render_fwd_bwd(d, t, r["M"], r["V"], r["P"], r["F"], r["out_dhw"], r["depth_step"],
                                     r["light"], num_mipmaps=mips)
        assert norm_err(b.images[k].cpu().numpy(), ref["img"]) < 1e-5, k
        _check_loss(b.losses[k].item(), b.images[k].cpu().numpy(), t, ref["img"], ref["loss"])
        assert norm_err(b.grads[k].cpu().numpy(), ref["gd"]) < 1e-4, k
    if mips:
        # the coarse reconstruction really reads its mip levels
        r = recons[2]
        pos = orc.sample_positions(r["M"], r["V"], r["P"], r["F"], r["grid"], r["out_dhw"], "TRANSFORM_LINDEPTH",
                                   lod=True)
        assert pos[..., 3].max() > 1.5


def test_batched_render_bricks_and_graph_replay(cuda):
    from gft.batch import BatchedRenderStep
    recons, dens, tgts, _ = _small_batch()
    b = BatchedRenderStep(recons, num_mipmaps=3)
    b.bind([_t(d) for d in dens], [_t(t) for t in tgts])
    results = {}
    for bricks in ([0, 0, 0, 0], [1, 1, 1, 1], [2, 2, 2, 2]):
        b.set_bricks(bricks)
        b.run()
        torch.cuda.synchronize()
        results[bricks[0]] = ([i.clone() for i in b.images], [g.clone() for g in b.grads],
                              [lo.clone() for lo in b.losses])
    for k in range(len(recons)):
        for br in (1, 2):
            assert torch.equal(results[br][0][k], results[0][0][k])
            assert torch.equal(results[br][2][k], results[0][2][k])
            assert norm_err(results[br][1][k].cpu().numpy(), results[0][1][k].cpu().numpy()) < 1e-6
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        b.run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        b.run()
    for _ in range(3):
        for i, gr in zip(b.images, b.grads):
            i.fill_(-1.0)
            gr.fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k in range(len(recons)):
            assert torch.equal(b.images[k], results[2][0][k]) and torch.equal(b.losses[k], results[2][2][k])
            assert torch.equal(b.grads[k], results[2][1][k])


def test_batched_render_edge_shapes(cuda):
    """Shapes at the edges of the batched passes, against the FP64 oracle with
    3 mip levels requested: a 1x1x1 grid (no level above 0), a grid whose
    chain stops at level 1 (an axis of 2 cells) seen by the largest view count
    (8) in one-pixel images, a single depth sample per ray (the column's LoD
    then steps to the next depth index, as the oracle's forward difference
    does), and an odd grid in wide images."""
    from gft.batch import BatchedRenderStep
    specs = [((1, 1, 1), 2, (5, 7), None, 0), ((5, 3, 2), 8, (1, 1), None, 1), ((7, 9, 6), 3, (4, 6), 1, 2),
             ((9, 17, 11), 1, (13, 40), None, 0)]
    recons, dens, tgts = [], [], []
    rng = np.random.Generator(np.random.PCG64(17))
    for k, (grid, views, image, S, brick) in enumerate(specs):
        sc = scene(grid=grid, image=image, views=views, seed=90 + k, yaw_range=(-170.0, 170.0))
        out = (S if S is not None else sc["out_dhw"][0],) + tuple(image)
        recons.append(dict(grid=grid, M=sc["M"], V=sc["V"], P=sc["P"], F=sc["F"], out_dhw=out,
                           depth_step=sc["step"], light=(1.0, 0.9, 0.7), brick=brick))
        dens.append(sc["data"] * 3.0)
        tgts.append(rng.uniform(0, 0.5, size=(views,) + tuple(image) + (3,)).astype(np.float32))
    print("edge shapes", [r["grid"] for r in recons], [r["out_dhw"] for r in recons], flush=True)
    b = BatchedRenderStep(recons, num_mipmaps=3)
    b.bind([_t(d) for d in dens], [_t(t) for t in tgts])
    b.run()
    torch.cuda.synchronize()
    for k, (r, d, t) in enumerate(zip(recons, dens, tgts)):
        ref = orc.mip_render_fwd_bwd(d, t, r["M"], r["V"], r["P"], r["F"], r["out_dhw"], r["depth_step"],
                                     r["light"], num_mipmaps=3)
        assert norm_err(b.images[k].cpu().numpy(), ref["img"]) < 1e-5, k
        _check_loss(b.losses[k].item(), b.images[k].cpu().numpy(), t, ref["img"], ref["loss"])
        assert norm_err(b.grads[k].cpu().numpy(), ref["gd"]) < 1e-4, k


def test_batched_render_validation(cuda):
    from gft import _lib
    from gft.batch import BatchedRenderStep
    recons, dens, tgts, _ = _small_batch()
    with pytest.raises(_lib.InvalidArgumentError):
        BatchedRenderStep(recons, num_mipmaps=4).bind([_t(d) for d in dens], [_t(t) for t in tgts])
    b = BatchedRenderStep(recons)
    with pytest.raises(_lib.InvalidArgumentError):
        b.bind([_t(d) for d in dens][:-1] + [_t(dens[0])], [_t(t) for t in tgts])
    bad = [dict(r) for r in recons]
    bad[1]["brick"] = 5
    with pytest.raises(_lib.InvalidArgumentError):
        BatchedRenderStep(bad).bind([_t(d) for d in dens], [_t(t) for t in tgts])


def c4_batch(count=32, seed=7):
    """The seeded c4 draw as batched-step records (densities and targets:
    Gaussian blobs, targets rendered from a second blob density with the same
    step)."""
    from gft.camera import camera_params_batch, model_matrices
    recons, dens, tdens = [], [], []
    for r in syn.batched_uneven(count, seed=seed):
        grid = r["grid"]
        cams, S = syn.arc_cameras(grid, r["image"], r["views"])
        M = model_matrices([syn.grid_transform(grid)])
        V, P, F = camera_params_batch(cams)
        recons.append(dict(grid=grid, M=M, V=V, P=P, F=F, out_dhw=(S,) + tuple(r["image"]),
                           depth_step=float(cams[0].depth_step)))
        dens.append(syn.gaussian_blobs(grid, r["blobs"], seed=r["seed"]))
        tdens.append(syn.gaussian_blobs(grid, r["blobs"], seed=r["seed"] + 2))
    return recons, dens, tdens


def test_c4_batched_draw(cuda):
    from gft.batch import BatchedRenderStep
    recons, dens, tdens = c4_batch()
    print("c4 batched: %d reconstructions" % len(recons), flush=True)
    # targets: the batch's own render of the second densities
    tb = BatchedRenderStep(recons, num_mipmaps=3)
    zeros = [torch.zeros((r["V"].shape[0],) + tuple(r["out_dhw"][1:]) + (3,), device="cuda") for r in recons]
    tb.bind([_t(d) for d in tdens], zeros)
    tb.run()
    targets = [i[..., :3].contiguous() for i in tb.images]
    b = BatchedRenderStep(recons, num_mipmaps=3)
    b.bind([_t(d) for d in dens], targets)
    b.run()
    torch.cuda.synchronize()
    eager = ([i.clone() for i in b.images], [g.clone() for g in b.grads], [lo.clone() for lo in b.losses])
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        b.run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        b.run()
    for _ in range(2):
        g.replay()
        torch.cuda.synchronize()
        for k in range(len(recons)):
            assert torch.equal(b.images[k], eager[0][k]) and torch.equal(b.losses[k], eager[2][k])
            assert torch.equal(b.grads[k], eager[1][k])
    checked = 0
    for k, (r, d) in enumerate(zip(recons, dens)):
        if np.prod(r["grid"]) > 64 * 112 * 64:
            continue
        ref = orc.mip_render_fwd_bwd(d, targets[k].cpu().numpy(), r["M"], r["V"], r["P"], r["F"], r["out_dhw"],
                                     r["depth_step"], num_mipmaps=3)
        assert norm_err(b.images[k].cpu().numpy(), ref["img"]) < 1e-5, k
        _check_loss(b.losses[k].item(), b.images[k].cpu().numpy(), targets[k].cpu().numpy(), ref["img"],
                    ref["loss"])
        assert norm_err(b.grads[k].cpu().numpy(), ref["gd"]) < 1e-4, k
        checked += 1
    print("c4 batched: %d reconstructions checked against the oracle" % checked, flush=True)
    assert checked >= 10


def test_c4_every_reconstruction_against_the_op_composition(cuda):
    """All 32 reconstructions of the seeded c4 draw at full size, against an
    independent GPU path: the per-op SampleGridTransform kernel (TRANSFORM_LINDEPTH,
    LINEAR mips, 3 levels; its EXACT adjoint is checked against the FP64
    oracle in the resample tests) followed by the BEER_LAMBERT blend and L2
    loss restated in float64 torch with autograd. Images 1e-5, gradients 1e-4
    (normalised max error), losses within what the image tolerance allows."""
    from gft.batch import BatchedRenderStep
    from gft.ops import sampling_ops
    recons, dens, tdens = c4_batch()
    rng = np.random.Generator(np.random.PCG64(11))
    targets = []
    for r in recons:
        nv, h, w = r["V"].shape[0], r["out_dhw"][1], r["out_dhw"][2]
        targets.append(_t(rng.uniform(0.0, 0.3, size=(nv, h, w, 3))))
    b = BatchedRenderStep(recons, num_mipmaps=3)
    b.bind([_t(d) for d in dens], targets)
    b.run()
    torch.cuda.synchronize()
    for k, (r, d) in enumerate(zip(recons, dens)):
        step = float(r["depth_step"])
        x = (_t(d) * step).requires_grad_(True)
        fr = sampling_ops.sample_grid_transform(x, r["M"], r["V"], r["P"], r["F"], r["out_dhw"], interpolation="LINEAR",
                                                boundary="BORDER", mipmapping="LINEAR", num_mipmaps=3,
                                                coordinate_mode="TRANSFORM_LINDEPTH", grad_mode="EXACT")
        sig = fr[0, ..., 0].double()  # [V, S, H, W]
        depth = torch.cumsum(sig, dim=1)
        emis = (sig * torch.exp(-depth)).sum(dim=1)  # [V, H, W]
        rgb = emis[..., None] * torch.ones(3, dtype=torch.float64, device="cuda")  # white light
        tg = targets[k].double()
        loss = ((rgb - tg) ** 2).sum() / rgb.numel()
        loss.backward()
        img = torch.cat([rgb, depth[:, -1, ..., None]], dim=-1).detach().float().cpu().numpy()
        got = b.images[k].cpu().numpy()
        assert norm_err(got, img) < 1e-5, k
        _check_loss(b.losses[k].item(), got, tg.cpu().numpy(), img, float(loss.detach()))
        gref = (x.grad * step).cpu().numpy()
        assert norm_err(b.grads[k].cpu().numpy(), gref) < 1e-4, k
        del fr, sig, depth, emis, rgb, x
    print("c4: %d reconstructions checked against the op composition" % len(recons), flush=True)
