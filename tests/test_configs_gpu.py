"""The remaining workloads at their full sizes, and the c2 gradients against
FP64 central finite differences.

c1 (64x112x64, 5 views of 228x128, 91 samples per ray): forward only, one
semi-Lagrangian advection then the render, both through the fused kernels
and through the per-op path (SampleGridTransform TRANSFORM_LINDEPTH into
each frustum, ReduceGridBlend along depth), against the FP64 oracle at 1e-5.

c2: the fused step's density and velocity gradients against FP64 central
finite differences (h = 1e-3) of the FP64 oracle's loss at 64 random voxels
each (drawn among voxels whose gradient is at least 1 % of the maximum),
within 1e-3 of the gradient's maximum. The difference quotients are exact
but local: a density or velocity perturbation at voxel j changes the
advected density only within a few cells of j (the MacCormack traces move at
most 1.5 cells), so the advection is recomputed on a crop around j and only
the pixels whose rays reach the changed box are rendered, through cameras
cropped to those pixels (the same rays). Both locality claims are checked
against full-size computations. Voxels whose MC_SMOOTH decisions or (for the
velocity) trace cells change within +-h are kinks of the loss and are
skipped deterministically.

c2 and c5 MC_SMOOTH branch agreement: the GPU's saved clamp mask against the
FP32 and FP64 oracles' branch choices, and the gradients on voxels whose
support saw no disagreeing branch and no trace near a cell face.

c3 (100x178x100, 20 frames, 5 views + the unseen one of 480x270, 142
samples per ray): the whole chain with checkpoints every 4 frames equal to
checkpoints every frame and to no checkpoints, less peak memory than no
checkpoints, and the first two frames against the FP64 oracle. Also the
semi-Lagrangian adjoint identity at c3's speed (2 cells per step).

c4: the seeded draw of 32 reconstructions captured as one graph (a branch per
reconstruction); every replay equals the eager run and every reconstruction
up to 64x112x64 matches the FP64 oracle.
"""
import copy

import numpy as np
import pytest
import torch

from gft import synthetic as syn
from helpers import norm_err
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def _workload(name, extra_yaws=()):
    from gft.camera import camera_params_batch, model_matrices
    w = syn.WORKLOADS[name]
    grid, image = w["grid"], w["image"]
    cams, S = syn.arc_cameras(grid, image, w["views"], extra_yaws=extra_yaws)
    M = model_matrices([syn.grid_transform(grid)])
    V, P, F = camera_params_batch(cams)
    return dict(grid=grid, cams=cams, M=M, V=V, P=P, F=F, out_dhw=(S,) + tuple(image),
                step=float(cams[0].depth_step), blobs=w["blobs"], max_speed=w["max_speed"])


def _target(w, d_target):
    from gft.pipeline import _Cameras, render_density_forward
    cams = _Cameras(w["M"], w["V"], w["P"], w["F"], w["out_dhw"])
    return render_density_forward(_t(d_target), cams, w["step"])[..., :3].contiguous()


# ------------------------------------------------------------------- c1
def test_c1_forward_fused_and_per_op_match_fp64_oracle(cuda):
    from gft.ops import blending_ops, sampling_ops
    from gft.pipeline import _Cameras, advect, render_density_forward
    w = _workload("c1")
    print("c1 forward", w["grid"], w["out_dhw"], flush=True)
    d0 = syn.gaussian_blobs(w["grid"], w["blobs"], seed=0)
    vel = syn.curl_noise_velocity(w["grid"], seed=1, max_speed=w["max_speed"])
    d1 = advect(_t(d0), _t(vel), 1.0, 1, "NONE")
    cams = _Cameras(w["M"], w["V"], w["P"], w["F"], w["out_dhw"])
    img = render_density_forward(d1, cams, w["step"])
    # per-op: the ambient light grid [d, d, d, d] * step resampled into each frustum and blended
    grid4 = torch.cat([d1, d1, d1, d1], dim=-1) * w["step"]
    fr = sampling_ops.sample_grid_transform(grid4, w["M"], w["V"], w["P"], w["F"], w["out_dhw"],
                                            coordinate_mode="TRANSFORM_LINDEPTH")
    img_op = blending_ops.reduce_grid_blend(fr[0], "BEER_LAMBERT")
    torch.cuda.synchronize()
    r1 = orc.advect_fwd(d0, vel, 1, "NONE", prec=1)
    assert norm_err(d1.cpu().numpy(), r1) < 1e-5
    ref = orc.render_density_fwd(r1, w["M"], w["V"], w["P"], w["F"], w["out_dhw"], w["step"])
    assert norm_err(img.cpu().numpy(), ref) < 1e-5
    g4 = np.concatenate([r1] * 4, axis=-1) * w["step"]
    fr_ref = orc.sample_transform_fwd(g4, w["M"], w["V"], w["P"], w["F"], w["out_dhw"], "LINEAR", "BORDER",
                                      coordinate_mode="TRANSFORM_LINDEPTH")
    img_ref = orc.reduce_blend_fwd(fr_ref[0], "BEER_LAMBERT")
    assert norm_err(img_op.cpu().numpy(), img_ref) < 1e-5


# ------------------------------------------------------- c2 finite differences
def _box(center, radius, grid):
    D, H, W = grid
    lo = [max(0, c - radius) for c in center]
    hi = [min(n, c + radius + 1) for c, n in zip(center, (D, H, W))]
    return tuple(slice(a, b) for a, b in zip(lo, hi))


def _crop_camera(cam, x0, y0, w, h, W, H):
    c = copy.copy(cam)
    t, r, b, l = cam.trbl
    sx, sy = (r - l) / W, (t - b) / H
    c.trbl = [b + (y0 + h) * sy, l + (x0 + w) * sx, b + y0 * sy, l + x0 * sx]
    return c


class _LocalLoss(object):
    """FP64 loss differences of the c2 step for perturbations at one voxel,
    evaluated on a crop of the advection and the pixels that see it."""

    def __init__(self, w, d0, vel, tgt):
        from gft.camera import camera_params_batch
        self.w, self._cpb = w, camera_params_batch
        self.d0, self.vel = d0.astype(np.float64), vel.astype(np.float64)
        self.tgt = tgt.astype(np.float64)
        self.d1, self.parts = orc.advect_fwd(self.d0, self.vel, 2, "MC_SMOOTH", prec=1, return_parts=True)
        _, _, self.H, self.W, _ = self.tgt.shape
        self.scale = 1.0 / self.tgt.size  # mean squared error over rgb
        Mr = w["M"][0].T.astype(np.float64)
        self.PVM = [w["P"][v].T.astype(np.float64) @ w["V"][v].T.astype(np.float64) @ Mr for v in range(len(w["cams"]))]

    def advect_box(self, d0, vel, center, radius):
        """advected density and MC_SMOOTH mask on the box of `radius` around
        center, from a crop 6 cells wider (every lookup of the box stays inside)"""
        grid = self.d0.shape[1:4]
        crop = _box(center, radius + 6, grid)
        b0 = (slice(0, 1),) + crop
        out, parts = orc.advect_fwd(d0[b0], vel[b0], 2, "MC_SMOOTH", prec=1, return_parts=True)
        box = _box(center, radius, grid)
        inner = tuple(slice(b.start - c.start, b.stop - c.start) for b, c in zip(box, crop))
        return box, out[(0,) + inner], parts["mask"][(0,) + inner]

    def pixel_rects(self, box):
        """per view the pixel rectangle whose rays can reach the box's support"""
        z, y, x = box
        lo = np.array([x.start - 0.5, y.start - 0.5, z.start - 0.5])  # object space = index + 0.5, support +-1
        hi = np.array([x.stop + 0.5, y.stop + 0.5, z.stop + 0.5])
        corners = np.array([[a, b, c, 1.0] for a in (lo[0], hi[0]) for b in (lo[1], hi[1]) for c in (lo[2], hi[2])])
        rects = []
        for A in self.PVM:
            clip = corners @ A.T
            ndc = clip[:, :2] / clip[:, 3:4]
            px = (ndc[:, 0] + 1) * 0.5 * self.W - 0.5
            py = (ndc[:, 1] + 1) * 0.5 * self.H - 0.5
            x0, x1 = max(0, int(np.floor(px.min())) - 1), min(self.W - 1, int(np.ceil(px.max())) + 1)
            y0, y1 = max(0, int(np.floor(py.min())) - 1), min(self.H - 1, int(np.ceil(py.max())) + 1)
            rects.append((x0, y0, x1 - x0 + 1, y1 - y0 + 1))
        return rects

    def crop_loss(self, box, values):
        """loss_scale * sum (rgb - target)^2 over every view's pixel rect of
        the box (padded to a common size inside the image), with the advected
        density replaced by `values` on the box; the render is the ambient one,
        a raymarch of [d, d, d, d] * step"""
        w = self.w
        if not hasattr(self, "g4"):
            self.g4 = np.concatenate([self.d1] * 4, axis=-1) * w["step"]
        rects = self.pixel_rects(box)
        cw = max(r[2] for r in rects)
        ch = max(r[3] for r in rects)
        cams, offs = [], []
        for cam, (x0, y0, _, _) in zip(w["cams"], rects):
            x0, y0 = min(x0, self.W - cw), min(y0, self.H - ch)
            cams.append(_crop_camera(cam, x0, y0, cw, ch, self.W, self.H))
            offs.append((x0, y0))
        V, P, F = self._cpb(cams)
        keep = self.g4[(0,) + box].copy()
        self.g4[(0,) + box] = values * w["step"]
        try:
            img = orc.raymarch_fwd(self.g4, w["M"], V, P, F, (w["out_dhw"][0], ch, cw))
        finally:
            self.g4[(0,) + box] = keep
        total = 0.0
        for v, (x0, y0) in enumerate(offs):
            e = img[0, v, ..., :3] - self.tgt[0, v, y0:y0 + ch, x0:x0 + cw]
            total += float((e * e).sum())
        return total * self.scale


def test_c2_gradients_match_fp64_finite_differences(cuda):
    from gft.pipeline import TransportRenderStep
    w = _workload("c2")
    d0 = syn.gaussian_blobs(w["grid"], w["blobs"], seed=0)
    vel = syn.curl_noise_velocity(w["grid"], seed=1, max_speed=w["max_speed"])
    tgt = _target(w, syn.gaussian_blobs(w["grid"], w["blobs"], seed=2))
    print("c2 FD64: GPU step", flush=True)
    step = TransportRenderStep(w["grid"], w["M"], w["V"], w["P"], w["F"], w["out_dhw"], w["step"])
    _, gd, gv = step.run(_t(d0), _t(vel), tgt)
    torch.cuda.synchronize()
    gd, gv = gd.cpu().numpy().astype(np.float64), gv.cpu().numpy().astype(np.float64)
    ll = _LocalLoss(w, d0, vel, tgt.cpu().numpy())
    h = 1e-3
    rng = np.random.Generator(np.random.PCG64(64))

    # locality checks on one voxel: the crop advection equals the full one on the
    # box, and a perturbation changes no pixel outside the rects
    j0 = np.unravel_index(int(np.argmax(np.abs(gd[0, ..., 0]))), gd.shape[1:4])
    box, a, _ = ll.advect_box(ll.d0, ll.vel, j0, 6)
    assert np.abs(a - ll.d1[(0,) + box]).max() <= 1e-14 * np.abs(ll.d1).max()
    dp = ll.d0.copy()
    dp[(0,) + tuple(j0) + (0,)] += 0.05
    full = orc.advect_fwd(dp, ll.vel, 2, "MC_SMOOTH", prec=1)
    changed = np.argwhere(full[0, ..., 0] != ll.d1[0, ..., 0])
    assert np.all(changed >= np.array([s.start for s in box])) and np.all(changed < np.array([s.stop for s in box]))
    img_a = orc.render_density_fwd(ll.d1, w["M"], w["V"], w["P"], w["F"], w["out_dhw"], w["step"])
    img_b = orc.render_density_fwd(full, w["M"], w["V"], w["P"], w["F"], w["out_dhw"], w["step"])
    rects = ll.pixel_rects(box)
    diff = np.abs(img_b - img_a).max(-1)[0]
    for v, (x0, y0, cw, ch) in enumerate(rects):
        ys, xs = np.nonzero(diff[v])
        assert len(ys) > 0
        assert xs.min() >= x0 and xs.max() < x0 + cw and ys.min() >= y0 and ys.max() < y0 + ch
    tg64 = ll.tgt

    def full_loss(img):
        e = img[..., :3] - tg64
        return float((e * e).sum()) * ll.scale
    want = full_loss(img_b) - full_loss(img_a)
    got = ll.crop_loss(box, full[(0,) + box]) - ll.crop_loss(box, ll.d1[(0,) + box])
    # the cropped cameras' float32 projections move the rays by rounding: 1e-5 relative
    assert abs(got - want) <= 1e-5 * abs(want)

    def fd_density(j):
        vals, masks = [], []
        for s in (1.0, -1.0):
            dd = ll.d0.copy()
            dd[(0,) + tuple(j) + (0,)] += s * h
            box, a, m = ll.advect_box(dd, ll.vel, j, 6)
            masks.append(m)
            vals.append(ll.crop_loss(box, a))
        if not (np.array_equal(masks[0], masks[1]) and np.array_equal(masks[0], ll.parts["mask"][(0,) + box])):
            return None  # an MC_SMOOTH decision switches within +-h: a kink
        return (vals[0] - vals[1]) / (2 * h)

    def fd_velocity(j, c):
        vals, masks = [], []
        for s in (1.0, -1.0):
            vv = ll.vel.copy()
            vv[(0,) + tuple(j) + (c,)] += s * h
            p = np.array(j[::-1], dtype=np.float64)  # x, y, z
            for pos in (p - vv[(0,) + tuple(j)], p + vv[(0,) + tuple(j)]):
                masks.append(tuple(np.floor(pos)) + tuple(np.ceil(pos)))
            box, a, m = ll.advect_box(ll.d0, vv, j, 4)
            masks.append(m.tobytes())
            vals.append(ll.crop_loss(box, a))
        if masks[:3] != masks[3:]:
            return None  # a trace corner or an MC_SMOOTH decision changes within +-h
        return (vals[0] - vals[1]) / (2 * h)

    for name, g, ncomp, fd in (("density", gd, 1, fd_density), ("velocity", gv, 3, fd_velocity)):
        gmax = np.abs(g).max()
        cand = np.argwhere(np.abs(g[0]) >= 0.01 * gmax)
        order = rng.permutation(len(cand))
        checked, worst = 0, 0.0
        for k in order[:200]:
            idx = tuple(int(i) for i in cand[k])
            j, c = idx[:3], idx[3]
            q = fd(j) if ncomp == 1 else fd(j, c)
            if q is None:
                continue
            err = abs(q - g[(0,) + idx]) / gmax
            worst = max(worst, err)
            assert err <= 1e-3, (name, idx, q, g[(0,) + idx])
            checked += 1
            if checked == 64:
                break
        print("c2 FD64 %s: %d voxels, worst %.2e of max" % (name, checked, worst), flush=True)
        assert checked == 64, (name, checked)


# ---------------------------------------------------- MC_SMOOTH branch agreement
def _gpu_mask(step, grid):
    V = grid[0] * grid[1] * grid[2]
    off = (V * 4 + 255) // 256 * 256
    return step.saved[off:off + V].cpu().numpy().reshape(grid).astype(bool)


def _near_face(vel, eps):
    """voxels whose trace x -/+ v sits within eps of a cell face on some axis"""
    D, H, W = vel.shape[1:4]
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    p = np.stack([x, y, z], -1).astype(np.float64)
    v = vel[0].astype(np.float64)
    out = np.zeros((D, H, W), bool)
    for q in (p - v, p + v):
        fr = q - np.round(q)
        out |= (np.abs(fr) < eps).any(-1)
    return out


@pytest.mark.parametrize("name,storage", [("c2", "float32"), ("c5", "bfloat16")])
def test_mc_smooth_branches_agree_with_the_oracle(cuda, name, storage):
    """Every MC_SMOOTH decision of the GPU equals the FP32 and FP64 oracles'
    except at near-ties of the corrected value with a corner extremum (within
    1e-5 of the density's maximum; both FP32 arithmetics round to about 1e-7).
    The gradients are then held to 1e-3 of their maximum on every voxel outside
    the exact dependency set of a disagreeing decision (the voxels the oracle's
    unclamped MacCormack adjoint reaches from those cells, with random weights
    so nothing cancels) and, for the velocity, away from traces within 1e-4 of
    a cell face; the excluded voxels' error stays below 0.25."""
    from gft.pipeline import TransportRenderStep
    w = _workload(name)
    print("MC_SMOOTH branches", name, storage, flush=True)
    d0 = syn.gaussian_blobs(w["grid"], w["blobs"], seed=0)
    vel = syn.curl_noise_velocity(w["grid"], seed=1, max_speed=w["max_speed"])
    tgt = _target(w, syn.gaussian_blobs(w["grid"], w["blobs"], seed=2))
    td, tv = _t(d0), _t(vel)
    if storage == "bfloat16":
        td, tv = td.to(torch.bfloat16), tv.to(torch.bfloat16)
        d0 = td.float().cpu().numpy()
        vel = tv.float().cpu().numpy()
    step = TransportRenderStep(w["grid"], w["M"], w["V"], w["P"], w["F"], w["out_dhw"], w["step"], storage=storage)
    _, gd, gv = step.run(td, tv, tgt)
    torch.cuda.synchronize()
    m_gpu = _gpu_mask(step, w["grid"])
    _, p32 = orc.advect_fwd(d0, vel, 2, "MC_SMOOTH", prec=0, return_parts=True)
    _, p64 = orc.advect_fwd(d0, vel, 2, "MC_SMOOTH", prec=1, return_parts=True)
    m32, m64 = p32["mask"][0, ..., 0], p64["mask"][0, ..., 0]
    dis32, dis64 = m_gpu != m32, m_gpu != m64
    corr, mn, mx = p64["corr"][0, ..., 0], p64["mn"][0, ..., 0], p64["mx"][0, ..., 0]
    tie = np.minimum(np.abs(corr - mn), np.abs(corr - mx)) <= 1e-5 * np.abs(d0).max()
    print("%s: decisions differing from the FP32 oracle %d, from FP64 %d (of %d); near-ties %d" % (
        name, dis32.sum(), dis64.sum(), m_gpu.size, tie.sum()), flush=True)
    assert not (dis32 & ~tie).any() and not (dis64 & ~tie).any()
    r = orc.step_fwd_bwd(d0, vel, tgt.cpu().numpy().astype(np.float64), w["M"], w["V"], w["P"], w["F"],
                         w["out_dhw"], w["step"], render_storage=storage)
    rng = np.random.Generator(np.random.PCG64(3))
    ind = (dis64 * rng.uniform(0.5, 1.5, size=dis64.shape))[None, ..., None]
    dep_d, dep_v = orc.advect_bwd(d0, vel, ind, 2, "NONE", prec=1)
    dep_d, dep_v = dep_d[0, ..., 0] != 0, (dep_v[0] != 0).any(-1)
    gd, gv = gd.cpu().numpy()[0, ..., 0], gv.cpu().numpy()[0]
    gdr, gvr = r["gd"][0, ..., 0], r["gv"][0]
    ok_d = ~dep_d
    ok_v = ~(dep_v | _near_face(vel, 1e-4))
    ed = np.abs(gd - gdr) / np.abs(gdr).max()
    ev = np.abs(gv - gvr).max(-1) / np.abs(gvr).max()
    print("%s: density grad max err %.2e on %.6f of voxels (%.2e overall); velocity %.2e on %.6f (%.2e)" % (
        name, ed[ok_d].max(), ok_d.mean(), ed.max(), ev[ok_v].max(), ok_v.mean(), ev.max()), flush=True)
    assert ed[ok_d].max() < 1e-3 and ok_d.mean() > 0.99
    assert ev[ok_v].max() < 1e-3 and ok_v.mean() > 0.99
    assert ed.max() < 0.25 and ev.max() < 0.25


# ------------------------------------------------------------------- c3
def _chain_workload():
    w = _workload("c3", extra_yaws=(syn.WORKLOADS["c3"]["unseen_yaw"],))
    T = 20
    d0 = syn.gaussian_blobs(w["grid"], w["blobs"], seed=0)
    vels = [syn.curl_noise_velocity(w["grid"], seed=100 + f, max_speed=w["max_speed"]) for f in range(T - 1)]
    tg = [_target(w, syn.gaussian_blobs(w["grid"], w["blobs"], seed=200 + f)) for f in range(T)]
    return w, d0, vels, tg


@pytest.fixture(scope="module")
def c3():
    return _chain_workload()


def test_c3_chain_checkpointing_at_full_size(cuda, c3):
    from gft.chain import TransportChain
    w, d0, vels, tg = c3
    assert len(w["cams"]) == 6 and w["out_dhw"] == (142, 480, 270)
    td, tv = _t(d0), [_t(v) for v in vels]
    res, peaks = {}, {}
    for k in (20, 4, 1):
        print("c3 chain k=%d" % k, flush=True)
        ch = TransportChain(w["grid"], w["M"], w["V"], w["P"], w["F"], w["out_dhw"], w["step"], checkpoint_every=k)
        torch.cuda.synchronize()
        base = torch.cuda.memory_allocated()
        torch.cuda.reset_peak_memory_stats()
        loss, gd0, gv = ch.run(td, tv, tg)
        torch.cuda.synchronize()
        peaks[k] = torch.cuda.max_memory_allocated() - base
        res[k] = (loss.item(), ch.frame_loss.cpu().numpy(), gd0.cpu().numpy(), [g.cpu().numpy() for g in gv])
        del ch, loss, gd0, gv
    for k in (4, 1):
        assert np.array_equal(res[k][1], res[20][1])  # the forward recomputes bit for bit
        assert norm_err(res[k][2], res[20][2]) < 1e-5  # adjoints accumulate with atomics
        assert all(norm_err(a, b) < 1e-5 for a, b in zip(res[k][3], res[20][3]))
    assert peaks[4] < peaks[20], peaks


def test_c3_first_two_frames_match_fp64_oracle(cuda, c3):
    from gft.chain import TransportChain
    w, d0, vels, tg = c3
    print("c3 chain T=2 vs oracle", flush=True)
    ch = TransportChain(w["grid"], w["M"], w["V"], w["P"], w["F"], w["out_dhw"], w["step"], checkpoint_every=4)
    loss, gd0, gv = ch.run(_t(d0), [_t(vels[0])], tg[:2])
    torch.cuda.synchronize()
    r = orc.chain_fwd_bwd(d0, vels[:1], [t.cpu().numpy().astype(np.float64) for t in tg[:2]], w["M"], w["V"],
                          w["P"], w["F"], w["out_dhw"], w["step"])
    assert abs(loss.item() - r["loss"]) <= 1e-5 * r["loss"]
    g, gr = gd0.cpu().numpy(), r["gd0"]
    assert np.mean(np.abs(g - gr) > 1e-3 * np.abs(gr).max()) <= 1e-4  # MC_SMOOTH near-ties, as at c2
    gvv = gv[0].cpu().numpy()
    assert np.mean(np.abs(gvv - r["gv"][0]) > 1e-3 * np.abs(r["gv"][0]).max()) <= 1e-4  # face crossings


def test_c3_semi_lagrangian_adjoint_identity(cuda, c3):
    from gft.pipeline import advect
    w, _, vels, _ = c3
    rng = np.random.Generator(np.random.PCG64(13))
    d = rng.uniform(-1, 1, size=(1,) + tuple(w["grid"]) + (1,)).astype(np.float32)
    g = rng.uniform(-1, 1, size=d.shape).astype(np.float32)
    td = _t(d).requires_grad_(True)
    out = advect(td, _t(vels[7]), 1.0, 1, "NONE")
    out.backward(_t(g))
    lhs = float(np.dot(out.detach().cpu().numpy().astype(np.float64).ravel(), g.astype(np.float64).ravel()))
    rhs = float(np.dot(d.astype(np.float64).ravel(), td.grad.cpu().numpy().astype(np.float64).ravel()))
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + 1.0)


# ------------------------------------------------------------------- c4
def test_c4_seeded_draw_in_one_graph(cuda):
    from gft.camera import camera_params_batch, model_matrices
    from gft.pipeline import TransportRenderStep, _Cameras, render_density_forward
    recs = syn.batched_uneven(32, seed=7)
    steps, inputs, meta = [], [], []
    for r in recs:
        grid = r["grid"]
        cams, S = syn.arc_cameras(grid, r["image"], r["views"])
        M = model_matrices([syn.grid_transform(grid)])
        V, P, F = camera_params_batch(cams)
        out_dhw = (S,) + tuple(r["image"])
        step = float(cams[0].depth_step)
        d0 = syn.gaussian_blobs(grid, r["blobs"], seed=r["seed"])
        vel = syn.curl_noise_velocity(grid, seed=r["seed"] + 1, max_speed=r["max_speed"])
        tgt = render_density_forward(_t(syn.gaussian_blobs(grid, r["blobs"], seed=r["seed"] + 2)),
                                     _Cameras(M, V, P, F, out_dhw), step)[..., :3].contiguous()
        steps.append(TransportRenderStep(grid, M, V, P, F, out_dhw, step))
        inputs.append((_t(d0), _t(vel), tgt))
        meta.append((grid, M, V, P, F, out_dhw, step, d0, vel))
    print("c4: %d reconstructions, grids %s" % (len(steps), sorted({m[0] for m in meta})), flush=True)
    for st, ins in zip(steps, inputs):
        st.run(*ins)
    torch.cuda.synchronize()
    eager = [(st.loss.clone(), st.images.clone(), st.grad_density.clone(), st.grad_velocity.clone()) for st in steps]
    branches = [torch.cuda.Stream() for _ in steps]
    for s_, st, ins in zip(branches, steps, inputs):
        s_.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s_):
            st.run(*ins)
        torch.cuda.current_stream().wait_stream(s_)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        cap = torch.cuda.current_stream()
        for s_, st, ins in zip(branches, steps, inputs):
            s_.wait_stream(cap)
            with torch.cuda.stream(s_):
                st.run(*ins)
        for s_ in branches:
            cap.wait_stream(s_)
    for rep in range(2):
        for st in steps:
            st.images.fill_(-1.0)
            st.grad_density.fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for st, (lo, img, gd, gv) in zip(steps, eager):
            assert torch.equal(st.images, img) and torch.equal(st.loss, lo)
            assert norm_err(st.grad_density.cpu().numpy(), gd.cpu().numpy()) < 1e-5
            assert norm_err(st.grad_velocity.cpu().numpy(), gv.cpu().numpy()) < 1e-5
    checked = 0
    for st, ins, (grid, M, V, P, F, out_dhw, step, d0, vel) in zip(steps, inputs, meta):
        if grid[0] * grid[1] * grid[2] > 64 * 112 * 64:
            continue
        r = orc.step_fwd_bwd(d0, vel, ins[2].cpu().numpy().astype(np.float64), M, V, P, F, out_dhw, step)
        assert abs(st.loss.item() - r["loss"]) <= 1e-5 * r["loss"]
        assert norm_err(st.images.cpu().numpy(), r["img"]) < 1e-5
        gd = st.grad_density.cpu().numpy()
        assert np.mean(np.abs(gd - r["gd"]) > 1e-3 * np.abs(r["gd"]).max()) <= 1e-4
        gv = st.grad_velocity.cpu().numpy()
        assert np.mean(np.abs(gv - r["gv"]) > 1e-3 * np.abs(r["gv"]).max()) <= 1e-4
        checked += 1
    print("c4: %d reconstructions checked against the oracle" % checked, flush=True)
    assert checked >= 10
