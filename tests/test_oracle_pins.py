"""Pins of the CPU oracle (no GPU).

The reference ships no golden vectors, so each oracle function is checked
against what the mathematics fixes independently of the restatement:
closed forms, exact reproduction cases (grid points, linear fields, identity
cameras), brute-force evaluation on tiny grids, adjoint identities and FP64
central finite differences for the EXACT gradients, and structural
invariants of the reference's normalised (REFERENCE) gradients.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as orc

RNG = np.random.Generator(np.random.PCG64(1234))


# ------------------------------------------------------------ small helpers
def _abs_lut(pos, lod=None):
    """Absolute-coordinate lookup table [1, n, 1, 1, 4] for positions (n, 3)."""
    pos = np.asarray(pos, dtype=np.float64)
    lut = np.zeros((1, pos.shape[0], 1, 1, 4))
    lut[0, :, 0, 0, :3] = pos
    if lod is not None:
        lut[0, :, 0, 0, 3] = lod
    return lut


def _lut_sample(data, pos, lod=None, **kw):
    kw.setdefault("relative_coords", False)
    out = orc.sample_lut_fwd(data, _abs_lut(pos, lod), **kw)
    return out[0, 0, :, 0, 0, :]


def _identity_camera(dhw):
    """Camera for which TRANSFORM maps output index q straight to grid index q
    (P = I, V = (index -> ndc), M = I), as renderer.py:197-198 packs an
    identity projection with frustum [-1,1,-1,1,1,-1]."""
    D, H, W = dhw
    A = np.array([[W / 2.0, 0, 0, W / 2.0], [0, H / 2.0, 0, H / 2.0], [0, 0, D / 2.0, D / 2.0], [0, 0, 0, 1]])
    V = np.linalg.inv(A).T.astype(np.float32)[None]
    P = np.eye(4, dtype=np.float32)[None]
    M = np.eye(4, dtype=np.float32)[None]
    F = np.array([[-1, 1, -1, 1, 1, -1]], dtype=np.float32)
    return M, V, P, F


def _trilinear_bruteforce(data, p, boundary):
    """Textbook trilinear interpolation of data[D,H,W,C] at index position p
    (x, y, z; cell centres at integers): sum over the 8 corners of the
    product weights. BORDER: 0 outside [-0.5, dims-0.5), corners outside read
    0; CLAMP clamps corner indices; WRAP wraps them."""
    D, H, W, C = data.shape
    dims = (W, H, D)
    if boundary == "BORDER" and not all(0.0 <= p[a] + 0.5 < dims[a] for a in range(3)):
        return np.zeros(C)
    base = [int(np.floor(p[a])) for a in range(3)]
    t = [p[a] - base[a] for a in range(3)]
    acc = np.zeros(C)
    for dx, dy, dz in itertools.product((0, 1), repeat=3):
        idx = [base[0] + dx, base[1] + dy, base[2] + dz]
        w = (t[0] if dx else 1 - t[0]) * (t[1] if dy else 1 - t[1]) * (t[2] if dz else 1 - t[2])
        if boundary == "BORDER":
            if not all(0 <= idx[a] < dims[a] for a in range(3)):
                continue
        elif boundary == "CLAMP":
            idx = [min(max(idx[a], 0), dims[a] - 1) for a in range(3)]
        else:
            idx = [idx[a] % dims[a] for a in range(3)]
        acc += w * data[idx[2], idx[1], idx[0]]
    return acc


# ------------------------------------------------------- sampling (A1, A8)
@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP", "WRAP"])
def test_trilinear_matches_bruteforce(boundary):
    data = RNG.uniform(-1, 1, size=(1, 3, 4, 5, 2))
    pos = RNG.uniform(-1.5, 6.0, size=(200, 3))
    got = _lut_sample(data, pos, boundary=boundary)
    for i in range(len(pos)):
        ref = _trilinear_bruteforce(data[0], pos[i], boundary)
        assert np.allclose(got[i], ref, atol=1e-12), (pos[i], got[i], ref)


@pytest.mark.parametrize("filt", ["NEAREST", "LINEAR", "MIN", "MAX"])
@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP", "WRAP"])
def test_grid_points_are_exact(filt, boundary):
    data = RNG.uniform(-1, 1, size=(1, 4, 5, 6, 4))
    idx = np.array([(x, y, z) for z in range(4) for y in range(5) for x in range(6)], dtype=np.float64)
    got = _lut_sample(data, idx, interpolation=filt, boundary=boundary)
    ref = data[0].reshape(-1, 4)
    assert np.array_equal(got, ref)


def test_linear_field_is_reproduced():
    D, H, W = 5, 6, 7
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    f = (0.3 * x - 0.7 * y + 1.1 * z + 2.0)[None, ..., None]
    pos = RNG.uniform(0, 1, size=(300, 3)) * np.array([W - 1, H - 1, D - 1])
    for boundary in ("BORDER", "CLAMP"):
        got = _lut_sample(f, pos, boundary=boundary)[:, 0]
        assert np.allclose(got, 0.3 * pos[:, 0] - 0.7 * pos[:, 1] + 1.1 * pos[:, 2] + 2.0, atol=1e-12)


def test_border_clamp_wrap_outside():
    data = RNG.uniform(0.5, 1, size=(1, 3, 3, 3, 1))
    far = np.array([[-5.0, 1.0, 1.0], [1.0, 9.0, 1.0], [1.0, 1.0, -3.2]])
    assert np.all(_lut_sample(data, far, boundary="BORDER") == 0)
    clamp = _lut_sample(data, far, boundary="CLAMP")[:, 0]
    assert clamp[0] == data[0, 1, 1, 0, 0] and clamp[1] == data[0, 1, 2, 1, 0] and clamp[2] == data[0, 0, 1, 1, 0]
    p = RNG.uniform(0, 3, size=(50, 3))
    for filt in ("LINEAR", "NEAREST"):
        a = _lut_sample(data, p, interpolation=filt, boundary="WRAP")
        # positive shifts only: the nearest index truncates toward zero
        # (int cast of p + 0.5), so negative coordinates are not periodic
        shift = np.array([3.0, -6.0, 9.0]) if filt == "LINEAR" else np.array([3.0, 6.0, 9.0])
        b = _lut_sample(data, p + shift, interpolation=filt, boundary="WRAP")
        assert np.allclose(a, b, atol=1e-12)


def test_nearest_rounds_half_up_and_truncates():
    data = np.arange(5, dtype=np.float64).reshape(1, 1, 1, 5, 1)
    pos = np.array([[1.49, 0, 0], [1.5, 0, 0], [2.51, 0, 0], [-0.4, 0, 0], [-0.6, 0, 0], [4.49, 0, 0]])
    got = _lut_sample(data, pos, interpolation="NEAREST", boundary="BORDER")[:, 0]
    # index = trunc(p + 0.5): -0.6 -> trunc(-0.1) = 0 (C truncation toward zero)
    assert list(got) == [1, 2, 3, 0, 0, 4]


def test_min_max_are_corner_extrema():
    data = RNG.uniform(0.1, 1, size=(1, 4, 4, 4, 2))
    pos = RNG.uniform(-0.7, 3.7, size=(300, 3))
    lo = _lut_sample(data, pos, interpolation="MIN", boundary="BORDER")
    hi = _lut_sample(data, pos, interpolation="MAX", boundary="BORDER")
    for i, p in enumerate(pos):
        if not all(0 <= p[a] + 0.5 < 4 for a in range(3)):
            assert np.all(lo[i] == 0) and np.all(hi[i] == 0)
            continue
        corners, invalid = [], False
        for c in itertools.product(*[sorted({int(np.floor(p[a])), int(np.ceil(p[a]))}) for a in range(3)]):
            if all(0 <= c[a] < 4 for a in range(3)):
                corners.append(data[0, c[2], c[1], c[0]])
            else:
                invalid = True
        corners = np.array(corners)
        start = 0.0 if invalid else None
        mn = corners.min(axis=0) if start is None else np.minimum(corners.min(axis=0), 0.0)
        mx = corners.max(axis=0) if start is None else np.maximum(corners.max(axis=0), 0.0)
        assert np.array_equal(lo[i], mn) and np.array_equal(hi[i], mx)


@pytest.mark.parametrize("mode", ["TRANSFORM", "TRANSFORM_REVERSE"])
def test_identity_camera_resample_is_identity(mode):
    dhw = (4, 8, 16)
    data = RNG.uniform(-1, 1, size=(1,) + dhw + (2,))
    M, V, P, F = _identity_camera(dhw)
    for filt in ("LINEAR", "NEAREST"):
        out = orc.sample_transform_fwd(data, M, V, P, F, dhw, filt, "BORDER", coordinate_mode=mode)
        assert np.allclose(out[:, 0], data, atol=1e-12)


def _numpy_geometry(M_cm, V_cm, P_cm, F, out_dhw, mode):
    """Object-space position of every output cell centre, computed with
    numpy's own inverse from the row-major OpenGL matrices (an independent
    restatement of transformations.hpp:120-145 for the forward modes)."""
    M, V, P = M_cm.T.astype(np.float64), V_cm.T.astype(np.float64), P_cm.T.astype(np.float64)
    D, H, W = out_dhw
    z, y, x = np.meshgrid(np.arange(D) + 0.5, np.arange(H) + 0.5, np.arange(W) + 0.5, indexing="ij")
    ndc = np.stack([x / W * 2 - 1, y / H * 2 - 1, z / D * 2 - 1, np.ones_like(x)], -1)
    if mode == "TRANSFORM_LINDEPTH":
        n, f = float(F[0]), float(F[1])
        zv = -(n + z / D * (f - n))
        ndc[..., 2] = -P[2, 2] - P[2, 3] / zv
    vs = ndc @ np.linalg.inv(P).T
    vs = vs / vs[..., 3:4]
    return (vs @ np.linalg.inv(V @ M).T)[..., :3]


@pytest.mark.parametrize("mode", ["TRANSFORM_LINDEPTH", "TRANSFORM"])
def test_forward_modes_match_numpy_geometry(mode):
    """An affine field a.p + b is reproduced exactly by trilinear sampling,
    so the resample at every in-hull cell equals the field at the position
    numpy computes from the camera."""
    from helpers import scene
    grid = (8, 10, 9)
    sc = scene(grid=grid, image=(7, 6), views=2, seed=2)
    a, b = np.array([0.3, -0.2, 0.45]), 0.7
    D, H, W = grid
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    field = (a[0] * (x + 0.5) + a[1] * (y + 0.5) + a[2] * (z + 0.5) + b)[None, ..., None]
    out_dhw = (11, 7, 6)
    out = orc.sample_transform_fwd(field, sc["M"], sc["V"], sc["P"], sc["F"], out_dhw, "LINEAR", "BORDER",
                                   coordinate_mode=mode)
    checked = 0
    for v in range(2):
        pos = _numpy_geometry(sc["M"][0], sc["V"][v], sc["P"][v], sc["F"][v], out_dhw, mode)
        inside = np.all((pos - 0.5 >= 0) & (pos - 0.5 <= np.array([W - 1, H - 1, D - 1])), axis=-1)
        expect = pos @ a + b
        assert np.allclose(out[0, v, ..., 0][inside], expect[inside], atol=1e-9)
        assert np.all(out[0, v, ..., 0][~inside & np.all((pos + 0.5 < 0) | (pos - 0.5 > np.array([W, H, D])),
                                                        axis=-1)] == 0)
        checked += inside.sum()
    assert checked > 50


def _numpy_lindepth_forward(M_cm, V_cm, P_cm, F, q, dims):
    """Object-space position of frustum index position q (any point, not only
    cell centres) under TRANSFORM_LINDEPTH, from numpy's inverses: the same
    restatement as _numpy_geometry, which the affine-field test pins."""
    M, V, P = M_cm.T.astype(np.float64), V_cm.T.astype(np.float64), P_cm.T.astype(np.float64)
    D, H, W = dims
    n, f = float(F[0]), float(F[1])
    ndc = np.stack([q[..., 0] / W * 2 - 1, q[..., 1] / H * 2 - 1, np.zeros_like(q[..., 0]), np.ones_like(q[..., 0])],
                   -1)
    zv = -(n + q[..., 2] / D * (f - n))
    ndc[..., 2] = -P[2, 2] - P[2, 3] / zv
    vs = ndc @ np.linalg.inv(P).T
    vs = vs / vs[..., 3:4]
    return (vs @ np.linalg.inv(V @ M).T)[..., :3]


def _numpy_perspective_reverse(M_cm, V_cm, P_cm, F, c, dims):
    """Frustum index position of object-space points c under a perspective
    camera with depth linear in view space, computed directly: x, y from the
    projected NDC, z from the view-space depth itself, t = (-z_view - near) /
    (far - near). The reference recovers z_view from NDC depth instead
    (transformations.hpp:106-117); for a perspective projection both agree."""
    M, V, P = M_cm.T.astype(np.float64), V_cm.T.astype(np.float64), P_cm.T.astype(np.float64)
    D, H, W = dims
    hom = np.concatenate([c, np.ones_like(c[..., :1])], -1)
    vs = hom @ (V @ M).T
    clip = vs @ P.T
    ndc = clip[..., :3] / clip[..., 3:4]
    t = (-vs[..., 2] - float(F[0])) / (float(F[1]) - float(F[0]))
    return np.stack([(ndc[..., 0] * 0.5 + 0.5) * W, (ndc[..., 1] * 0.5 + 0.5) * H, t * D], -1)


def _cell_centres(dhw):
    D, H, W = dhw
    z, y, x = np.meshgrid(np.arange(D) + 0.5, np.arange(H) + 0.5, np.arange(W) + 0.5, indexing="ij")
    return np.stack([x, y, z], -1)


def test_lindepth_reverse_positions_match_numpy():
    """TRANSFORM_LINDEPTH_REVERSE maps every grid cell centre to the frustum
    index position numpy computes directly from the camera (perspective)."""
    from helpers import scene
    grid = (9, 12, 10)
    sc = scene(grid=grid, image=(13, 7), views=3, seed=4)
    fr = (17, 13, 7)
    pos = orc.sample_positions(sc["M"], sc["V"], sc["P"], sc["F"], fr, grid, "TRANSFORM_LINDEPTH_REVERSE")
    c = _cell_centres(grid)
    for v in range(3):
        want = _numpy_perspective_reverse(sc["M"][0], sc["V"][v], sc["P"][v], sc["F"][v], c, fr)
        assert np.abs(pos[v, ..., :3] - want).max() <= 1e-9 * max(1.0, np.abs(want).max())
        assert np.ptp(want[..., 2]) > 3  # the depth index really varies over the grid


@pytest.mark.parametrize("perspective", [True, False])
def test_lindepth_reverse_inverts_lindepth(perspective):
    """LINDEPTH o LINDEPTH_REVERSE is the identity: the reverse position of a
    cell centre, mapped forward again with the (pinned) numpy restatement of
    TRANSFORM_LINDEPTH, is that cell centre. Orthographic cameras too, whose
    NDC depth the reference's linear-depth formula treats as perspective."""
    from gft import synthetic as syn
    from gft.camera import camera_params_batch, model_matrices
    from helpers import ortho_camera, scene
    grid = (8, 11, 9)
    fr = (15, 12, 10)
    if perspective:
        sc = scene(grid=grid, image=(10, 12), views=2, seed=6)
        M, V, P, F = sc["M"], sc["V"], sc["P"], sc["F"]
    else:
        cams = [ortho_camera(grid, (11, 9), 15, yaw=y) for y in (0.0, 33.0)]
        M = model_matrices([syn.grid_transform(grid)])
        V, P, F = camera_params_batch(cams)
    pos = orc.sample_positions(M, V, P, F, fr, grid, "TRANSFORM_LINDEPTH_REVERSE")
    c = _cell_centres(grid)
    for v in range(2):
        back = _numpy_lindepth_forward(M[0], V[v], P[v], F[v], pos[v, ..., :3], fr)
        assert np.abs(back - c).max() <= 1e-8 * np.abs(c).max()


@pytest.mark.parametrize("k", [1, 2])
@pytest.mark.parametrize("mode", ["TRANSFORM_REVERSE", "TRANSFORM_LINDEPTH_REVERSE"])
def test_reverse_lod_of_a_finer_input_is_k(mode, k):
    """Identity camera with an input 2^k times finer per axis than the output
    grid: each output cell step moves 2^k input cells, so LoD = k."""
    dhw = (4, 8, 8)
    M, V, P, F = _identity_camera(dhw)
    fine = tuple(s << k for s in dhw)
    pos = orc.sample_positions(M, V, P, F, fine, dhw, mode, lod=True)
    assert np.allclose(pos[..., 3], k, atol=1e-12)


def test_reverse_lod_uses_jacobian_columns():
    """The reverse modes' LoD is log2 of the largest Jacobian COLUMN norm of
    the cell -> frustum-index map (sampling.hpp:124-136; the forward modes use
    rows), with forward differences p(c) - p(c + e_a). Checked against numpy's
    direct map under an anisotropic perspective camera, where rows and columns
    give different LoDs."""
    from gft.camera import Camera, camera_params_batch, model_matrices
    from gft import synthetic as syn
    from gft.transform import GridTransform, Transform
    grid = (6, 7, 8)
    fr = (40, 21, 33)  # fine, anisotropic frustum grid: LoD > 0 everywhere
    t = GridTransform([fr[0], fr[1], fr[2]], translation=[0, 0, 14],
                      parent=Transform(rotation_deg=[20, 35, 10], translation=[4, 3.5, 3]))
    cam = Camera(t, nearFar=[8.0, 20.0], fov=40, aspect=fr[2] / fr[1])
    M = model_matrices([syn.grid_transform(grid)])
    V, P, F = camera_params_batch([cam])
    pos = orc.sample_positions(M, V, P, F, fr, grid, "TRANSFORM_LINDEPTH_REVERSE", lod=True)
    c = _cell_centres(grid)
    p0 = _numpy_perspective_reverse(M[0], V[0], P[0], F[0], c, fr)
    J = [p0 - _numpy_perspective_reverse(M[0], V[0], P[0], F[0], c + e, fr) for e in np.eye(3)]
    cols = np.stack([np.linalg.norm(j, axis=-1) for j in J], -1)
    rows = np.stack([np.linalg.norm(np.stack([j[..., a] for j in J], -1), axis=-1) for a in range(3)], -1)
    lod_cols = np.maximum(0, np.log2(cols.max(-1)))
    lod_rows = np.maximum(0, np.log2(rows.max(-1)))
    assert np.abs(lod_cols - lod_rows).max() > 0.05
    assert lod_cols.min() > 0.5
    assert np.abs(pos[0, ..., 3] - lod_cols).max() <= 1e-8


def test_raymarch_equals_depth_sum_of_resample_for_ortho():
    """With an orthographic camera the raymarch samples (uniform in object
    space between the near and far cell centres) coincide with the TRANSFORM
    resample's cell centres, so ADDITIVE raymarch == sum over depth."""
    from gft import synthetic as syn
    from gft.camera import camera_params_batch, model_matrices
    from helpers import ortho_camera
    grid = (6, 8, 7)
    cam = ortho_camera(grid, (8, 7), 9, yaw=25.0)
    M = model_matrices([syn.grid_transform(grid)])
    V, P, F = camera_params_batch([cam])
    out_dhw = (9, 8, 7)
    data = RNG.uniform(0, 1, size=(1,) + grid + (1,))
    ray = orc.raymarch_fwd(data, M, V, P, F, out_dhw, "LINEAR", "BORDER", "ADDITIVE")
    fr = orc.sample_transform_fwd(data, M, V, P, F, out_dhw, "LINEAR", "BORDER", coordinate_mode="TRANSFORM")
    assert np.abs(fr).max() > 0.1
    assert np.allclose(ray[0, 0], fr[0, 0].sum(0), atol=1e-10)


# ------------------------------------------------------------ mips (A2)
def test_mips_are_block_means():
    data = RNG.uniform(0, 1, size=(9, 8, 12, 3))
    lv = orc.build_mips(data, 3)
    assert [l.shape for l in lv] == [(4, 4, 6, 3), (2, 2, 3, 3), (1, 1, 1, 3)]
    for m, l in enumerate(lv, start=1):
        s = 2 ** m
        for k, j, i in itertools.product(range(l.shape[0]), range(l.shape[1]), range(l.shape[2])):
            block = data[k * s:(k + 1) * s, j * s:(j + 1) * s, i * s:(i + 1) * s]
            assert np.allclose(l[k, j, i], block.reshape(-1, 3).mean(0), atol=1e-12)
    const = np.full((8, 8, 8, 1), 0.37)
    assert all(np.allclose(l, 0.37, atol=1e-15) for l in orc.build_mips(const, 3))


@pytest.mark.parametrize("k", [1, 2])
def test_lod_selects_level_of_downscaled_identity(k):
    """Output cells 2^k times larger than the grid cells: LoD = k exactly, so
    NEAREST mips read level k at its own cell centres."""
    dhw = (8, 16, 16)
    data = RNG.uniform(0, 1, size=(1,) + dhw + (1,))
    out_dhw = tuple(s >> k for s in dhw)
    M, V, P, F = _identity_camera(dhw)
    # index map of the output: A_out maps output ndc to grid index, so reuse
    # the grid's camera; the output grid is coarser by 2^k per axis
    out = orc.sample_transform_fwd(data, M, V, P, F, out_dhw, "LINEAR", "CLAMP", "NEAREST", 3, 0.0, "TRANSFORM")
    ref = orc.build_mips(data[0], 3)[k - 1]
    assert np.allclose(out[0, 0], ref, atol=1e-12)


def test_linear_mip_lerp_weight_is_fract_lod():
    """Checkerboard along x: level 1 is 0 everywhere, so a LINEAR-mip sample
    at a level-0 cell centre is (1 - fract(lod)) * value; NEAREST mips switch
    level at lod 0.5."""
    D, H, W = 4, 4, 8
    x = np.arange(W)
    data = np.broadcast_to(np.where(x % 2 == 0, 1.0, -1.0), (1, D, H, W))[..., None].copy()
    pos = np.array([[i + 0.5, 1.5, 1.5] for i in range(W)])
    for lod in (0.0, 0.25, 0.5, 0.8):
        got = _lut_sample(data, pos, lod=np.full(W, lod), boundary="CLAMP", mipmapping="LINEAR", num_mipmaps=1,
                          cell_center_offset=0.5)[:, 0]
        assert np.allclose(got, (1 - lod) * data[0, 1, 1, :, 0], atol=1e-12)
        nearest = _lut_sample(data, pos, lod=np.full(W, lod), boundary="CLAMP", mipmapping="NEAREST", num_mipmaps=1,
                              cell_center_offset=0.5)[:, 0]
        assert np.allclose(nearest, data[0, 1, 1, :, 0] if lod < 0.5 else 0.0, atol=1e-12)


@pytest.mark.parametrize("inverse", [False, True])
def test_lod_transform_matches_its_own_jacobian(inverse):
    from helpers import scene
    sc = scene(grid=(8, 10, 12), image=(9, 7), views=2, seed=3)
    MV = np.stack([(sc["V"][i].T @ sc["M"][0].T).T for i in range(2)]).astype(np.float32)
    grid_dhw, fr_dhw = (8, 10, 12), tuple(sc["out_dhw"])
    in_dhw, out_dhw = (fr_dhw, grid_dhw) if inverse else (grid_dhw, fr_dhw)
    out = orc.lod_transform(in_dhw, MV, sc["P"], sc["F"], out_dhw, inverse)
    pos = out[..., :3]
    # central differences of the position output at interior cells
    jx = (pos[:, 1:-1, 1:-1, 2:] - pos[:, 1:-1, 1:-1, :-2]) / 2
    jy = (pos[:, 1:-1, 2:, 1:-1] - pos[:, 1:-1, :-2, 1:-1]) / 2
    jz = (pos[:, 2:, 1:-1, 1:-1] - pos[:, :-2, 1:-1, 1:-1]) / 2
    foot = np.sqrt(jx ** 2 + jy ** 2 + jz ** 2)  # per component a: |d p_a / d (x, y, z)|
    lod = np.maximum(0.0, np.log2(foot.max(-1)))
    assert np.allclose(out[:, 1:-1, 1:-1, 1:-1, 3], lod, atol=1e-9)


# ---------------------------------------------------------- raymarch (A5)
def _cams(views=2, seed=0):
    from helpers import scene
    return scene(grid=(6, 8, 7), image=(5, 4), views=views, seed=seed)


def test_raymarch_constant_density_closed_forms():
    sc = _cams()
    S = sc["out_dhw"][0]
    e, s = np.array([0.3, 0.5, 0.9]), 0.07
    data = np.broadcast_to(np.concatenate([e, [s]]), (1, 6, 8, 7, 4)).copy()
    i = np.arange(1, S + 1)
    expect = {"BEER_LAMBERT": np.concatenate([e * np.exp(-i * s).sum(), [S * s]]),
              "ADDITIVE": S * np.concatenate([e, [s]]),
              "ALPHA": np.concatenate([e * ((1 - s) ** (i - 1)).sum(), [1 - (1 - s) ** S]]),
              "ALPHA_ADDITIVE": np.concatenate([S * (1 - s) * e, [S * s]])}
    for blend, ref in expect.items():
        out = orc.raymarch_fwd(data, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], "LINEAR", "CLAMP", blend)
        assert np.allclose(out, ref, rtol=1e-12, atol=1e-12), blend
    d1 = np.full((1, 6, 8, 7, 1), s)
    assert np.allclose(orc.raymarch_fwd(d1, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], boundary="CLAMP"),
                       S * s, atol=1e-12)
    d2 = np.broadcast_to(np.array([0.4, s]), (1, 6, 8, 7, 2)).copy()
    out2 = orc.raymarch_fwd(d2, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], boundary="CLAMP")
    assert np.allclose(out2[..., 0], 0.4 * np.exp(-i * s).sum(), atol=1e-12)


# ---------------------------------------------------- EXACT adjoints (FD64)
def _fd(fun, x, idx, h):
    xp, xm = x.copy(), x.copy()
    xp.flat[idx] += h
    xm.flat[idx] -= h
    return (fun(xp) - fun(xm)) / (2 * h)


@pytest.mark.parametrize("blend,channels", [("BEER_LAMBERT", 4), ("BEER_LAMBERT", 2), ("ALPHA", 4),
                                            ("ALPHA_ADDITIVE", 4), ("ADDITIVE", 1)])
@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP"])
def test_raymarch_exact_gradient_fd(blend, channels, boundary):
    sc = _cams(seed=5)
    data = RNG.uniform(0.0, 0.15, size=(1, 6, 8, 7, channels))
    out = orc.raymarch_fwd(data, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], "LINEAR", boundary, blend)
    y = RNG.uniform(-1, 1, size=out.shape)
    g = orc.raymarch_bwd(data, out, y, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], boundary, blend,
                         grad_mode="EXACT")

    def loss(d):
        return float((orc.raymarch_fwd(d, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], "LINEAR", boundary,
                                       blend) * y).sum())
    cand = np.argsort(-np.abs(g).ravel())[:6].tolist() + RNG.integers(0, data.size, 4).tolist()
    for idx in cand:
        assert abs(_fd(loss, data, idx, 1e-5) - g.flat[idx]) <= 1e-6 * max(np.abs(g).max(), 1e-12)


@pytest.mark.parametrize("mode", ["TRANSFORM_LINDEPTH", "TRANSFORM_REVERSE"])
@pytest.mark.parametrize("filt,mip", [("LINEAR", "NONE"), ("NEAREST", "NONE"), ("LINEAR", "LINEAR"),
                                      ("NEAREST", "NEAREST")])
def test_resample_exact_adjoint_identity(mode, filt, mip):
    sc = _cams(seed=6)
    rev = mode.endswith("REVERSE")
    in_dhw = tuple(sc["out_dhw"]) if rev else (6, 8, 7)
    out_dhw = (6, 8, 7) if rev else tuple(sc["out_dhw"])
    x = RNG.uniform(-1, 1, size=(1,) + in_dhw + (2,))
    kw = dict(interpolation=filt, boundary="BORDER", mipmapping=mip, num_mipmaps=2 if mip != "NONE" else 0,
              mip_bias=0.3, coordinate_mode=mode)
    ax = orc.sample_transform_fwd(x, sc["M"], sc["V"], sc["P"], sc["F"], out_dhw, **kw)
    y = RNG.uniform(-1, 1, size=ax.shape)
    aty = orc.sample_transform_bwd(x, y, sc["M"], sc["V"], sc["P"], sc["F"], grad_mode="EXACT", **kw)
    lhs, rhs = float((ax * y).sum()), float((x * aty).sum())
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), 1.0)


@pytest.mark.parametrize("filt", ["LINEAR", "NEAREST", "MIN", "MAX"])
@pytest.mark.parametrize("boundary", ["BORDER", "CLAMP", "WRAP"])
def test_lut_exact_input_adjoint(filt, boundary):
    """The sample is linear in the data for LINEAR/NEAREST and piecewise
    linear (a selection) for MIN/MAX: <A x, y> == <x, A^T y>."""
    x = RNG.uniform(0.1, 1, size=(1, 4, 5, 6, 2))
    lut = RNG.uniform(-2, 2, size=(2, 3, 4, 5, 4))
    kw = dict(interpolation=filt, boundary=boundary)
    ax = orc.sample_lut_fwd(x, lut, **kw)
    y = RNG.uniform(-1, 1, size=ax.shape)
    gin, _ = orc.sample_lut_bwd(x, lut, y, grad_mode="EXACT", **kw)
    assert abs(float((ax * y).sum()) - float((x * gin).sum())) <= 1e-10 * max(abs(float((ax * y).sum())), 1)


@pytest.mark.parametrize("normalized", [False, True])
def test_lut_exact_lookup_gradient_fd(normalized):
    x = RNG.uniform(0.1, 1, size=(1, 4, 5, 6, 2))
    lut = RNG.uniform(-2, 2, size=(1, 3, 4, 5, 4))
    kw = dict(interpolation="LINEAR", boundary="CLAMP")
    if normalized:
        lut[..., :3] = RNG.uniform(0.0, 1.0, size=lut[..., :3].shape)
        kw.update(relative_coords=False, normalized_coords=True, cell_center_offset=0.5)
    y = RNG.uniform(-1, 1, size=(1, 1, 3, 4, 5, 2))
    _, glut = orc.sample_lut_bwd(x, lut, y, grad_mode="EXACT", **kw)

    def loss(l):
        return float((orc.sample_lut_fwd(x, l, **kw) * y).sum())
    for idx in RNG.integers(0, lut.size, 12):
        if idx % 4 == 3:
            continue
        assert abs(_fd(loss, lut, int(idx), 1e-7) - glut.flat[idx]) <= 1e-5 * max(np.abs(glut).max(), 1e-9)


@pytest.mark.parametrize("blend", ["BEER_LAMBERT", "ALPHA", "ALPHA_ADDITIVE", "ADDITIVE"])
@pytest.mark.parametrize("keep", [False, True])
def test_reduce_blend_exact_gradient_fd(blend, keep):
    x = RNG.uniform(0.0, 0.3, size=(1, 5, 2, 3, 4))
    out = orc.reduce_blend_fwd(x, blend, keep)
    y = RNG.uniform(-1, 1, size=out.shape)
    g = orc.reduce_blend_bwd(x, out, y, blend, keep, grad_mode="EXACT")

    def loss(d):
        return float((orc.reduce_blend_fwd(d, blend, keep) * y).sum())
    for idx in range(0, x.size, 7):
        assert abs(_fd(loss, x, idx, 1e-6) - g.flat[idx]) <= 1e-7 * max(np.abs(g).max(), 1)


def test_reduce_blend_beer_lambert_closed_form():
    s = RNG.uniform(0, 0.3, size=(1, 6, 2, 2, 1))
    e = RNG.uniform(0, 1, size=(1, 6, 2, 2, 3))
    out = orc.reduce_blend_fwd(np.concatenate([e, s], -1), "BEER_LAMBERT", False)
    tau = np.cumsum(s, axis=1)
    assert np.allclose(out[..., 3], tau[:, -1, ..., 0], atol=1e-12)
    assert np.allclose(out[..., :3], (np.exp(-tau) * e).sum(1), atol=1e-12)


# -------------------------------------------- REFERENCE-mode invariants
def test_reference_raymarch_gradient_normalises_per_camera():
    """Constant output gradient c with ADDITIVE: each camera's scatter is
    divided by its own weight sum, so every touched voxel gets c per camera."""
    sc = _cams(views=1, seed=7)
    data = RNG.uniform(0, 1, size=(1, 6, 8, 7, 1))
    out = orc.raymarch_fwd(data, sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], "LINEAR", "BORDER", "ADDITIVE")
    g = orc.raymarch_bwd(data, out, np.full(out.shape, 0.7), sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"],
                         "BORDER", "ADDITIVE", grad_mode="REFERENCE")
    touched = g != 0
    assert touched.mean() > 0.3
    assert np.allclose(g[touched], 0.7, atol=1e-12)


def test_reference_resample_gradient_counts_every_corner():
    """Identity camera, constant output gradient c: every output sample sits
    on a grid point, where floor == ceil, so each of the 8 corner visits hits
    the same voxel: weight sum 1, count 8, gradient c / 8 (weight-independent
    counting of the reference)."""
    dhw = (4, 8, 8)
    M, V, P, F = _identity_camera(dhw)
    x = RNG.uniform(size=(1,) + dhw + (1,))
    g = orc.sample_transform_bwd(x, np.full((1, 1) + dhw + (1,), 0.8), M, V, P, F, coordinate_mode="TRANSFORM",
                                 grad_mode="REFERENCE")
    assert np.allclose(g, 0.1, atol=1e-12)
    ge = orc.sample_transform_bwd(x, np.full((1, 1) + dhw + (1,), 0.8), M, V, P, F, coordinate_mode="TRANSFORM",
                                  grad_mode="EXACT")
    assert np.allclose(ge, 0.8, atol=1e-12)


def test_reference_lut_gradient_vanishes_at_grid_points():
    """read3DGrad uses floor/ceil corners, so at integer positions its spatial
    derivative is 0; the EXACT gradient is the one-sided slope."""
    D, H, W = 4, 5, 6
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    f = (0.5 * x + 0.25 * y - 1.0 * z)[None, ..., None]
    lut = np.zeros((1, D, H, W, 4))
    og = np.ones((1, 1, D, H, W, 1))
    _, gref = orc.sample_lut_bwd(f, lut, og, boundary="CLAMP", grad_mode="REFERENCE")
    _, gex = orc.sample_lut_bwd(f, lut, og, boundary="CLAMP", grad_mode="EXACT")
    assert np.all(gref == 0)
    inner = gex[0, :-1, :-1, :-1]
    assert np.allclose(inner[..., 0], 0.5) and np.allclose(inner[..., 1], 0.25) and np.allclose(inner[..., 2], -1.0)


# ------------------------------------------------------- advection (A10)
def _linear_field(D, H, W):
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    return (0.2 * x - 0.3 * y + 0.4 * z + 1.0)[None, ..., None]


@pytest.mark.parametrize("order,clamp", [(1, "NONE"), (2, "NONE"), (2, "MC_SMOOTH"), (1, "NEGATIVE")])
def test_advect_zero_velocity_is_identity(order, clamp):
    d = RNG.uniform(0, 1, size=(1, 5, 6, 7, 1))
    out = orc.advect_fwd(d, np.zeros((1, 5, 6, 7, 3)), order, clamp)
    assert np.allclose(out, d, atol=1e-15)


@pytest.mark.parametrize("order", [1, 2])
def test_advect_integer_translation_is_a_shift(order):
    d = RNG.uniform(0, 1, size=(1, 5, 6, 7, 1))
    v = np.zeros((1, 5, 6, 7, 3))
    v[..., 0], v[..., 2] = 2.0, -1.0  # x by +2 cells, z by -1 cell
    out = orc.advect_fwd(d, v, order, "MC_SMOOTH")
    # out(x, y, z) = d(x - 2, y, z + 1) with CLAMP edge replication
    xs = np.clip(np.arange(7) - 2, 0, 6)
    zs = np.clip(np.arange(5) + 1, 0, 4)
    ref = d[:, zs][:, :, :, xs]
    # MacCormack's back-trace b(x) = f(x + v) returns d(x) wherever neither
    # trace leaves the grid: x <= 4, z >= 1
    inner = (slice(None), slice(1, 5), slice(None), slice(0, 5))
    assert np.allclose(out[inner], ref[inner], atol=1e-12)
    if order == 1:
        assert np.allclose(out, ref, atol=1e-12)


def test_maccormack_is_exact_on_linear_fields():
    d = _linear_field(8, 9, 10)
    v = np.broadcast_to(np.array([0.37, -0.61, 0.23]), (1, 8, 9, 10, 3)).copy()
    for order in (1, 2):
        out = orc.advect_fwd(d, v, order, "NONE")
        ref = _linear_field(8, 9, 10) - (0.2 * 0.37 - 0.3 * -0.61 + 0.4 * 0.23)
        inner = (slice(None), slice(2, -2), slice(2, -2), slice(2, -2))
        assert np.allclose(out[inner], ref[inner], atol=1e-12)


def test_mc_smooth_is_bounded_or_first_order():
    d = RNG.uniform(0, 1, size=(1, 8, 8, 8, 1))
    v = RNG.uniform(-1.5, 1.5, size=(1, 8, 8, 8, 3))
    out, parts = orc.advect_fwd(d, v, 2, "MC_SMOOTH", return_parts=True)
    inside = (out >= parts["mn"]) & (out <= parts["mx"])
    first = out == parts["f"]
    assert np.all(inside | first)
    assert np.array_equal(out[parts["mask"]], parts["f"][parts["mask"]])
    assert parts["mask"].any() and (~parts["mask"]).any()


@pytest.mark.parametrize("order,clamp", [(1, "NONE"), (2, "NONE"), (2, "MC_SMOOTH")])
def test_advect_exact_gradient_fd(order, clamp):
    d = RNG.uniform(0, 1, size=(1, 5, 6, 7, 1))
    v = RNG.uniform(-1.2, 1.2, size=(1, 5, 6, 7, 3))
    y = RNG.uniform(-1, 1, size=d.shape)
    gd, gv = orc.advect_bwd(d, v, y, order, clamp, grad_mode="EXACT")
    mask0 = orc.advect_fwd(d, v, order, clamp, return_parts=True)[1].get("mask")
    h = 1e-6

    def central(x, idx, fun):
        """Central difference, or None where the MC_SMOOTH mask switches."""
        xp, xm = x.copy(), x.copy()
        xp.flat[idx] += h
        xm.flat[idx] -= h
        op, pp = fun(xp)
        om, pm = fun(xm)
        if mask0 is not None and not (np.array_equal(pp["mask"], mask0) and np.array_equal(pm["mask"], mask0)):
            return None
        return float(((op - om) * y).sum()) / (2 * h)

    checked = 0
    for idx in RNG.integers(0, d.size, 12):
        fd = central(d, int(idx), lambda dd: orc.advect_fwd(dd, v, order, clamp, return_parts=True))
        if fd is not None:
            assert abs(fd - gd.flat[idx]) <= 1e-6 * np.abs(gd).max()
            checked += 1
    for idx in RNG.integers(0, v.size, 40):
        fd = central(v, int(idx), lambda vv: orc.advect_fwd(d, vv, order, clamp, return_parts=True))
        if fd is not None:
            assert abs(fd - gv.flat[idx]) <= 1e-5 * np.abs(gv).max()
            checked += 1
    assert checked >= 20


def test_transport_render_step_gradient_fd():
    from helpers import scene
    sc = scene(grid=(6, 8, 7), image=(5, 4), views=2, seed=9)
    d0 = sc["data"].astype(np.float64)
    v = RNG.uniform(-0.8, 0.8, size=d0.shape[:-1] + (3,))
    tgt = RNG.uniform(0, 0.5, size=(1, 2) + tuple(sc["out_dhw"][1:]) + (3,))
    args = (sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"])
    r = orc.step_fwd_bwd(d0, v, tgt, *args, order=1, clamp="NONE")

    def loss_d(dd):
        return orc.step_fwd_bwd(dd, v, tgt, *args, order=1, clamp="NONE")["loss"]
    for idx in np.argsort(-np.abs(r["gd"]).ravel())[:5]:
        assert abs(_fd(loss_d, d0, int(idx), 1e-6) - r["gd"].flat[idx]) <= 1e-5 * np.abs(r["gd"]).max()


# ------------------------------------------------- staggered velocity (A10)
def _stag_shapes(D, H, W):
    return [(1, D, H, W + 1, 1), (1, D, H + 1, W, 1), (1, D + 1, H, W, 1)]


def test_face_velocity_interpolates_linear_components_at_the_faces():
    """Each component linear in its own face coordinates: resampled to the
    faces of another component it must equal the same linear function at
    those faces (x-face (i,j,k) sits at centred (i-0.5, j, k), y-face at
    (i, j-0.5, k), z-face at (i, j, k-0.5))."""
    D, H, W = 6, 7, 8
    a = {"X": (0.3, -0.2, 0.1), "Y": (-0.4, 0.25, 0.05), "Z": (0.15, 0.1, -0.3)}
    shift = {"X": (-0.5, 0, 0), "Y": (0, -0.5, 0), "Z": (0, 0, -0.5)}

    def field(name, shape):
        z, y, x = np.meshgrid(np.arange(shape[1]), np.arange(shape[2]), np.arange(shape[3]), indexing="ij")
        s = shift[name]
        cx, cy, cz = x + s[0], y + s[1], z + s[2]  # centred coordinates of the faces
        return (a[name][0] * cx + a[name][1] * cy + a[name][2] * cz)[None, ..., None]
    comps = [field(n, s) for n, s in zip("XYZ", _stag_shapes(D, H, W))]
    for ti, target in enumerate("XYZ"):
        fv = orc.face_velocity(comps, target)
        shape = _stag_shapes(D, H, W)[ti]
        z, y, x = np.meshgrid(np.arange(shape[1]), np.arange(shape[2]), np.arange(shape[3]), indexing="ij")
        s = shift[target]
        cx, cy, cz = x + s[0], y + s[1], z + s[2]
        for ci, n in enumerate("XYZ"):
            expect = a[n][0] * cx + a[n][1] * cy + a[n][2] * cz
            # away from the CLAMP edges, where the half-cell offset leaves the component's grid
            inner = (slice(1, -1),) * 3
            assert np.allclose(fv[0, ..., ci][inner], expect[inner], atol=1e-12), (target, n)


def test_velocity_self_advection_zero_and_uniform():
    D, H, W = 5, 6, 7
    zero = [np.zeros(s) for s in _stag_shapes(D, H, W)]
    out = orc.velocity_self_advect(zero, 2, "MC_SMOOTH")
    assert all(np.all(o == 0) for o in out)
    # a uniform velocity (1, 0, 0) cell/step: every component field is constant, so it is unchanged
    uni = [np.full(s, v) for s, v in zip(_stag_shapes(D, H, W), (1.0, 0.0, 0.0))]
    out = orc.velocity_self_advect(uni, 2, "MC_SMOOTH")
    assert all(np.allclose(o, u, atol=1e-14) for o, u in zip(out, uni))


def test_velocity_self_advection_exact_gradient_fd():
    D, H, W = 4, 5, 6
    comps = [RNG.uniform(-0.6, 0.6, size=s) for s in _stag_shapes(D, H, W)]
    ws = [RNG.uniform(-1, 1, size=s) for s in _stag_shapes(D, H, W)]
    g = orc.velocity_self_advect_bwd(comps, ws, 2, "NONE")

    def loss(cs):
        return float(sum((o * w).sum() for o, w in zip(orc.velocity_self_advect(cs, 2, "NONE"), ws)))
    h = 1e-6
    gmax = max(np.abs(x).max() for x in g)
    for ci in range(3):
        for idx in RNG.integers(0, comps[ci].size, 6):
            cp = [c.copy() for c in comps]
            cm = [c.copy() for c in comps]
            cp[ci].flat[idx] += h
            cm[ci].flat[idx] -= h
            fd = (loss(cp) - loss(cm)) / (2 * h)
            assert abs(fd - g[ci].flat[idx]) <= 1e-5 * gmax


# --------------------------------------------------- transport chain (A12)
def test_chain_exact_gradient_is_the_fd_of_the_summed_loss():
    from helpers import scene
    sc = scene(grid=(5, 7, 6), image=(5, 4), views=2, seed=13)
    d0 = sc["data"].astype(np.float64)
    T = 3
    vels = [RNG.uniform(-0.7, 0.7, size=d0.shape[:-1] + (3,)) for _ in range(T - 1)]
    tg = [RNG.uniform(0, 0.4, size=(1, 2) + tuple(sc["out_dhw"][1:]) + (3,)) for _ in range(T)]
    args = (sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"])
    r = orc.chain_fwd_bwd(d0, vels, tg, *args, order=1, clamp="NONE")
    assert len(r["losses"]) == T

    def loss_d(dd):
        return orc.chain_fwd_bwd(dd, vels, tg, *args, order=1, clamp="NONE")["loss"]
    for idx in np.argsort(-np.abs(r["gd0"]).ravel())[:4]:
        assert abs(_fd(loss_d, d0, int(idx), 1e-6) - r["gd0"].flat[idx]) <= 1e-5 * np.abs(r["gd0"]).max()
    for f in range(T - 1):
        def loss_v(vv, f=f):
            vs = list(vels)
            vs[f] = vv
            return orc.chain_fwd_bwd(d0, vs, tg, *args, order=1, clamp="NONE")["loss"]
        for idx in np.argsort(-np.abs(r["gv"][f]).ravel())[:3]:
            assert abs(_fd(loss_v, vels[f], int(idx), 1e-6) - r["gv"][f].flat[idx]) <= \
                1e-5 * np.abs(r["gv"][f]).max()


@pytest.mark.parametrize("decay", [0.0, 0.9])
def test_chain_reference_recursion_closed_form(decay):
    """With zero velocity every advection is the identity (A_f = I for the
    density), so the reference's warped-gradient recursion
    (optimization.py:1017-1031, BWD order) has a closed form in the frames'
    own render gradients g_f, here for T = 3 with three different targets:
    decay a > 0: lambda_2 = (1-a) g_2, lambda_1 = (1-a) g_1 + a lambda_2,
    grad d0 = g_0 + lambda_1; decay 0: lambda_1 = g_1 + g_2 and the warped term
    applied to frame 0 is divided by num_step = 2: grad d0 = g_0 + (g_1 + g_2)/2.
    The velocity gradients at decay 0 are the EXACT ones divided by
    num_step = T-1-f (frame 1: 1, frame 0: 2), at decay 0.9 equal to the
    EXACT adjoint of the decayed accumulator."""
    from helpers import scene
    sc = scene(grid=(5, 7, 6), image=(5, 4), views=2, seed=17)
    d0 = sc["data"].astype(np.float64)
    T = 3
    zero = [np.zeros(d0.shape[:-1] + (3,)) for _ in range(T - 1)]
    tg = [RNG.uniform(0, 0.4, size=(1, 2) + tuple(sc["out_dhw"][1:]) + (3,)) for _ in range(T)]
    args = (sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"])
    g = []
    for f in range(T):  # each frame's own render-loss gradient of the same (unmoved) density
        img = orc.render_density_fwd(d0, *args)
        _, gimg = orc.l2_loss_and_grad(img, tg[f])
        g.append(orc.render_density_bwd(d0, img, gimg, *args, grad_mode="EXACT"))
    assert all(np.abs(g[f] - g[0]).max() > 1e-3 * np.abs(g[0]).max() for f in (1, 2))
    r = orc.chain_fwd_bwd(d0, zero, tg, *args, order=2, clamp="MC_SMOOTH", grad_mode="REFERENCE", decay=decay)
    a = decay
    want = g[0] + (g[1] + g[2]) / 2 if a == 0 else g[0] + (1 - a) * g[1] + a * (1 - a) * g[2]
    assert np.abs(r["gd0"] - want).max() <= 1e-12 * np.abs(want).max()
    ex = orc.chain_fwd_bwd(d0, zero, tg, *args, order=2, clamp="MC_SMOOTH", grad_mode="EXACT")
    assert np.abs(ex["gd0"] - (g[0] + g[1] + g[2])).max() <= 1e-12 * np.abs(ex["gd0"]).max()
    if a == 0:
        for f, n in ((0, 2), (1, 1)):
            assert np.abs(r["gv"][f] - ex["gv"][f] / n).max() <= 1e-12 * np.abs(ex["gv"][f]).max()
    else:
        # frame 1's warped term is A_1^T lambda_2 = (1-a) A_1^T g_2: a scaled copy of EXACT's
        assert np.abs(r["gv"][1] - (1 - a) * ex["gv"][1]).max() <= 1e-12 * np.abs(ex["gv"][1]).max()
    # a single frame: nothing is warped, the gradient is the frame's own
    r1 = orc.chain_fwd_bwd(d0, [], tg[:1], *args, grad_mode="REFERENCE", decay=decay)
    assert np.array_equal(r1["gd0"], g[0])


# ----------------------------------------------------- lit density render
def test_lit_render_reduces_to_the_ambient_render_and_fd():
    """A constant light factor k with ambient a is the ambient render of
    intensity k + a (closed form of the light grid); the gradients w.r.t. the
    density and the light factor are the central differences of <img, w>."""
    from helpers import scene
    sc = scene(grid=(6, 7, 5), image=(6, 5), views=2, seed=23)
    d = sc["data"].astype(np.float64)
    args = (sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"])
    color, a = (0.9, 0.6, 0.3), 0.4
    lam = np.full(d.shape, 0.7)
    img = orc.render_lit_fwd(d, lam, *args, color=color, ambient=a)
    ref = orc.render_density_fwd(d, *args, light_color=color, light_intensity=0.7 + a)
    assert np.abs(img - ref).max() <= 1e-13 * np.abs(ref).max()
    lam = RNG.uniform(0, 1.5, size=d.shape)
    img = orc.render_lit_fwd(d, lam, *args, color=color, ambient=a)
    w = RNG.uniform(-1, 1, size=img.shape)
    gd, gl = orc.render_lit_bwd(d, lam, img, w, *args, color=color, ambient=a)

    def loss_d(x):
        return float((orc.render_lit_fwd(x, lam, *args, color=color, ambient=a) * w).sum())

    def loss_l(x):
        return float((orc.render_lit_fwd(d, x, *args, color=color, ambient=a) * w).sum())
    for g, fn, x in ((gd, loss_d, d), (gl, loss_l, lam)):
        assert np.abs(g).max() > 0
        for idx in np.argsort(-np.abs(g).ravel())[:5]:
            assert abs(_fd(fn, x, int(idx), 1e-6) - g.flat[idx]) <= 1e-6 * np.abs(g).max()


# ------------------------------------------------- loss / optimizer epilogue
@pytest.mark.parametrize("kind", ["SE", "AE"])
@pytest.mark.parametrize("with_bkg", [False, True])
def test_image_loss_gradient_fd(kind, with_bkg):
    """The image loss gradient (rgb and, with a background, the optical depth
    channel) is the central difference of the loss; errors are kept away from
    the AE kink so the difference is exact to O(h^2)."""
    rng = np.random.Generator(np.random.PCG64(21))
    img = rng.uniform(0.1, 1.0, size=(1, 2, 3, 4, 4))
    tgt = img[..., :3] + rng.choice([-1.0, 1.0], size=img[..., :3].shape) * rng.uniform(0.05, 0.3,
                                                                                          size=img[..., :3].shape)
    bkg = rng.uniform(0.0, 1.0, size=tgt.shape) if with_bkg else None
    if with_bkg:
        tgt = tgt + bkg * np.exp(-img[..., 3:4])
    loss, g = orc.image_loss_and_grad(img, tgt, kind, bkg, scale=0.37)
    h = 1e-6
    for idx in [(0, 0, 0, 0, 0), (0, 1, 2, 3, 2), (0, 1, 1, 2, 3), (0, 0, 2, 1, 3)]:
        a, b = img.copy(), img.copy()
        a[idx] += h
        b[idx] -= h
        fd = (orc.image_loss_and_grad(a, tgt, kind, bkg, 0.37)[0] - orc.image_loss_and_grad(b, tgt, kind, bkg,
                                                                                           0.37)[0]) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, g[idx])
    if not with_bkg:
        assert np.all(g[..., 3] == 0)


def test_divergence_loss_linear_field_and_fd():
    """Components linear along their own axis (slopes a, b, c per cell) have
    div = wx a + wy b + wz c in every cell; the gradient is the central
    difference of the loss."""
    D, H, W = 4, 5, 6
    ws = (0.5, 2.0, 1.5)
    a, b, c = 0.3, -0.7, 0.2
    x = np.broadcast_to(a * np.arange(W + 1, dtype=np.float64)[None, None, None, :, None], (1, D, H, W + 1, 1))
    y = np.broadcast_to(b * np.arange(H + 1, dtype=np.float64)[None, None, :, None, None], (1, D, H + 1, W, 1))
    z = np.broadcast_to(c * np.arange(D + 1, dtype=np.float64)[None, :, None, None, None], (1, D + 1, H, W, 1))
    loss, div, _ = orc.divergence_loss_and_grad(x, y, z, ws, scale=0.25)
    want = ws[0] * a + ws[1] * b + ws[2] * c
    assert np.allclose(div, want, atol=1e-12)
    assert abs(loss - 0.25 * D * H * W * want * want) <= 1e-12
    rng = np.random.Generator(np.random.PCG64(5))
    comps = [rng.uniform(-1, 1, size=s) for s in ((1, D, H, W + 1, 1), (1, D, H + 1, W, 1), (1, D + 1, H, W, 1))]
    loss, _, grads = orc.divergence_loss_and_grad(*comps, ws, scale=0.25)
    h = 1e-6
    for ci, idx in [(0, (0, 1, 2, 0, 0)), (0, (0, 3, 4, 6, 0)), (1, (0, 2, 5, 3, 0)), (2, (0, 4, 0, 5, 0)),
                    (2, (0, 0, 3, 2, 0))]:
        up = [cc.copy() for cc in comps]
        dn = [cc.copy() for cc in comps]
        up[ci][idx] += h
        dn[ci][idx] -= h
        fd = (orc.divergence_loss_and_grad(*up, ws, 0.25)[0] - orc.divergence_loss_and_grad(*dn, ws, 0.25)[0]) / (2 * h)
        assert abs(fd - grads[ci][idx]) <= 1e-6 * max(1.0, abs(fd))


def test_adam_tf_closed_forms():
    """First step: p - lr g / (|g| + eps / sqrt(1 - b2)). A constant gradient
    moves every step by lr g / (|g| + eps / sqrt(1 - b2^t)) (m_t = (1-b1^t) g,
    v_t = (1-b2^t) g^2)."""
    lr, b1, b2, eps = 0.01, 0.9, 0.999, 1e-3
    p0 = np.array([1.0, -2.0, 0.5, 0.0])
    g = np.array([0.5, -1e-3, 2.0, 0.0])
    p, m, v = p0, np.zeros(4), np.zeros(4)
    for t in range(1, 6):
        prev = p
        p, m, v = orc.adam_tf(p, g, m, v, lr, b1, b2, eps, t)
        want = prev - lr * g / (np.abs(g) + eps / np.sqrt(1.0 - b2 ** t))
        assert np.allclose(p, want, rtol=0, atol=1e-13), t
    assert p[3] == p0[3]


def test_bf16_round_matches_torch():
    """bfloat16 rounding (nearest even) equals torch's float32 -> bfloat16
    conversion, ties and subnormals included."""
    import torch
    rng = np.random.Generator(np.random.PCG64(31))
    x = np.concatenate([rng.uniform(-3, 3, 5000), rng.uniform(0, 1e-38, 200),
                        (np.arange(1, 2000) * 2.0 ** -8 + 1.0) - 2.0 ** -9, [0.0, -0.0, 1.0, 65504.0]])
    x = x.astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).float().numpy().astype(np.float64)
    assert np.array_equal(orc.bf16_round(x), want)


def test_blend_beer_lambert_closed_form_and_fd():
    """The non-fused BEER_LAMBERT blend: constant sigma c over S cells gives
    rgb = light c sum_{i=1..S} e^{-ic} and depth S c; its sigma gradient is the
    central difference of <rgb, w>."""
    S = 7
    sig = np.full((1, S, 2, 3), 0.3)
    img, _ = orc.blend_beer_lambert_fwd_bwd(sig, (0.5, 1.0, 2.0))
    want = 0.3 * np.exp(-0.3 * np.arange(1, S + 1)).sum()
    assert np.allclose(img[..., 0], 0.5 * want, atol=1e-14) and np.allclose(img[..., 3], S * 0.3, atol=1e-14)
    rng = np.random.Generator(np.random.PCG64(8))
    sig = rng.uniform(0, 0.5, size=(2, 5, 3, 4))
    w = rng.uniform(-1, 1, size=(2, 3, 4, 3))
    light = (0.7, 0.2, 1.1)
    _, g = orc.blend_beer_lambert_fwd_bwd(sig, light, w)

    def loss(x):
        return float((orc.blend_beer_lambert_fwd_bwd(x, light)[0][..., :3] * w).sum())
    for idx in [(0, 0, 0, 0), (1, 4, 2, 3), (0, 2, 1, 2), (1, 1, 0, 1)]:
        a, b = sig.copy(), sig.copy()
        a[idx] += 1e-6
        b[idx] -= 1e-6
        assert abs((loss(a) - loss(b)) / 2e-6 - g[idx]) <= 1e-8


def test_mip_render_gradient_is_the_loss_derivative():
    """The batched step's composition (mip_render_fwd_bwd: LINDEPTH resample
    with LINEAR mips, BEER_LAMBERT blend, L2 loss, exact adjoint): its density
    gradient equals the FP64 central difference of its own loss, on a scene
    whose far samples read mip levels 1-3 (the mips, the level choice and the
    LoD do not depend on the density, so the loss is smooth in it)."""
    from helpers import scene
    sc = scene(grid=(16, 18, 20), image=(7, 6), views=2, seed=3)
    r = (sc["M"], sc["V"], sc["P"], sc["F"], sc["out_dhw"], sc["step"])
    pos = orc.sample_positions(sc["M"], sc["V"], sc["P"], sc["F"], (16, 18, 20), sc["out_dhw"], "TRANSFORM_LINDEPTH",
                               lod=True)
    assert pos[..., 3].max() > 1.0
    d = sc["data"].astype(np.float64) * 3.0
    rng = np.random.Generator(np.random.PCG64(4))
    tgt = rng.uniform(0, 0.4, size=(2, 7, 6, 3))
    light = (0.9, 0.6, 1.2)
    ref = orc.mip_render_fwd_bwd(d, tgt, *r, light=light, num_mipmaps=3)
    g = ref["gd"]
    flat = np.argsort(np.abs(g).ravel())[::-1][:10]
    h = 1e-4
    for f in flat:
        idx = np.unravel_index(f, g.shape)
        a, b = d.copy(), d.copy()
        a[idx] += h
        b[idx] -= h
        fd = (orc.mip_render_fwd_bwd(a, tgt, *r, light=light)["loss"] -
              orc.mip_render_fwd_bwd(b, tgt, *r, light=light)["loss"]) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-6 * np.abs(g).max() + 1e-12, idx
