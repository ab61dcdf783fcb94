"""Python face of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the reference results or as the timed CPU baseline; the product path
never does. Wraps liboracle.so (oracle.cpp, built by oracle/Makefile) with
numpy arrays and restates, in numpy, the Python-level compositions of the
reference:

* MacCormack advection with the MC_SMOOTH revert clamp
  (phitest/render/data_structures.py:870-905) and its adjoint, both the
  reference's (TF autodiff through count-normalised SampleGridLut gradients,
  cuda/ops_loader.py:35-51) and the exact one;
* the density render of Renderer.render_density with an ambient light
  (renderer.py:603-651, lighting.py:23-26): light = c*i*d, data scaled by the
  camera depth step (renderer.py:420-437), fused raymarch, image = rgb;
* the scalar L2 image loss and the chain rule back to density and velocity.

``prec`` is 1 for FP64 (the oracle proper) and 0 for FP32 (the reference's
arithmetic).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FILTER = {"NEAREST": 0, "LINEAR": 1, "MIN": 2, "MAX": 3}
BOUNDARY = {"BORDER": 0, "CLAMP": 1, "WRAP": 2}
BLEND = {"BEER_LAMBERT": 0, "ALPHA": 1, "ALPHA_ADDITIVE": 2, "ADDITIVE": 3}
MIP = {"NONE": 0, "NEAREST": 1, "LINEAR": 2}
COORD = {"TRANSFORM_LINDEPTH": 0, "TRANSFORM_LINDEPTH_REVERSE": 1, "TRANSFORM": 2, "TRANSFORM_REVERSE": 3}
GRAD = {"REFERENCE": 0, "EXACT": 1}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        _LIB = ctypes.CDLL(path)
    return _LIB


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _p(a):
    # the numpy ctypes object keeps the array alive for the duration of the call
    return a.ctypes


def _mats(M, V, P, F):
    M = _f(M).reshape(-1, 16)
    V = _f(V).reshape(-1, 16)
    P = _f(P).reshape(-1, 16)
    F = _f(F).reshape(-1, 6)
    return M, V, P, F


def _check(rc, name):
    if rc != 0:
        raise RuntimeError("oracle %s failed (%d)" % (name, rc))


# ---------------------------------------------------------------- raymarch
def raymarch_fwd(data, M, V, P, F, out_dhw, interpolation="LINEAR", boundary="BORDER", blending_mode="BEER_LAMBERT",
                 separate_camera_batch=True, prec=1):
    """raymarch_grid_transform (keep_dims=False): returns [N', V', H, W, C]."""
    data = _d(data)
    N, C = data.shape[0], data.shape[4]
    M, V, P, F = _mats(M, V, P, F)
    nCam = V.shape[0]
    out_dhw = _i64(out_dhw)
    nv = nCam if separate_camera_batch else 1
    out = np.zeros((N, nv, int(out_dhw[1]), int(out_dhw[2]), C), dtype=np.float64)
    _check(lib().orc_raymarch_fwd(ctypes.c_int(prec), _p(data), _p(_i64(data.shape)), _p(M), _p(V), _p(P), _p(F),
                                  ctypes.c_int(nCam), _p(out_dhw), ctypes.c_int(FILTER[interpolation]),
                                  ctypes.c_int(BOUNDARY[boundary]), ctypes.c_int(BLEND[blending_mode]),
                                  ctypes.c_int(int(separate_camera_batch)), _p(out)), "raymarch_fwd")
    return out


def raymarch_bwd(data, output, output_grad, M, V, P, F, out_dhw, boundary="BORDER", blending_mode="BEER_LAMBERT",
                 separate_camera_batch=True, grad_mode="EXACT", prec=1):
    data = _d(data)
    M, V, P, F = _mats(M, V, P, F)
    grad = np.zeros_like(data)
    _check(lib().orc_raymarch_bwd(ctypes.c_int(prec), ctypes.c_int(GRAD[grad_mode]), _p(data), _p(_i64(data.shape)),
                                  _p(_d(output)), _p(_d(output_grad)), _p(M), _p(V), _p(P), _p(F),
                                  ctypes.c_int(V.shape[0]), _p(_i64(out_dhw)), ctypes.c_int(BOUNDARY[boundary]),
                                  ctypes.c_int(BLEND[blending_mode]), ctypes.c_int(int(separate_camera_batch)),
                                  _p(grad)), "raymarch_bwd")
    return grad


# ------------------------------------------------------------ reduce blend
def reduce_blend_fwd(data, blending_mode="BEER_LAMBERT", keep_dims=False, prec=1):
    data = _d(data)
    N, D, H, W, C = data.shape
    out = np.zeros(data.shape if keep_dims else (N, H, W, C), dtype=np.float64)
    _check(lib().orc_reduce_blend_fwd(ctypes.c_int(prec), _p(data), _p(_i64(data.shape)),
                                      ctypes.c_int(BLEND[blending_mode]), ctypes.c_int(int(keep_dims)), _p(out)),
           "reduce_blend_fwd")
    return out


def reduce_blend_bwd(data, output, output_grad, blending_mode="BEER_LAMBERT", keep_dims=False, grad_mode="REFERENCE",
                     prec=1):
    data = _d(data)
    grad = np.zeros_like(data)
    _check(lib().orc_reduce_blend_bwd(ctypes.c_int(prec), ctypes.c_int(GRAD[grad_mode]), _p(data),
                                      _p(_i64(data.shape)), _p(_d(output)), _p(_d(output_grad)),
                                      ctypes.c_int(BLEND[blending_mode]), ctypes.c_int(int(keep_dims)), _p(grad)),
           "reduce_blend_bwd")
    return grad


# -------------------------------------------------------- sample transform
def sample_transform_fwd(data, M, V, P, F, out_dhw, interpolation="LINEAR", boundary="BORDER", mipmapping="NONE",
                         num_mipmaps=0, mip_bias=0.0, coordinate_mode="TRANSFORM_LINDEPTH",
                         separate_camera_batch=True, prec=1):
    data = _d(data)
    N, C = data.shape[0], data.shape[4]
    M, V, P, F = _mats(M, V, P, F)
    nCam = V.shape[0]
    nv = nCam if separate_camera_batch else 1
    out_dhw = _i64(out_dhw)
    out = np.zeros((N, nv) + tuple(int(s) for s in out_dhw) + (C,), dtype=np.float64)
    _check(lib().orc_sample_transform_fwd(
        ctypes.c_int(prec), _p(data), _p(_i64(data.shape)), _p(M), _p(V), _p(P), _p(F), ctypes.c_int(nCam),
        _p(out_dhw), ctypes.c_int(FILTER[interpolation]), ctypes.c_int(BOUNDARY[boundary]),
        ctypes.c_int(MIP[mipmapping]), ctypes.c_int(num_mipmaps), ctypes.c_double(mip_bias),
        ctypes.c_int(COORD[coordinate_mode]), ctypes.c_int(int(separate_camera_batch)), _p(out)),
        "sample_transform_fwd")
    return out


def sample_transform_bwd(data, output_grad, M, V, P, F, interpolation="LINEAR", boundary="BORDER",
                         mipmapping="NONE", num_mipmaps=0, mip_bias=0.0, coordinate_mode="TRANSFORM_LINDEPTH",
                         separate_camera_batch=True, grad_mode="REFERENCE", prec=1):
    data = _d(data)
    og = _d(output_grad)
    M, V, P, F = _mats(M, V, P, F)
    grad = np.zeros_like(data)
    out_dhw = _i64(og.shape[2:5])
    _check(lib().orc_sample_transform_bwd(
        ctypes.c_int(prec), ctypes.c_int(GRAD[grad_mode]), _p(data), _p(_i64(data.shape)), _p(og), _p(M), _p(V),
        _p(P), _p(F), ctypes.c_int(V.shape[0]), _p(out_dhw), ctypes.c_int(FILTER[interpolation]),
        ctypes.c_int(BOUNDARY[boundary]), ctypes.c_int(MIP[mipmapping]), ctypes.c_int(num_mipmaps),
        ctypes.c_double(mip_bias), ctypes.c_int(COORD[coordinate_mode]), ctypes.c_int(int(separate_camera_batch)),
        _p(grad)), "sample_transform_bwd")
    return grad


# --------------------------------------------------------------------- LuT
def sample_lut_fwd(data, lut, interpolation="LINEAR", boundary="BORDER", mipmapping="NONE", num_mipmaps=0,
                   mip_bias=0.0, cell_center_offset=0.0, separate_camera_batch=True, relative_coords=True,
                   normalized_coords=False, prec=1):
    data = _d(data)
    lut = _d(lut)
    N, C = data.shape[0], data.shape[4]
    nv = lut.shape[0] if separate_camera_batch else 1
    out = np.zeros((N, nv) + lut.shape[1:4] + (C,), dtype=np.float64)
    _check(lib().orc_sample_lut_fwd(
        ctypes.c_int(prec), _p(data), _p(_i64(data.shape)), _p(lut), _p(_i64(lut.shape)),
        ctypes.c_int(FILTER[interpolation]), ctypes.c_int(BOUNDARY[boundary]), ctypes.c_int(MIP[mipmapping]),
        ctypes.c_int(num_mipmaps), ctypes.c_double(mip_bias), ctypes.c_double(cell_center_offset),
        ctypes.c_int(int(separate_camera_batch)), ctypes.c_int(int(relative_coords)),
        ctypes.c_int(int(normalized_coords)), _p(out)), "sample_lut_fwd")
    return out


def sample_lut_bwd(data, lut, output_grad, interpolation="LINEAR", boundary="BORDER", cell_center_offset=0.0,
                   separate_camera_batch=True, relative_coords=True, normalized_coords=False, grad_mode="REFERENCE",
                   prec=1):
    data = _d(data)
    lut = _d(lut)
    gin = np.zeros_like(data)
    glut = np.zeros_like(lut)
    _check(lib().orc_sample_lut_bwd(
        ctypes.c_int(prec), ctypes.c_int(GRAD[grad_mode]), _p(data), _p(_i64(data.shape)), _p(lut),
        _p(_i64(lut.shape)), _p(_d(output_grad)), ctypes.c_int(FILTER[interpolation]),
        ctypes.c_int(BOUNDARY[boundary]), ctypes.c_double(cell_center_offset),
        ctypes.c_int(int(separate_camera_batch)), ctypes.c_int(int(relative_coords)),
        ctypes.c_int(int(normalized_coords)), _p(gin), _p(glut)), "sample_lut_bwd")
    return gin, glut


def lod_transform(in_dhw, MV, P, F, out_dhw, inverse_transform=False, prec=1):
    MV = _f(MV).reshape(-1, 16)
    P = _f(P).reshape(-1, 16)
    F = _f(F).reshape(-1, 6)
    out = np.zeros((MV.shape[0],) + tuple(int(s) for s in out_dhw) + (4,), dtype=np.float64)
    _check(lib().orc_lod_transform(ctypes.c_int(prec), _p(_i64(in_dhw)), _p(MV), _p(P), _p(F),
                                   ctypes.c_int(MV.shape[0]), _p(_i64(out_dhw)), ctypes.c_int(int(inverse_transform)),
                                   _p(out)), "lod_transform")
    return out


def sample_positions(M, V, P, F, in_dhw, out_dhw, coordinate_mode="TRANSFORM_LINDEPTH", lod=False, prec=1):
    """Per output cell centre (index + 0.5) of every camera: the position the
    coordinate mode maps it to (object space for the forward modes, input
    index space for the reverse ones) and its LoD (samplePos3D,
    sampling.hpp:59-136). Returns [nCam, D, H, W, 4] = (x, y, z, lod)."""
    M = _f(M).reshape(-1, 16)[:1]
    V = _f(V).reshape(-1, 16)
    P = _f(P).reshape(-1, 16)
    F = _f(F).reshape(-1, 6)
    out = np.zeros((V.shape[0],) + tuple(int(s) for s in out_dhw) + (4,), dtype=np.float64)
    _check(lib().orc_sample_positions(ctypes.c_int(prec), _p(M), _p(V), _p(P), _p(F), ctypes.c_int(V.shape[0]),
                                      _p(_i64(in_dhw)), _p(_i64(out_dhw)), ctypes.c_int(COORD[coordinate_mode]),
                                      ctypes.c_int(int(lod)), _p(out)), "sample_positions")
    return out


def build_mips(data_dhwc, levels, prec=1):
    data = _d(data_dhwc)
    D, H, W, C = data.shape
    dims = []
    total = 0
    for m in range(1, levels + 1):
        s = (D >> m, H >> m, W >> m)
        dims.append(s)
        total += s[0] * s[1] * s[2] * C
    buf = np.zeros(total, dtype=np.float64)
    _check(lib().orc_build_mips(ctypes.c_int(prec), _p(data), _p(_i64(data.shape)), ctypes.c_int(levels), _p(buf)),
           "build_mips")
    out, o = [], 0
    for s in dims:
        n = s[0] * s[1] * s[2] * C
        out.append(buf[o:o + n].reshape(s + (C,)))
        o += n
    return out


# ------------------------------------------------------- advection (A10)
def _lut_from_velocity(vel_c, dt=1.0):
    """centered_lut_grid (data_structures.py:864-868) padded with a zero LoD
    channel (renderer.py:448-454): [1,D,H,W,4]."""
    v = _d(vel_c)
    lut = np.concatenate([-v * dt, np.zeros(v.shape[:-1] + (1,))], axis=-1)
    return lut


def advect_fwd(d, vel_c, order=2, clamp="MC_SMOOTH", dt=1.0, boundary="CLAMP", prec=1, return_parts=False):
    """VelocityGrid.warp (data_structures.py:870-905) on a [1,D,H,W,C] grid."""
    lut = _lut_from_velocity(vel_c, dt)
    kw = dict(boundary=boundary, separate_camera_batch=False, relative_coords=True, prec=prec)
    f = sample_lut_fwd(d, lut, **kw)[:, 0]
    parts = {"f": f}
    out = f
    if order == 2:
        nlut = lut.copy()
        nlut[..., :3] *= -1
        b = sample_lut_fwd(f, nlut, **kw)[:, 0]
        corr = f + 0.5 * (_d(d) - b)
        parts.update(b=b, corr=corr)
        if clamp in ("MC", "MC_SMOOTH"):
            mn = sample_lut_fwd(d, lut, interpolation="MIN", **kw)[:, 0]
            mx = sample_lut_fwd(d, lut, interpolation="MAX", **kw)[:, 0]
            mask = np.logical_or(corr < mn, corr > mx)
            parts.update(mn=mn, mx=mx, mask=mask)
            if clamp == "MC":
                raise NotImplementedError("MC clamp raises in the reference (data_structures.py:890-893)")
            corr = np.where(mask, f, corr)
        out = corr
    if clamp == "NEGATIVE":
        out = np.maximum(out, 0)
    return (out, parts) if return_parts else out


def advect_bwd(d, vel_c, out_grad, order=2, clamp="MC_SMOOTH", dt=1.0, boundary="CLAMP", grad_mode="EXACT",
               prec=1):
    """Gradients of advect_fwd w.r.t. the density and the centred velocity.

    REFERENCE follows TF autodiff through the composition with the
    reference's SampleGridLut gradient op; EXACT uses the exact LuT adjoint.
    MIN/MAX samples and the clamp mask carry no gradient (tf.less/greater)."""
    lut = _lut_from_velocity(vel_c, dt)
    kw = dict(boundary=boundary, separate_camera_batch=False, relative_coords=True, grad_mode=grad_mode, prec=prec)
    out, parts = advect_fwd(d, vel_c, order, clamp, dt, boundary, prec, return_parts=True)
    g = _d(out_grad)
    if clamp == "NEGATIVE":
        g = g * (out > 0)
    gd = np.zeros_like(_d(d))
    glut = np.zeros_like(lut)
    gf = np.zeros_like(g)
    if order == 2:
        if "mask" in parts:
            gf += np.where(parts["mask"], g, 0.0)
            gc = np.where(parts["mask"], 0.0, g)
        else:
            gc = g
        gf += gc
        gd += 0.5 * gc
        gb = -0.5 * gc
        nlut = lut.copy()
        nlut[..., :3] *= -1
        gin, gnlut = sample_lut_bwd(parts["f"], nlut, gb[:, None], **kw)
        gf += gin
        glut[..., :3] -= gnlut[..., :3]
    else:
        gf = g
    gin, gl = sample_lut_bwd(d, lut, gf[:, None], **kw)
    gd += gin
    glut[..., :3] += gl[..., :3]
    gvel = -glut[..., :3] * dt
    return gd, gvel


# -------------------------------------------------- density render (A11)
def render_density_fwd(d, M, V, P, F, out_dhw, depth_step, light_color=(1.0, 1.0, 1.0), light_intensity=1.0,
                       boundary="BORDER", prec=1):
    """Ambient-light density render: grid = [c*i*d, d] * depth_step, fused
    BEER_LAMBERT raymarch. Returns [1,V,H,W,4] (rgb, optical depth)."""
    d = _d(d)
    light = d * (np.asarray(light_color, dtype=np.float64) * light_intensity)
    grid = np.concatenate([light, d], axis=-1) * depth_step
    return raymarch_fwd(grid, M, V, P, F, out_dhw, "LINEAR", boundary, "BEER_LAMBERT", True, prec)


def render_density_bwd(d, out, out_grad, M, V, P, F, out_dhw, depth_step, light_color=(1.0, 1.0, 1.0),
                       light_intensity=1.0, boundary="BORDER", grad_mode="EXACT", prec=1):
    d = _d(d)
    cvec = np.asarray(light_color, dtype=np.float64) * light_intensity
    grid = np.concatenate([d * cvec, d], axis=-1) * depth_step
    g4 = raymarch_bwd(grid, out, out_grad, M, V, P, F, out_dhw, boundary, "BEER_LAMBERT", True, grad_mode, prec)
    return depth_step * (g4[..., :3] @ cvec + g4[..., 3])[..., None]


def render_lit_fwd(d, lam, M, V, P, F, out_dhw, depth_step, color=(1.0, 1.0, 1.0), ambient=0.0,
                   boundary="BORDER", prec=1):
    """Density render under a spot/point light and an ambient light whose
    colours are proportional (Renderer.render_density over _build_light_grid,
    renderer.py:573-651; SpotLight.grid_lighting lighting.py:143-147): light
    grid = color * d * (lam + ambient) with lam the light's per-cell factor,
    grid = [light, d] * depth_step, BEER_LAMBERT raymarch."""
    d, lam = _d(d), _d(lam)
    c = np.asarray(color, dtype=np.float64)
    grid = np.concatenate([d * (lam + ambient) * c, d], axis=-1) * depth_step
    return raymarch_fwd(grid, M, V, P, F, out_dhw, "LINEAR", boundary, "BEER_LAMBERT", True, prec)


def render_lit_bwd(d, lam, out, out_grad, M, V, P, F, out_dhw, depth_step, color=(1.0, 1.0, 1.0), ambient=0.0,
                   boundary="BORDER", grad_mode="EXACT", prec=1):
    """Gradients of render_lit_fwd w.r.t. the density and the light factor:
    the raymarch gradient g4 of the 4-channel grid chained through
    light = color * d * (lam + ambient) * step and sigma = d * step."""
    d, lam = _d(d), _d(lam)
    c = np.asarray(color, dtype=np.float64)
    grid = np.concatenate([d * (lam + ambient) * c, d], axis=-1) * depth_step
    g4 = raymarch_bwd(grid, out, out_grad, M, V, P, F, out_dhw, boundary, "BEER_LAMBERT", True, grad_mode, prec)
    gl = (g4[..., :3] @ c)[..., None] * depth_step
    return gl * (lam + ambient) + depth_step * g4[..., 3:4], gl * d


def l2_loss_and_grad(img, target):
    """loss = mean((rgb - target)^2); returns (loss, d loss / d img) with the
    optical-depth channel receiving zero gradient."""
    rgb = img[..., :3]
    diff = rgb - target
    loss = float(np.mean(diff * diff))
    g = np.zeros_like(img)
    g[..., :3] = 2.0 * diff / diff.size
    return loss, g


def image_loss_and_grad(img, target, kind="SE", background=None, scale=1.0):
    """Image loss of loss_dens_target / loss_dens_target_raw
    (optimization.py:396-430) with the SE or AE loss function
    (optimization.py:94-130): loss = scale * sum_c f(e_c), e = rgb + bkg *
    exp(-D) - target (no bkg term without a background), f = e^2 or |e|
    (d|e|/de = sign(e), 0 at 0). Returns (loss, d loss / d img [..., 4])."""
    img = np.asarray(img, dtype=np.float64)
    rgb, D = img[..., :3], img[..., 3:4]
    e = rgb - np.asarray(target, dtype=np.float64)
    tr = None
    if background is not None:
        tr = np.exp(-D)
        e = e + np.asarray(background, dtype=np.float64) * tr
    if kind == "AE":
        loss, fp = float(np.sum(np.abs(e))), np.sign(e)
    else:
        loss, fp = float(np.sum(e * e)), 2.0 * e
    g = np.zeros_like(img)
    g[..., :3] = scale * fp
    if background is not None:
        g[..., 3] = -scale * np.sum(np.asarray(background, dtype=np.float64) * fp, axis=-1) * tr[..., 0]
    return scale * loss, g


def divergence_loss_and_grad(x, y, z, world_scale=(1.0, 1.0, 1.0), scale=1.0):
    """VelocityGrid.divergence (data_structures.py:929-936) of staggered
    components x [1,D,H,W+1,1], y [1,D,H+1,W,1], z [1,D+1,H,W,1] and the SE
    divergence loss scale * sum(div^2) (optimization.py:824-838). The
    gradient is the adjoint of the forward differences: a face receives
    2 scale w (div of the cell before it - div of the cell after it)."""
    x, y, z = (np.asarray(a, dtype=np.float64) for a in (x, y, z))
    div = ((x[:, :, :, 1:] - x[:, :, :, :-1]) * world_scale[0] + (y[:, :, 1:] - y[:, :, :-1]) * world_scale[1]
           + (z[:, 1:] - z[:, :-1]) * world_scale[2])
    loss = scale * float(np.sum(div * div))
    gx, gy, gz = np.zeros_like(x), np.zeros_like(y), np.zeros_like(z)
    t = 2.0 * scale * div
    gx[:, :, :, 1:] += t * world_scale[0]
    gx[:, :, :, :-1] -= t * world_scale[0]
    gy[:, :, 1:] += t * world_scale[1]
    gy[:, :, :-1] -= t * world_scale[1]
    gz[:, 1:] += t * world_scale[2]
    gz[:, :-1] -= t * world_scale[2]
    return loss, div, (gx, gy, gz)


def adam_tf(p, g, m, v, lr, beta1, beta2, epsilon, t):
    """One update as tf.train.AdamOptimizer documents it (the reference's
    optimizer, reconstruct_sequence.py:1899-1905), in float64:
    lr_t = lr sqrt(1 - b2^t) / (1 - b1^t); m = b1 m + (1 - b1) g;
    v = b2 v + (1 - b2) g^2; p = p - lr_t m / (sqrt(v) + eps)."""
    p, g, m, v = (np.asarray(a, dtype=np.float64) for a in (p, g, m, v))
    lr_t = lr * np.sqrt(1.0 - beta2 ** t) / (1.0 - beta1 ** t)
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    return p - lr_t * m / (np.sqrt(v) + epsilon), m, v


def bf16_round(a):
    """Values rounded to bfloat16 (nearest even), as float64: the float32
    value's upper 16 bits after adding half an ulp of the bfloat16 grid."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def step_fwd_bwd(d0, vel_c, target, M, V, P, F, out_dhw, depth_step, order=2, clamp="MC_SMOOTH", prec=1,
                 grad_mode="EXACT", boundary_render="BORDER", loss_kind=None, background=None, loss_scale=None,
                 render_storage="float32"):
    """One advect -> render -> image loss step and its gradients w.r.t. d0
    and the centred velocity. Default: the mean squared error
    (l2_loss_and_grad); with loss_kind / background / loss_scale the image
    loss of image_loss_and_grad. render_storage "bfloat16" renders the
    advected density rounded to bfloat16 (the stored precision), its gradient
    passed through the rounding. Returns dict(loss, img, d1, d1r, gd, gv)."""
    d1 = advect_fwd(d0, vel_c, order, clamp, 1.0, "CLAMP", prec)
    d1r = bf16_round(d1) if render_storage == "bfloat16" else d1
    img = render_density_fwd(d1r, M, V, P, F, out_dhw, depth_step, boundary=boundary_render, prec=prec)
    if loss_kind is None and background is None and loss_scale is None:
        loss, gimg = l2_loss_and_grad(img, target)
    else:
        scale = loss_scale if loss_scale is not None else 1.0 / np.asarray(target).size
        loss, gimg = image_loss_and_grad(img, target, loss_kind or "SE", background, scale)
    gd1 = render_density_bwd(d1r, img, gimg, M, V, P, F, out_dhw, depth_step, boundary=boundary_render,
                             grad_mode=grad_mode, prec=prec)
    gd, gv = advect_bwd(d0, vel_c, gd1, order, clamp, 1.0, "CLAMP", grad_mode, prec)
    return dict(loss=loss, img=img, d1=d1, d1r=d1r, gd=gd, gv=gv, gd1=gd1)


# --------------------------------------- mipmapped non-fused render (c4)
def blend_beer_lambert_fwd_bwd(sigma, light, img_grad_rgb=None):
    """_blend_grid BEER_LAMBERT on a density frustum [V,S,H,W] whose light
    channels are light * sigma (renderer.py:488-499): D_i = sum_{j<=i}
    sigma_j, rgb = light * sum_i sigma_i exp(-D_i), depth = D_{S-1}. With an
    rgb gradient [V,H,W,3], also d loss / d sigma: with G = light . g_rgb,
    G (u_k - sum_{i>=k} u_i sigma_i), u = exp(-D)."""
    sig = _d(sigma)
    D = np.cumsum(sig, axis=1)
    u = np.exp(-D)
    E = (u * sig).sum(1)
    light = np.asarray(light, dtype=np.float64)
    img = np.concatenate([E[..., None] * light, D[:, -1][..., None]], axis=-1)
    if img_grad_rgb is None:
        return img, None
    G = _d(img_grad_rgb) @ light
    R = np.flip(np.cumsum(np.flip(u * sig, 1), 1), 1)
    return img, G[:, None] * (u - R)


def mip_render_fwd_bwd(d, target, M, V, P, F, out_dhw, depth_step, light=(1.0, 1.0, 1.0), num_mipmaps=3,
                       loss_scale=None, prec=1):
    """The density-reconstruction step of the batched render for one
    reconstruction: the non-fused ambient render (Renderer.render_density with
    fused=False, renderer.py:555-568,603-651), i.e. SampleGridTransform
    TRANSFORM_LINDEPTH of the step-scaled density with LINEAR mips
    (num_mipmaps levels, LoD per sample), the BEER_LAMBERT blend, loss =
    loss_scale * sum (rgb - target)^2 (default: the mean), and the exact
    density gradient through blend, resample and mips. target [V,H,W,3].
    Returns dict(img [V,H,W,4], loss, gd [1,D,H,W,1])."""
    dd = _d(d) * depth_step
    mm = "LINEAR" if num_mipmaps > 0 else "NONE"
    fr = sample_transform_fwd(dd, M, V, P, F, out_dhw, "LINEAR", "BORDER", mm, num_mipmaps, 0.0,
                              "TRANSFORM_LINDEPTH", prec=prec)
    sig = fr[0, ..., 0]
    img, _ = blend_beer_lambert_fwd_bwd(sig, light)
    tg = _d(target)
    scale = 1.0 / tg.size if loss_scale is None else loss_scale
    e = img[..., :3] - tg
    loss = float(scale * (e * e).sum())
    _, gsig = blend_beer_lambert_fwd_bwd(sig, light, 2.0 * scale * e)
    gd = sample_transform_bwd(dd, gsig[None, ..., None], M, V, P, F, "LINEAR", "BORDER", mm, num_mipmaps, 0.0,
                              "TRANSFORM_LINDEPTH", grad_mode="EXACT", prec=prec) * depth_step
    return dict(img=img, loss=loss, gd=gd)


# ------------------------------------------- renderer gather gradient (A4)
def sample_transform_gather_grad(output_grad, M, V, P, F, in_dhw, interpolation="LINEAR", boundary="BORDER",
                                 mipmapping="NONE", num_mipmaps=0, mip_bias=0.0,
                                 coordinate_mode="TRANSFORM_LINDEPTH", prec=1):
    """Custom gradient of Renderer._sample_camera_transform / _sample_transform
    (renderer.py:200-227,246-272): per data batch i, resample every view's
    gradient grid dy[i, j] back with camera j and the reverse coordinate mode
    (mips built on dy) and sum over the views."""
    rev = {"TRANSFORM_LINDEPTH": "TRANSFORM_LINDEPTH_REVERSE", "TRANSFORM_LINDEPTH_REVERSE": "TRANSFORM_LINDEPTH",
           "TRANSFORM": "TRANSFORM_REVERSE", "TRANSFORM_REVERSE": "TRANSFORM"}[coordinate_mode]
    og = _d(output_grad)
    N, Vc = og.shape[:2]
    M = _f(M).reshape(-1, 16)
    out = []
    for i in range(N):
        dxv = sample_transform_fwd(og[i], np.repeat(M[i:i + 1], Vc, axis=0), V, P, F, in_dhw, interpolation, boundary,
                                   mipmapping, num_mipmaps, mip_bias, rev, separate_camera_batch=False, prec=prec)
        out.append(dxv.sum(axis=(0, 1)))
    return np.stack(out)


# ----------------------------------------------- staggered velocity (A10)
_FACE_OFFSETS = {  # (component sampled, target component) -> offset xyz (data_structures.py:835-862)
    ("X", "Y"): (0.5, -0.5, 0.0), ("X", "Z"): (0.5, 0.0, -0.5),
    ("Y", "X"): (-0.5, 0.5, 0.0), ("Y", "Z"): (0.0, 0.5, -0.5),
    ("Z", "X"): (-0.5, 0.0, 0.5), ("Z", "Y"): (0.0, -0.5, 0.5),
}


def _offset_lut(shape_dhw, offset):
    lut = np.zeros((1,) + tuple(shape_dhw) + (4,))
    lut[..., 0], lut[..., 1], lut[..., 2] = offset
    return lut


def face_velocity(comps, target, prec=1):
    """Velocity at the faces of component ``target``: its own values and the
    other two resampled with half-cell offsets (LINEAR, CLAMP, relative LuT)."""
    names = ("X", "Y", "Z")
    tshape = comps[names.index(target)].shape[1:4]
    out = []
    for n, c in zip(names, comps):
        if n == target:
            out.append(_d(c))
        else:
            out.append(sample_lut_fwd(c, _offset_lut(tshape, _FACE_OFFSETS[(n, target)]), boundary="CLAMP",
                                      separate_camera_batch=False, relative_coords=True, prec=prec)[:, 0])
    return np.concatenate(out, axis=-1)


def face_velocity_bwd(comps, target, g_faces, prec=1):
    """Adjoint of face_velocity (EXACT LuT input adjoints) -> grads per component."""
    names = ("X", "Y", "Z")
    tshape = comps[names.index(target)].shape[1:4]
    grads = []
    for i, (n, c) in enumerate(zip(names, comps)):
        g = _d(g_faces[..., i:i + 1])
        if n == target:
            grads.append(g)
        else:
            gin, _ = sample_lut_bwd(c, _offset_lut(tshape, _FACE_OFFSETS[(n, target)]), g[:, None], boundary="CLAMP",
                                    separate_camera_batch=False, relative_coords=True, grad_mode="EXACT", prec=prec)
            grads.append(gin)
    return grads


def velocity_self_advect(comps, order=2, clamp="MC_SMOOTH", dt=1.0, prec=1):
    """VelocityGrid.warped() with no external velocity (data_structures.py:594-651):
    each component warped by the velocity at its own faces; only the
    MacCormack clamps apply to components."""
    c = clamp if clamp in ("MC", "MC_SMOOTH") else "NONE"
    return [advect_fwd(comp, face_velocity(comps, n, prec), order, c, dt, "CLAMP", prec)
            for n, comp in zip(("X", "Y", "Z"), comps)]


def velocity_self_advect_bwd(comps, out_grads, order=2, clamp="MC_SMOOTH", dt=1.0, prec=1):
    """Exact gradient of sum_c <out_grads[c], warped_c> w.r.t. the components."""
    c = clamp if clamp in ("MC", "MC_SMOOTH") else "NONE"
    total = [np.zeros_like(_d(x)) for x in comps]
    for i, (n, comp) in enumerate(zip(("X", "Y", "Z"), comps)):
        vf = face_velocity(comps, n, prec)
        gd, gv = advect_bwd(comp, vf, out_grads[i], order, c, dt, "CLAMP", "EXACT", prec)
        total[i] += gd
        for j, g in enumerate(face_velocity_bwd(comps, n, gv, prec)):
            total[j] += g
    return total


# ------------------------------------------------- transport chain (A12)
def chain_fwd_bwd(d0, velocities, targets, M, V, P, F, out_dhw, depth_step, order=2, clamp="MC_SMOOTH",
                  grad_mode="EXACT", decay=0.0, light=(1.0, 1.0, 1.0), prec=1, vel_decay=None):
    """Density chain d_{f+1} = advect(d_f, v_f), loss = sum_f mean((render(d_f)
    - target_f)^2) and the backward of optStep_sequence in BWD order
    (optimization.py:592-602,974-1047), with g_f frame f's own render-loss
    gradient and A_f^T the advection adjoint.

    EXACT: lambda_f = g_f + A_f^T lambda_{f+1}; grad d0 = lambda_0 and the
    velocity gradient of frame f is the velocity part of A_f^T lambda_{f+1}.

    REFERENCE (the reference's warped-gradient accumulation, :1017-1031):
    the accumulator starts at lambda_{T-1} = lerp(g_{T-1}, 0, a) and is
    lambda_f = lerp(g_f, A_f^T lambda_{f+1}, a) = (1-a) g_f + a A_f^T lambda_{f+1}
    for a decay a > 0 and g_f + A_f^T lambda_{f+1} for a = 0. The gradient
    applied to frame f is its own g_f plus the warped term, which is divided by
    num_step = T-1-f when the decay is 0 (:1017-1018): grad d0 = g_0 +
    s_0 (A_0^T lambda_1)_d. The velocity gradient of frame f is
    s_f (A_f^T lambda_{f+1})_v, with its own decay (warp_vel_grads_decay,
    :1019-1020; vel_decay defaults to decay)."""
    T = len(targets)
    ds = [_d(d0)]
    for f in range(T - 1):
        ds.append(advect_fwd(ds[-1], velocities[f], order, clamp, 1.0, "CLAMP", prec))
    losses, g_own = [], []
    for f in range(T):
        img = render_density_fwd(ds[f], M, V, P, F, out_dhw, depth_step, light, prec=prec)
        loss, gimg = l2_loss_and_grad(img, _d(targets[f]))
        losses.append(loss)
        g_own.append(render_density_bwd(ds[f], img, gimg, M, V, P, F, out_dhw, depth_step, light,
                                        grad_mode="EXACT", prec=prec))
    ref = grad_mode == "REFERENCE"
    a = float(decay)
    av = a if vel_decay is None else float(vel_decay)

    def lerp(l, d):  # optimization.py:1031: lerp(l, d, a) if a > 0 else l + d
        return (1.0 - a) * l + a * d if (ref and a > 0) else l + d
    if T == 1:
        return dict(loss=float(np.sum(losses)), losses=losses, gd0=g_own[0], gv=[], densities=ds)
    lam = lerp(g_own[-1], 0.0 * g_own[-1])
    gv = [None] * (T - 1)
    gd0 = None
    for f in range(T - 2, -1, -1):
        num_step = T - 1 - f
        gd, gvf = advect_bwd(ds[f], velocities[f], lam, order, clamp, 1.0, "CLAMP", "EXACT", prec)
        sd = 1.0 / num_step if (ref and a == 0) else 1.0
        sv = 1.0 / num_step if (ref and av == 0) else 1.0
        gv[f] = gvf * sv
        if f == 0:
            gd0 = g_own[0] + sd * gd
        lam = lerp(g_own[f], gd)
    return dict(loss=float(np.sum(losses)), losses=losses, gd0=gd0, gv=gv, densities=ds)
