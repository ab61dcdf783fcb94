// Fused density render: emission-absorption raymarch of a scalar density lit
// by an ambient light, and its backward with the L2 image loss folded in.
//
// This is Renderer.render_density with an ambient Light
// (phitest/render/renderer.py:603-651, lighting.py:23-26) through the fused
// raymarch (raymarch_grid.cu.cc:107-143) without materialising the 4-channel
// [c*i*d, d] * depth_step grid: the density is sampled once per step and
//   sigma_i = s * d(p_i),  L_i = (c*i) * sigma_i,
//   D_i = max(D_{i-1} + sigma_i, 0),  rgb += exp(-D_i) * L_i   (BEER_LAMBERT)
// so rgb = (c*i) * E with one scalar accumulator E = sum exp(-D_i) sigma_i.
// Output pixels are float4 (r, g, b, optical depth), like the reference.
//
// Backward, from an image gradient or with the L2 loss
// loss = scale * sum (rgb - target)^2 folded in (each lane then forms g_rgb
// from its own pixel). The density gradient of each sample is
//   g_d = s * (exp(-D_i) * G - G * sum_{j>=i} exp(-D_j) sigma_j + g_D),
//   G = sum_c (c*i)_c g_rgb_c,
// and spreading g_d to the 8 trilinear corners is done in one of three ways:
//   BORDER : atomic-free transpose. The forward saves, per sample,
//            c_i = exp(-D_i) + E_{i-1} (so that g_d = alpha c_i + beta with
//            per-pixel alpha, beta) and per pixel the clipped sample range and
//            E; a pixel pass forms alpha, beta from the loss gradient; a
//            voxel-parallel gather then enumerates, for each view, exactly the
//            samples inside the voxel's trilinear support (box -> projected
//            pixel range -> ray/box slab) and sums g_d * w in a fixed order:
//            deterministic, one write per voxel.
//   CLAMP  : samples outside the grid land on boundary voxels, so the support
//            is unbounded; a tiled walk from far to near (recovering
//            D_{i-1} = max(D_i - sigma_i, 0)) accumulates the scatter in an LDS
//            brick per 16x16-pixel tile and K-sample segment and flushes it
//            with coalesced global atomics.
//   WRAP   : the same walk with per-sample global atomics.
// Samples of a BORDER ray outside the grid read 0 and are skipped (exact).
#include "gft_device.hpp"
#include "quad.hpp"

namespace gft {

constexpr int kRTileX = 16;
constexpr int kRTileY = 16;

// ----------------------------------------------------------------- helpers
__device__ __forceinline__ void pixel_ray(const RayView& v, const RayDims& rd, int x, int y, float3& nearp,
                                          float3& step) {
  const float dinvx = 1.f / (float)rd.out_w, dinvy = 1.f / (float)rd.out_h, dinvz = 1.f / (float)rd.out_d;
  const float zmax = (float)max(rd.out_d - 1, 1);
  const float izm = 1.f / zmax;
  nearp = idx_to_os(v, x + 0.5f, y + 0.5f, 0.5f, dinvx, dinvy, dinvz);
  float3 f = idx_to_os(v, x + 0.5f, y + 0.5f, zmax + 0.5f, dinvx, dinvy, dinvz);
  step = make_float3((f.x - nearp.x) * izm, (f.y - nearp.y) * izm, (f.z - nearp.z) * izm);
}

// one thread per stored entry (coalesced writes); padding entries of partial
// bricks are written as zeros and never read. BORDER: a corner outside the
// grid is 0. CLAMP: corners take the clamped voxel, so that an entry at the
// copy's edge holds the edge value in all its slots and a tap whose index is
// clamped into the copy reads the CLAMP value whatever its weights.
// With `fac` non-null the copy holds the product lin * fac per voxel (the
// spot-lit emission grid d * lambda of the lit render).
template <int BM, class Q = float4>
__global__ __launch_bounds__(256) void k_quad_copy(const float* __restrict__ lin, const float* __restrict__ fac,
                                                   Q* __restrict__ quad, Dims3 d, QuadDims q) {
  using B = QuadBrick<Q>;
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= quad_entries<Q>(q)) return;
  const unsigned b = (unsigned)(j / B::N), o = (unsigned)(j % B::N);
  const unsigned ix = (b % q.nbx) * B::X + (o & (B::X - 1));
  const unsigned iy = ((b / q.nbx) % q.nby) * B::Y + ((o >> B::Xl) & (B::Y - 1));
  const unsigned iz = (b / (q.nbx * q.nby)) * B::Z + (o >> (B::Xl + B::Yl));
  const int x = (int)ix - 1, y = (int)iy - 1, z = (int)iz - 1;
  auto at = [&](size_t f) -> float { return fac ? lin[f] * fac[f] : lin[f]; };
  auto v = [&](int xx, int yy) -> float {
    if (BM == B_CLAMP) return at(flat3(clampi(xx, 0, d.x - 1), clampi(yy, 0, d.y - 1), clampi(z, 0, d.z - 1), d));
    return (((unsigned)xx < (unsigned)d.x) & ((unsigned)yy < (unsigned)d.y) & ((unsigned)z < (unsigned)d.z))
               ? at(flat3(xx, yy, z, d))
               : 0.f;
  };
  quad_store(quad + j, v(x, y), v(x + 1, y), v(x, y + 1), v(x + 1, y + 1));
}

// Branch-free read of U consecutive samples i0, i0+dir, ... of one ray (dir =
// +1 or -1) from the corner-quad copy: all 2U loads are issued before any
// value is used, with the addresses clamped into the copy. Samples past
// `last` (in walk direction) read 0 and report inside = false; so do, with
// BORDER, samples outside [-0.5, dims-0.5). With CLAMP every other sample is
// inside and reads the clamped value from the CLAMP copy.
template <int U, int BM = B_BORDER, class Q = float4>
__device__ __forceinline__ void read_border_batch(const Q* __restrict__ quad, Dims3 d, QuadDims q, float3 n,
                                                  float3 st, int i0, int dir, int last, float val[U],
                                                  bool inside[U]) {
  unsigned a[U], a1[U];
  float w[U][3];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int i = i0 + dir * u;
    const float fi = (float)i;
    const float px = fmaf(fi, st.x, n.x) - 0.5f, py = fmaf(fi, st.y, n.y) - 0.5f, pz = fmaf(fi, st.z, n.z) - 0.5f;
    const float qx = px + 0.5f, qy = py + 0.5f, qz = pz + 0.5f;
    const bool in_range = dir > 0 ? i <= last : i >= last;
    inside[u] = BM == B_CLAMP ? in_range
                              : in_range & (0.f <= qx) & (qx < (float)d.x) & (0.f <= qy) & (qy < (float)d.y) &
                                    (0.f <= qz) & (qz < (float)d.z);
    const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    w[u][0] = px - fx;
    w[u][1] = py - fy;
    w[u][2] = pz - fz;
    const unsigned ix = (unsigned)(clampi((int)fx, -1, d.x - 1) + 1), iy = (unsigned)(clampi((int)fy, -1, d.y - 1) + 1),
                   iz = (unsigned)(clampi((int)fz, -1, d.z - 1) + 1);
    a[u] = quad_index<Q>(q, ix, iy, iz);
    a1[u] = quad_index<Q>(q, ix, iy, iz + 1);
  }
  Q r0[U], r1[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    r0[u] = quad[a[u]];
    r1[u] = quad[a1[u]];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const float4 c0 = quad_value(r0[u]), c1 = quad_value(r1[u]);
    const float v0 = lerpf(lerpf(c0.x, c0.y, w[u][0]), lerpf(c0.z, c0.w, w[u][0]), w[u][1]);
    const float v1 = lerpf(lerpf(c1.x, c1.y, w[u][0]), lerpf(c1.z, c1.w, w[u][0]), w[u][1]);
    val[u] = inside[u] ? lerpf(v0, v1, w[u][2]) : 0.f;
  }
}

// As read_border_batch for two corner-quad copies of the same grid shape (the
// density and the lit emission grid d * lambda), read with the same corners and
// weights; all 4U loads are in flight before any value is used.
template <int U>
__device__ __forceinline__ void read_border_batch2(const float4* __restrict__ qa, const float4* __restrict__ qb,
                                                   Dims3 d, QuadDims q, float3 n, float3 st, int i0, int last,
                                                   float va[U], float vb[U], bool inside[U]) {
  unsigned a[U], a1[U];
  float w[U][3];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int i = i0 + u;
    const float fi = (float)i;
    const float px = fmaf(fi, st.x, n.x) - 0.5f, py = fmaf(fi, st.y, n.y) - 0.5f, pz = fmaf(fi, st.z, n.z) - 0.5f;
    const float qx = px + 0.5f, qy = py + 0.5f, qz = pz + 0.5f;
    inside[u] = (i <= last) & (0.f <= qx) & (qx < (float)d.x) & (0.f <= qy) & (qy < (float)d.y) & (0.f <= qz) &
                (qz < (float)d.z);
    const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    w[u][0] = px - fx;
    w[u][1] = py - fy;
    w[u][2] = pz - fz;
    const unsigned ix = (unsigned)(clampi((int)fx, -1, d.x - 1) + 1), iy = (unsigned)(clampi((int)fy, -1, d.y - 1) + 1),
                   iz = (unsigned)(clampi((int)fz, -1, d.z - 1) + 1);
    a[u] = quad_index(q, ix, iy, iz);
    a1[u] = quad_index(q, ix, iy, iz + 1);
  }
  float4 c0[U], c1[U], e0[U], e1[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    c0[u] = qa[a[u]];
    c1[u] = qa[a1[u]];
    e0[u] = qb[a[u]];
    e1[u] = qb[a1[u]];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const float v0 = lerpf(lerpf(c0[u].x, c0[u].y, w[u][0]), lerpf(c0[u].z, c0[u].w, w[u][0]), w[u][1]);
    const float v1 = lerpf(lerpf(c1[u].x, c1[u].y, w[u][0]), lerpf(c1[u].z, c1[u].w, w[u][0]), w[u][1]);
    const float g0 = lerpf(lerpf(e0[u].x, e0[u].y, w[u][0]), lerpf(e0[u].z, e0[u].w, w[u][0]), w[u][1]);
    const float g1 = lerpf(lerpf(e1[u].x, e1[u].y, w[u][0]), lerpf(e1[u].z, e1[u].w, w[u][0]), w[u][1]);
    va[u] = inside[u] ? lerpf(v0, v1, w[u][2]) : 0.f;
    vb[u] = inside[u] ? lerpf(g0, g1, w[u][2]) : 0.f;
  }
}

// samples per forward round trip (every corner-quad load of the batch in
// flight): 4 keeps the forward at 87 VGPRs, 5 waves per SIMD, on the bricked
// copy (c2: 0.157 ms; 2: 0.181, 3: 0.185, 6: 0.164, 8 (147 VGPRs): 0.163)
constexpr int kBatch = 4;
// Per-pixel record (float4): the clipped sample range lo, hi (int bits in x, y;
// lo > hi: no sample, or after the backward's pixel pass, no gradient) and
// coefficients in (z, w): (E, 0) from the forward (E the pixel's emission
// sum), (alpha, beta) after the pixel pass. One 16-byte load per candidate
// pixel in the gather.
__device__ __forceinline__ float4 make_record(int lo, int hi, float a, float b) {
  return make_float4(__int_as_float(lo), __int_as_float(hi), a, b);
}
// saved value of a sample outside the grid (quiet NaN; never the value of a
// sample of a finite density)
#define kOutsideMark __builtin_nanf("")

__device__ __forceinline__ int wave_min(int v) {
  for (int m = 32; m > 0; m >>= 1) v = min(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
  for (int m = 32; m > 0; m >>= 1) v = max(v, __shfl_xor(v, m, 64));
  return v;
}

// ----------------------------------------------------------------- forward
// BORDER and CLAMP read `dens` as the corner-quad copy of that boundary
// (float4, QuadDims q); WRAP reads the x-fastest grid. BORDER marches the
// clipped range of each ray, CLAMP every sample. With gs != nullptr (BORDER, training) the
// forward also saves what the backward needs (RenderSaved): per pixel the
// sample range and E, per sample of the wave's clipped range a value c_i at
// samples[sample_base(pixel) + i * kSampleRun] (wave-major: one instruction's
// stores fill two whole lines; a NaN mark for a sample outside the grid).
template <int BM, class Q = float4>
__global__ __launch_bounds__(256) void k_render_density_fwd(const float* __restrict__ dens, float* __restrict__ img,
                                                            RayDims rd, RayLaunch L, RenderParams rp, QuadDims q,
                                                            RenderSaved sv) {
  const RayView& v = L.view[blockIdx.z];
  const int x = blockIdx.x * kRTileX + (threadIdx.x % kRTileX);
  const int y = blockIdx.y * kRTileY + (threadIdx.x / kRTileX);
  const bool valid = x < rd.out_w && y < rd.out_h;
  float D = 0.f, E = 0.f;
  if (BM == B_BORDER || BM == B_CLAMP) {
    // every lane of the wave takes part in the range reduction
    float3 nearp = make_float3(0.f, 0.f, 0.f), step = make_float3(0.f, 0.f, 0.f);
    int lo = 1, hi = 0;
    if (valid) {
      pixel_ray(v, rd, x, y, nearp, step);
      if (BM == B_BORDER) {
        clip_ray(nearp, step, rd.in, rd.out_d, 0.f, lo, hi);
      } else {
        lo = 0;
        hi = rd.out_d - 1;
      }
    }
    const size_t plane = size_t(rd.out_w) * rd.out_h;
    const size_t pix = valid ? size_t(y) * rd.out_w + x : 0;
    float* gview = (sv.samples && valid) ? sv.samples + sample_base(rd.out_w, rd.out_h, rd.out_d, v.out_slot, x, y)
                                         : nullptr;
    // march the wave's union of ranges with one sample index for all lanes, so
    // that the saved-value stores of one instruction fill one wave-major run
    const int wlo = wave_min(lo <= hi ? lo : rd.out_d), whi = wave_max(lo <= hi ? hi : -1);
    for (int i = wlo; i <= whi; i += kBatch) {
      float val[kBatch];
      bool in[kBatch];
      read_border_batch<kBatch, BM, Q>(reinterpret_cast<const Q*>(dens), rd.in, q, nearp, step, i, 1, hi, val, in);
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        // samples outside the grid or the lane's range read 0 and leave D and E unchanged
        const bool act = in[u] & (i + u >= lo);
        const float sigma = act ? rp.step * val[u] : 0.f;
        D = act ? fmaxf(D + sigma, 0.f) : D;
        const float e = __expf(-D);
        // saved for the backward: c_i = exp(-D_i) + E_{i-1}; a NaN mark for a
        // sample outside the grid (its gradient is 0)
        if (gview && i + u <= whi) gview[(size_t)(i + u) * kSampleRun] = act ? e + E : kOutsideMark;
        E = fmaf(e, sigma, E);
      }
    }
    if (gview) sv.records[(size_t)v.out_slot * plane + pix] = make_record(lo, hi, E, 0.f);
  } else {
    if (!valid) return;
    float3 nearp, step;
    pixel_ray(v, rd, x, y, nearp, step);
    const float* buf = dens + rd.in_batch_stride * v.in_batch;
    for (int i = 0; i < rd.out_d; ++i) {
      const float fi = (float)i;
      const float sigma = rp.step * read_linear<1, BM>(buf, rd.in, fmaf(fi, step.x, nearp.x) - 0.5f,
                                                       fmaf(fi, step.y, nearp.y) - 0.5f,
                                                       fmaf(fi, step.z, nearp.z) - 0.5f).v[0];
      D = fmaxf(D + sigma, 0.f);
      E = fmaf(__expf(-D), sigma, E);
    }
  }
  if (!valid) return;
  const float4 o = make_float4(rp.light[0] * E, rp.light[1] * E, rp.light[2] * E, D);
  reinterpret_cast<float4*>(img + rd.out_view_stride * v.out_slot)[size_t(y) * rd.out_w + x] = o;
}

// ------------------------------------------------------------ lit forward
// Density lit by a spot or point light and an ambient term (BORDER): per cell
// the emission is color * d * (lambda + a), with lambda the light's per-cell
// factor (intensity * shadow transmittance * falloff, lighting.py:106-147) and
// a the ambient intensity relative to `color`; the absorption is d. Both are
// scaled by the depth step and sampled trilinearly from the corner-quad copies
// of d and of d * lambda:
//   sigma_i = s d(p_i), eps_i = s (d lambda)(p_i),
//   D_i = max(D_{i-1} + sigma_i, 0), u_i = exp(-D_i),
//   E = sum_i u_i (eps_i + a sigma_i), rgb = color * E.
// With sv.samples the forward saves per sample c_i = a u_i + E_{i-1} (a NaN
// mark outside the grid) and u_i (sv.samples2): the backward's absorption
// gradient is alpha c_i + beta and its emission gradient alpha u_i.
constexpr int kLitBatch = 4;
__global__ __launch_bounds__(256) void k_render_density_lit_fwd(const float4* __restrict__ dq,
                                                                const float4* __restrict__ eq,
                                                                float* __restrict__ img, RayDims rd, RayLaunch L,
                                                                RenderParams rp, QuadDims q, RenderSaved sv) {
  const RayView& v = L.view[blockIdx.z];
  const int x = blockIdx.x * kRTileX + (threadIdx.x % kRTileX);
  const int y = blockIdx.y * kRTileY + (threadIdx.x / kRTileX);
  const bool valid = x < rd.out_w && y < rd.out_h;
  float D = 0.f, E = 0.f;
  float3 nearp = make_float3(0.f, 0.f, 0.f), step = make_float3(0.f, 0.f, 0.f);
  int lo = 1, hi = 0;
  if (valid) {
    pixel_ray(v, rd, x, y, nearp, step);
    clip_ray(nearp, step, rd.in, rd.out_d, 0.f, lo, hi);
  }
  const size_t plane = size_t(rd.out_w) * rd.out_h;
  const size_t pix = valid ? size_t(y) * rd.out_w + x : 0;
  const size_t vbase = valid ? sample_base(rd.out_w, rd.out_h, rd.out_d, v.out_slot, x, y) : 0;
  float* cview = (sv.samples && valid) ? sv.samples + vbase : nullptr;
  float* uview = (sv.samples2 && valid) ? sv.samples2 + vbase : nullptr;
  const int wlo = wave_min(lo <= hi ? lo : rd.out_d), whi = wave_max(lo <= hi ? hi : -1);
  const float a = rp.ambient;
  for (int i = wlo; i <= whi; i += kLitBatch) {
    float vd[kLitBatch], ve[kLitBatch];
    bool in[kLitBatch];
    read_border_batch2<kLitBatch>(dq, eq, rd.in, q, nearp, step, i, hi, vd, ve, in);
#pragma unroll
    for (int u = 0; u < kLitBatch; ++u) {
      const bool act = in[u] & (i + u >= lo);
      const float sigma = act ? rp.step * vd[u] : 0.f;
      const float eps = act ? rp.step * ve[u] : 0.f;
      D = act ? fmaxf(D + sigma, 0.f) : D;
      const float e = __expf(-D);
      if (cview && i + u <= whi) {
        cview[(size_t)(i + u) * kSampleRun] = act ? fmaf(a, e, E) : kOutsideMark;
        uview[(size_t)(i + u) * kSampleRun] = act ? e : 0.f;
      }
      E = fmaf(e, fmaf(a, sigma, eps), E);
    }
  }
  if (!valid) return;
  if (cview) sv.records[(size_t)v.out_slot * plane + pix] = make_record(lo, hi, E, 0.f);
  const float4 o = make_float4(rp.light[0] * E, rp.light[1] * E, rp.light[2] * E, D);
  reinterpret_cast<float4*>(img + rd.out_view_stride * v.out_slot)[size_t(y) * rd.out_w + x] = o;
}

// Per-pixel loss gradient: G (light-weighted rgb gradient), g_D (optical depth
// gradient) and the pixel's loss term. The folded loss compares
// e = rgb (+ bkg * exp(-D)) - target per channel: SE sums e^2 (gradient 2e),
// AE sums |e| (gradient sign(e), 0 at 0), as loss_dens_target and
// loss_dens_target_raw with the reference's SE / AE loss functions
// (optimization.py:94-130,396-430); the composited background makes the
// optical depth channel receive -exp(-D) sum_c bkg_c dL/de_c.
template <bool L2>
__device__ __forceinline__ void pixel_grad(const float* __restrict__ img, const float* __restrict__ second,
                                           const RayDims& rd, const RayView& v, size_t pix, const RenderParams& rp,
                                           float& G, float& gD, float& D, float& sq) {
  const float4 o = reinterpret_cast<const float4*>(img + rd.out_view_stride * v.out_slot)[pix];
  D = o.w;
  if (L2) {
    const size_t at3 = ((size_t)v.out_slot * rd.out_w * rd.out_h + pix) * 3;
    const float* t = second + at3;
    float e0 = o.x - t[0], e1 = o.y - t[1], e2 = o.z - t[2];
    float tr = 0.f, b0 = 0.f, b1 = 0.f, b2 = 0.f;
    if (rp.bkg) {
      tr = expf(-D);
      b0 = rp.bkg[at3];
      b1 = rp.bkg[at3 + 1];
      b2 = rp.bkg[at3 + 2];
      e0 = (o.x + b0 * tr) - t[0];
      e1 = (o.y + b1 * tr) - t[1];
      e2 = (o.z + b2 * tr) - t[2];
    }
    float f0, f1, f2;
    if (rp.loss_kind == LOSS_AE) {
      sq = fabsf(e0) + fabsf(e1) + fabsf(e2);
      f0 = (float)(e0 > 0.f) - (float)(e0 < 0.f);
      f1 = (float)(e1 > 0.f) - (float)(e1 < 0.f);
      f2 = (float)(e2 > 0.f) - (float)(e2 < 0.f);
    } else {
      sq = e0 * e0 + e1 * e1 + e2 * e2;
      f0 = 2.f * e0;
      f1 = 2.f * e1;
      f2 = 2.f * e2;
    }
    G = rp.loss_scale * (rp.light[0] * f0 + rp.light[1] * f1 + rp.light[2] * f2);
    gD = rp.bkg ? -(rp.loss_scale * tr) * (b0 * f0 + b1 * f1 + b2 * f2) : 0.f;
  } else {
    const float4 go = reinterpret_cast<const float4*>(second + rd.out_view_stride * v.out_slot)[pix];
    G = rp.light[0] * go.x + rp.light[1] * go.y + rp.light[2] * go.z;
    gD = go.w;
    sq = 0.f;
  }
}

__device__ __forceinline__ void block_loss(float sq, float* __restrict__ loss, float scale) {
  __shared__ float red[256 / 64];
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int w = 0; w < 256 / 64; ++w) s += red[w];
    atomic_addf(loss, s * scale);
  }
}

// -------------------------------------------------- BORDER: pixel pass + gather
// With the forward's D_i = max(D_{i-1} + sigma_i, 0) (no clamp active),
// E = sum_i exp(-D_i) sigma_i and the pixel gradients G (light-weighted rgb)
// and g_D (optical depth), the gradient of sample k is
//   dL/dd_k = s * (G * (exp(-D_k) - sum_{i>=k} exp(-D_i) sigma_i) + g_D)
//           = alpha * c_k + beta,  c_k = exp(-D_k) + E_{k-1},
//   alpha = s * G,  beta = s * (g_D - G * E).
// The forward saves c_k per sample and E per pixel, so the backward needs no
// walk along the ray: this pass turns each pixel's loss gradient into (alpha,
// beta) (and the L2 loss), the gather evaluates alpha * c_k + beta per sample.
// A pixel with no gradient gets lo > hi so the gather skips it.
// loss = scale * sum of the n tile partials, in a fixed order (deterministic):
// 256 strided sums, then a tree over them
__device__ __forceinline__ void loss_sum_block(const float* partials, int n, float scale, float* loss, float* red) {
  float t = 0.f;
  for (int i = threadIdx.x; i < n; i += blockDim.x) t += partials[i];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = red[0] * scale;
}

template <bool L2>
__global__ __launch_bounds__(256) void k_render_density_bwd_pixel(const float* __restrict__ img,
                                                                  const float* __restrict__ second, RenderSaved sv,
                                                                  RayDims rd, RayLaunch L, RenderParams rp) {
  const RayView& v = L.view[blockIdx.z];
  const int x = blockIdx.x * kRTileX + (threadIdx.x % kRTileX);
  const int y = blockIdx.y * kRTileY + (threadIdx.x / kRTileX);
  float sq = 0.f;
  if (x < rd.out_w && y < rd.out_h) {
    const size_t pix = size_t(y) * rd.out_w + x;
    float G, gD, D;
    pixel_grad<L2>(img, second, rd, v, pix, rp, G, gD, D, sq);
    const size_t at = (size_t)v.out_slot * rd.out_w * rd.out_h + pix;
    const float4 rec = sv.records[at];
    const int lo = __float_as_int(rec.x), hi = __float_as_int(rec.y);
    if (lo <= hi)
      sv.records[at] = (G == 0.f && gD == 0.f) ? make_record(1, 0, 0.f, 0.f)
                                               : make_record(lo, hi, rp.step * G, rp.step * fmaf(-G, rec.z, gD));
  }
  if (L2) {
    // per-tile partial sum (fixed order); the first workgroup of the gather
    // that follows adds the tiles of every view
    __shared__ float red[256 / 64];
    for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
    __syncthreads();
    if (threadIdx.x == 0) {
      float t = 0.f;
      for (int w = 0; w < 256 / 64; ++w) t += red[w];
      sv.partials[(size_t)v.out_slot * gridDim.x * gridDim.y + blockIdx.y * gridDim.x + blockIdx.x] = t;
    }
  }
}

// Gather: one wave per block of WX x WY x WZ voxels, accumulated in a
// wave-private FP64 LDS region. For each view the wave enumerates the pixels
// whose rays can cross the block's support box (x0-1, x0+WX) x ... (the
// projected box; lanes take consecutive pixels in row order), cuts each ray by
// the box and adds every inside sample's gradient alpha c_i + beta,
// trilinear-weighted, to the corner voxels that lie in the block (a sample
// q reaches voxel c with weight prod_a max(0, 1 - |q_a - c_a|)). The pixel loop
// is amortised over a whole block. Only this wave touches its region, so the
// sums run in a fixed order (deterministic) and one store per voxel follows.
constexpr int kGatherBatch = 4;  // saved values loaded per round trip (2: 0.187 ms, 8: 0.211 ms at c2)
// voxel blocks (one per wave) per workgroup: every c2 block is resident at
// once, so 4 per workgroup costs no waiting there (c2: 0.175 ms; 1: 0.178);
// c5 runs several rounds (1: 1.648 ms against 1.676)
constexpr int kRGatherWaves = 4;
//
// LIT (the spot-lit render): a second region sums the emission gradients
// alpha u_i * w of the d * lambda grid; the store then writes
// grad = sum_sigma + lambda * sum_eps and lambda_grad = d * sum_eps.
//
// With `loss` non-null (the last launch of a backward with the image loss
// folded in), workgroup 0 first sums the pixel pass's per-tile loss partials
// of every view in a fixed order; the kernel boundary after the pixel pass
// makes them visible, so no separate launch is needed.
template <int WX, int WY, int WZ, bool LIT = false>
__global__ __launch_bounds__(256) void k_render_density_bwd_gather(RenderSaved sv, float* __restrict__ grad,
                                                                   RayDims rd, ProjLaunch PL,
                                                                   const float* __restrict__ dens,
                                                                   const float* __restrict__ lam,
                                                                   float* __restrict__ lam_grad,
                                                                   float* __restrict__ loss, float loss_scale) {
  if (loss && blockIdx.x == 0) {
    __shared__ float red[256];
    loss_sum_block(sv.partials, sv.n_partials, loss_scale, loss, red);
  }
  // rows padded to WX + 1 doubles: lanes at one x in neighbouring rows fall on
  // different LDS banks (16 doubles are 128 bytes, the banks' width)
  constexpr int kRX = WX + 1;
  constexpr int kVox = kRX * WY * WZ;
  __shared__ double region[LIT ? 2 * kRGatherWaves : kRGatherWaves][kVox];
  typedef __attribute__((address_space(3))) double lds_double;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  lds_double* acc = (lds_double*)region[wave];
  lds_double* acc2 = (lds_double*)region[LIT ? kRGatherWaves + wave : wave];
  const Dims3 d = rd.in;
  const int nbx = (d.x + WX - 1) / WX, nby = (d.y + WY - 1) / WY, nbz = (d.z + WZ - 1) / WZ;
  const int blk = blockIdx.x * kRGatherWaves + wave;
  const bool live = blk < nbx * nby * nbz;  // a wave past the last block only joins the barriers
  const int x0 = (blk % nbx) * WX, y0 = ((blk / nbx) % nby) * WY, z0 = (blk / (nbx * nby)) * WZ;
  for (int k = lane; k < kVox; k += 64) acc[k] = 0.0;
  if (LIT)
    for (int k = lane; k < kVox; k += 64) acc2[k] = 0.0;
  const size_t plane = size_t(rd.out_w) * rd.out_h;
  const float fx0 = (float)x0, fy0 = (float)y0, fz0 = (float)z0;
  // support box in ray coordinates (which carry the +0.5 cell-centre offset)
  const float bx0 = fx0 - 0.5f, by0 = fy0 - 0.5f, bz0 = fz0 - 0.5f;
  const float bx1 = bx0 + (float)(WX + 1), by1 = by0 + (float)(WY + 1), bz1 = bz0 + (float)(WZ + 1);
  const float hx = 0.5f * (float)(WX + 1), hy = 0.5f * (float)(WY + 1), hz = 0.5f * (float)(WZ + 1);
  for (int vi = 0; vi < PL.num_views && live; ++vi) {
    const ProjView& pv = PL.view[vi];
    const float* A = pv.pmv;
    const float4 cc = mat_vec(A, fx0 + 0.5f * (float)WX, fy0 + 0.5f * (float)WY, fz0 + 0.5f * (float)WZ, 1.f);
    float mnx = 3.4e38f, mxx = -3.4e38f, mny = 3.4e38f, mxy = -3.4e38f;
    bool behind = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float sx = (k & 1) ? hx : -hx, sy = (k & 2) ? hy : -hy, sz = (k & 4) ? hz : -hz;
      const float cxk = fmaf(sz, A[8], fmaf(sy, A[4], fmaf(sx, A[0], cc.x)));
      const float cyk = fmaf(sz, A[9], fmaf(sy, A[5], fmaf(sx, A[1], cc.y)));
      const float cwk = fmaf(sz, A[11], fmaf(sy, A[7], fmaf(sx, A[3], cc.w)));
      behind |= !(cwk > 0.f);
      const float iw = __builtin_amdgcn_rcpf(cwk);
      const float px = fmaf(cxk * iw, 0.5f, 0.5f) * (float)rd.out_w, py = fmaf(cyk * iw, 0.5f, 0.5f) * (float)rd.out_h;
      mnx = fminf(mnx, px);
      mxx = fmaxf(mxx, px);
      mny = fminf(mny, py);
      mxy = fmaxf(mxy, py);
    }
    int xs = 0, xe = rd.out_w - 1, ys = 0, ye = rd.out_h - 1;
    if (!behind) {
      xs = max(xs, (int)ceilf(mnx - 0.5f - 1e-2f));
      xe = min(xe, (int)floorf(mxx - 0.5f + 1e-2f));
      ys = max(ys, (int)ceilf(mny - 0.5f - 1e-2f));
      ye = min(ye, (int)floorf(mxy - 0.5f + 1e-2f));
    }
    if (xs > xe || ys > ye) continue;

    const float4* rview = sv.records + (size_t)pv.out_slot * plane;
    {
      const int nx = xe - xs + 1, cnt = nx * (ye - ys + 1);
      // lane k of the box's row-major pixel list; the next record (k + 64) is
      // loaded while the current one is walked
      const int dy64 = 64 / nx, dx64 = 64 - dy64 * nx;
      int px = xs + lane % nx, py = ys + lane / nx;
      float4 rec_next = lane < cnt ? rview[size_t(py) * rd.out_w + px] : make_float4(0.f, 0.f, 0.f, 0.f);
      for (int k = lane; k < cnt; k += 64) {
        const size_t pix = size_t(py) * rd.out_w + px;
        const float fpx = (float)px, fpy = (float)py;
        const float4 rec = rec_next;
        px += dx64;
        py += dy64;
        if (px > xe) {
          px -= nx;
          ++py;
        }
        if (k + 64 < cnt) rec_next = rview[size_t(py) * rd.out_w + px];
        const int lo = __float_as_int(rec.x), hi = __float_as_int(rec.y);
        if (lo > hi) continue;
        const float4 gn = make_float4(fmaf(fpy, pv.gy[0], fmaf(fpx, pv.gx[0], pv.near0[0])),
                                      fmaf(fpy, pv.gy[1], fmaf(fpx, pv.gx[1], pv.near0[1])),
                                      fmaf(fpy, pv.gy[2], fmaf(fpx, pv.gx[2], pv.near0[2])),
                                      fmaf(fpy, pv.gy[3], fmaf(fpx, pv.gx[3], pv.near0[3])));
        const float4 gf = make_float4(fmaf(fpy, pv.gy[0], fmaf(fpx, pv.gx[0], pv.far0[0])),
                                      fmaf(fpy, pv.gy[1], fmaf(fpx, pv.gx[1], pv.far0[1])),
                                      fmaf(fpy, pv.gy[2], fmaf(fpx, pv.gx[2], pv.far0[2])),
                                      fmaf(fpy, pv.gy[3], fmaf(fpx, pv.gx[3], pv.far0[3])));
        const float nw = __builtin_amdgcn_rcpf(gn.w), fw = __builtin_amdgcn_rcpf(gf.w);
        const float nx_ = gn.x * nw, ny_ = gn.y * nw, nz_ = gn.z * nw;
        const float sx = (gf.x * fw - nx_) * PL.inv_zmax, sy = (gf.y * fw - ny_) * PL.inv_zmax,
                    sz = (gf.z * fw - nz_) * PL.inv_zmax;
        const float irx = __builtin_amdgcn_rcpf(sx), iry = __builtin_amdgcn_rcpf(sy), irz = __builtin_amdgcn_rcpf(sz);
        // both slab ends as products: a zero step component (reciprocal inf)
        // then gives (-inf, inf) for a ray inside that slab and an empty
        // interval outside it
        const float ax = (bx0 - nx_) * irx, bx = (bx1 - nx_) * irx;
        const float ay = (by0 - ny_) * iry, by = (by1 - ny_) * iry;
        const float az = (bz0 - nz_) * irz, bz = (bz1 - nz_) * irz;
        const float tlo = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
        const float thi = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
        const int i0 = max(lo, (int)ceilf(fmaxf(tlo, -1.f) - 1e-3f));
        const int i1 = min(hi, (int)floorf(fminf(thi, 1e9f) + 1e-3f));
        if (i0 > i1) continue;
        const float alpha = rec.z, beta = rec.w;
        // sample position in cell coordinates relative to the block origin
        const float fi0 = (float)i0;
        float qx = fmaf(fi0, sx, nx_) - (fx0 + 0.5f), qy = fmaf(fi0, sy, ny_) - (fy0 + 0.5f),
              qz = fmaf(fi0, sz, nz_) - (fz0 + 0.5f);
        const size_t sb = sample_base(rd.out_w, rd.out_h, rd.out_d, pv.out_slot, (int)fpx, (int)fpy) +
                          (size_t)i0 * kSampleRun;
        const float* gp = sv.samples + sb;
        const float* up = LIT ? sv.samples2 + sb : nullptr;
        for (int ib = i0; ib <= i1; ib += kGatherBatch) {
          // the batch's saved values are all loaded before the first is used
          float cs[kGatherBatch], us[kGatherBatch];
#pragma unroll
          for (int u = 0; u < kGatherBatch; ++u) cs[u] = ib + u <= i1 ? gp[u * kSampleRun] : kOutsideMark;
          if (LIT) {
#pragma unroll
            for (int u = 0; u < kGatherBatch; ++u) us[u] = ib + u <= i1 ? up[u * kSampleRun] : 0.f;
            up += kGatherBatch * kSampleRun;
          }
          gp += kGatherBatch * kSampleRun;
#pragma unroll
          for (int u = 0; u < kGatherBatch; ++u) {
          const float c = cs[u];
          if (c == c) {  // NaN mark: sample outside the grid (or past the segment)
            const float g = fmaf(alpha, c, beta);
            const float flx = floorf(qx), fly = floorf(qy), flz = floorf(qz);
            const float tx = qx - flx, ty = qy - fly, tz = qz - flz;
            const int lx = (int)flx, ly = (int)fly, lz = (int)flz;
            const float wx[2] = {g * (1.f - tx), g * tx}, wy[2] = {1.f - ty, ty}, wz[2] = {1.f - tz, tz};
            const float ge = LIT ? alpha * us[u] : 0.f;
            const float ex[2] = {ge * (1.f - tx), ge * tx};
            // corners outside the block are skipped (a zero-weight LDS atomic
            // costs as much as any other: the LDS atomic rate bounds this loop)
#pragma unroll
            for (int kz = 0; kz < 2; ++kz) {
              const int cz = lz + kz;
              if ((unsigned)cz >= (unsigned)WZ) continue;
#pragma unroll
              for (int ky = 0; ky < 2; ++ky) {
                const int cy = ly + ky;
                if ((unsigned)cy >= (unsigned)WY) continue;
                const float wyz = wy[ky] * wz[kz];
#pragma unroll
                for (int kx = 0; kx < 2; ++kx) {
                  const int cx = lx + kx;
                  if ((unsigned)cx >= (unsigned)WX) continue;
                  __hip_atomic_fetch_add(acc + (cz * WY + cy) * kRX + cx, (double)(wx[kx] * wyz), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
                  if (LIT)
                    __hip_atomic_fetch_add(acc2 + (cz * WY + cy) * kRX + cx, (double)(ex[kx] * wyz),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
              }
            }
          }
          qx += sx;
          qy += sy;
          qz += sz;
          }
        }
      }
    }
  }
  __syncthreads();
  if (!live) return;
  for (int k0 = lane; k0 < WX * WY * WZ; k0 += 64) {
    const int x = x0 + k0 % WX, y = y0 + (k0 / WX) % WY, z = z0 + k0 / (WX * WY);
    const int k = (k0 / WX) * kRX + k0 % WX;
    if (x < d.x && y < d.y && z < d.z) {
      const size_t f = flat3(x, y, z, d);
      float* g = grad + f;
      if (LIT) {
        const double se = acc2[k];
        const float v = (float)(acc[k] + (double)lam[f] * se), vl = (float)((double)dens[f] * se);
        *g = PL.accumulate ? *g + v : v;
        lam_grad[f] = PL.accumulate ? lam_grad[f] + vl : vl;
      } else {
        const float v = (float)acc[k];
        *g = PL.accumulate ? *g + v : v;
      }
    }
  }
}

// ----------------------------------------- CLAMP: tiled LDS pre-reduction
// A workgroup owns a 16x16 pixel tile and walks its rays in lock-step
// segments of K samples from far to near. The samples of a segment lie in the
// convex hull of the tile's four corner rays at the segment's first and last
// sample (the index->object map is projective), so their trilinear corners fit
// in an axis-aligned brick; the scatter is accumulated there with LDS atomics
// and flushed once per segment with coalesced global atomics (zero entries
// skipped). A brick over the LDS capacity, or a corner outside the brick
// through rounding, goes straight to a global atomic.
constexpr int kSegK = 8;
constexpr int kCap = 6144;  // FP64 brick cells: ds_add_f64 is ~20x faster than ds_add_f32 on gfx950

struct Brick {
  int lo[3], n[3];
  int size;
};

__device__ __forceinline__ Brick segment_brick(const float3* cn, const float3* cs, int ia, int ib, Dims3 d) {
  float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
#pragma unroll
  for (int c = 0; c < 4; ++c) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const float fi = (float)(e ? ib : ia);
      const float p[3] = {fmaf(fi, cs[c].x, cn[c].x) - 0.5f, fmaf(fi, cs[c].y, cn[c].y) - 0.5f,
                          fmaf(fi, cs[c].z, cn[c].z) - 0.5f};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        mn[a] = fminf(mn[a], p[a]);
        mx[a] = fmaxf(mx[a], p[a]);
      }
    }
  }
  const int dim[3] = {d.x, d.y, d.z};
  Brick b;
  b.size = 1;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    // every corner index is clamp(floor(p)) or clamp(floor(p)+1)
    const int lo = clampi((int)floorf(mn[a] - 1e-3f), 0, dim[a] - 1);
    const int hi = clampi((int)floorf(mx[a] + 1e-3f) + 1, 0, dim[a] - 1);
    b.lo[a] = lo;
    b.n[a] = hi - lo + 1;
    b.size *= b.n[a];
  }
  return b;
}

// `quad` is the corner-quad copy of the density for the boundary (the
// forward's copy): a segment's 8 samples are read with all their loads in
// flight, and give the forward's sigma_i bit for bit (the same corners,
// weights and lerp order).
// BORDER (the recompute alternative to the saved-state gather: nothing is
// kept from the forward but the images) walks only the segments of the
// tile's clipped sample ranges; samples outside the grid read 0 and corners
// outside it take no gradient. Its global float atomics make the summation
// order, and so the last bits, vary from run to run.
template <bool L2, int BM = B_CLAMP>
__global__ __launch_bounds__(256) void k_render_density_bwd_clamp(const float4* __restrict__ quad, QuadDims q,
                                                                  const float* __restrict__ img,
                                                                  const float* __restrict__ second,
                                                                  float* __restrict__ grad, float* __restrict__ loss,
                                                                  RayDims rd, RayLaunch L, RenderParams rp) {
  __shared__ double lg[kCap];
  __shared__ int seg_range[2];
  const RayView& v = L.view[blockIdx.z];
  const int x0 = blockIdx.x * kRTileX, y0 = blockIdx.y * kRTileY;
  const int x = x0 + (threadIdx.x % kRTileX), y = y0 + (threadIdx.x / kRTileX);
  const bool active = x < rd.out_w && y < rd.out_h;
  float* gout = grad + rd.in_batch_stride * v.in_batch;
  float3 cn[4], cs[4];
  {
    const int x1 = min(x0 + kRTileX - 1, rd.out_w - 1), y1 = min(y0 + kRTileY - 1, rd.out_h - 1);
    const int px[4] = {x0, x1, x0, x1}, py[4] = {y0, y0, y1, y1};
#pragma unroll
    for (int c = 0; c < 4; ++c) pixel_ray(v, rd, px[c], py[c], cn[c], cs[c]);
  }
  float sq = 0.f, G = 0.f, gD = 0.f, D = 0.f;
  float3 nearp = make_float3(0.f, 0.f, 0.f), step = nearp;
  if (active) {
    pixel_grad<L2>(img, second, rd, v, size_t(y) * rd.out_w + x, rp, G, gD, D, sq);
    pixel_ray(v, rd, x, y, nearp, step);
  }
  const bool walk = active && (G != 0.f || gD != 0.f);
  for (int i = threadIdx.x; i < kCap; i += 256) lg[i] = 0.0;
  const int S = rd.out_d;
  int sg_lo = 0, sg_hi = (S + kSegK - 1) / kSegK - 1;
  if (BM == B_BORDER) {
    // the union of the tile's clipped ranges, in whole segments
    int lo = 1, hi = 0;
    if (walk) clip_ray(nearp, step, rd.in, S, 0.f, lo, hi);
    const int wlo = wave_min(lo <= hi ? lo : S), whi = wave_max(lo <= hi ? hi : -1);
    if (threadIdx.x == 0) {
      seg_range[0] = S;
      seg_range[1] = -1;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
      atomicMin(&seg_range[0], wlo);
      atomicMax(&seg_range[1], whi);
    }
    __syncthreads();
    sg_lo = seg_range[0] / kSegK;
    sg_hi = seg_range[1] < 0 ? -1 : seg_range[1] / kSegK;
  } else {
    __syncthreads();
  }
  for (int sgi = sg_hi; sgi >= sg_lo; --sgi) {
    const int ia = sgi * kSegK, ib = min(ia + kSegK, S) - 1;
    const Brick br = segment_brick(cn, cs, ia, ib, rd.in);
    const bool use_lds = br.size <= kCap;
    if (walk) {
      float val[kSegK];
      bool inr[kSegK];
      read_border_batch<kSegK, BM>(quad, rd.in, q, nearp, step, ib, -1, ia, val, inr);
#pragma unroll
      for (int u = 0; u < kSegK; ++u) {
        const int i = ib - u;
        if (i < ia) break;
        const float fi = (float)i;
        const Taps tp = make_taps<BM>(fmaf(fi, step.x, nearp.x) - 0.5f, fmaf(fi, step.y, nearp.y) - 0.5f,
                                      fmaf(fi, step.z, nearp.z) - 0.5f, rd.in);
        const float sigma = rp.step * val[u];
        const float e = __expf(-D);
        D = fmaxf(D - sigma, 0.f);
        const float esg = e * sigma * G;
        const float gds = rp.step * (e * G + (gD - esg));
        gD -= esg;
        if (BM == B_BORDER && !inr[u]) continue;  // outside the grid: no gradient
        float wx[2] = {1.f - tp.wx, tp.wx}, wy[2] = {1.f - tp.wy, tp.wy}, wz[2] = {1.f - tp.wz, tp.wz};
        if (BM == B_CLAMP) {
          // an axis whose two corners clamp to one cell (the sample lies outside
          // the grid, or in the last half cell, along it) gives that cell both
          // weights: one tap of weight 1 instead of two. Most of a CLAMP ray's
          // samples lie outside the grid and would otherwise spend 8 atomics on
          // 4, 2 or 1 distinct cells.
          if (tp.ix[0] == tp.ix[1]) { wx[0] = 1.f; wx[1] = 0.f; }
          if (tp.iy[0] == tp.iy[1]) { wy[0] = 1.f; wy[1] = 0.f; }
          if (tp.iz[0] == tp.iz[1]) { wz[0] = 1.f; wz[1] = 0.f; }
        }
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) {
              const float w = wx[ii] * wy[j] * wz[k];
              if (w == 0.f) continue;
              if (BM == B_BORDER && !(tp.vx[ii] & tp.vy[j] & tp.vz[k])) continue;
              const int bx = tp.ix[ii] - br.lo[0], by = tp.iy[j] - br.lo[1], bz = tp.iz[k] - br.lo[2];
              const bool in = use_lds & ((unsigned)bx < (unsigned)br.n[0]) & ((unsigned)by < (unsigned)br.n[1]) &
                              ((unsigned)bz < (unsigned)br.n[2]);
              if (in) atomicAdd(&lg[(bz * br.n[1] + by) * br.n[0] + bx], (double)(gds * w));
              else atomic_addf(gout + flat3(tp.ix[ii], tp.iy[j], tp.iz[k], rd.in), gds * w);
            }
      }
    }
    if (use_lds) {
      __syncthreads();
      const int nx = br.n[0], ny = br.n[1];
      const int q = 256 / nx, r = 256 - q * nx;
      int bx = threadIdx.x % nx, t = threadIdx.x / nx;
      int by = t % ny, bz = t / ny;
      for (int idx = threadIdx.x; idx < br.size; idx += 256) {
        const float val = (float)lg[idx];
        if (val != 0.f) {
          atomic_addf(gout + flat3(br.lo[0] + bx, br.lo[1] + by, br.lo[2] + bz, rd.in), val);
          lg[idx] = 0.0;
        }
        bx += r;
        by += q;
        if (bx >= nx) {
          bx -= nx;
          ++by;
        }
        while (by >= ny) {
          by -= ny;
          ++bz;
        }
      }
      __syncthreads();
    }
  }
  if (L2) block_loss(sq, loss, rp.loss_scale);
}

// -------------------------------------------------- WRAP: global atomics
template <bool L2>
__global__ __launch_bounds__(256) void k_render_density_bwd_wrap(const float* __restrict__ dens,
                                                                 const float* __restrict__ img,
                                                                 const float* __restrict__ second,
                                                                 float* __restrict__ grad, float* __restrict__ loss,
                                                                 RayDims rd, RayLaunch L, RenderParams rp) {
  const RayView& v = L.view[blockIdx.z];
  const int x = blockIdx.x * kRTileX + (threadIdx.x % kRTileX);
  const int y = blockIdx.y * kRTileY + (threadIdx.x / kRTileX);
  float sq = 0.f;
  if (x < rd.out_w && y < rd.out_h) {
    float G, gD, D;
    pixel_grad<L2>(img, second, rd, v, size_t(y) * rd.out_w + x, rp, G, gD, D, sq);
    float3 nearp, step;
    pixel_ray(v, rd, x, y, nearp, step);
    const float* buf = dens + rd.in_batch_stride * v.in_batch;
    float* gout = grad + rd.in_batch_stride * v.in_batch;
    if (G != 0.f || gD != 0.f) {
      for (int i = rd.out_d - 1; i >= 0; --i) {
        const float fi = (float)i;
        const Taps tp = make_taps<B_WRAP>(fmaf(fi, step.x, nearp.x) - 0.5f, fmaf(fi, step.y, nearp.y) - 0.5f,
                                          fmaf(fi, step.z, nearp.z) - 0.5f, rd.in);
        float c[2][2][2];
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) c[k][j][ii] = buf[flat3(tp.ix[ii], tp.iy[j], tp.iz[k], rd.in)];
        const float v0 = lerpf(lerpf(c[0][0][0], c[0][0][1], tp.wx), lerpf(c[0][1][0], c[0][1][1], tp.wx), tp.wy);
        const float v1 = lerpf(lerpf(c[1][0][0], c[1][0][1], tp.wx), lerpf(c[1][1][0], c[1][1][1], tp.wx), tp.wy);
        const float sigma = rp.step * lerpf(v0, v1, tp.wz);
        const float e = __expf(-D);
        D = fmaxf(D - sigma, 0.f);
        const float esg = e * sigma * G;
        const float gds = rp.step * (e * G + (gD - esg));
        gD -= esg;
        const float wx[2] = {1.f - tp.wx, tp.wx}, wy[2] = {1.f - tp.wy, tp.wy}, wz[2] = {1.f - tp.wz, tp.wz};
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) {
              const float w = wx[ii] * wy[j] * wz[k];
              if (w != 0.f) atomic_addf(gout + flat3(tp.ix[ii], tp.iy[j], tp.iz[k], rd.in), gds * w);
            }
      }
    }
  }
  if (L2) block_loss(sq, loss, rp.loss_scale);
}

// ---------------------------------------------------------------- launches
static dim3 tile_grid(const RayDims& rd, int views) {
  return dim3((rd.out_w + kRTileX - 1) / kRTileX, (rd.out_h + kRTileY - 1) / kRTileY, views);
}

size_t quad_copy_floats(Dims3 d) { return quad_entries(make_quad_dims(d)) * 4; }
size_t quad_copy_bf16_bytes(Dims3 d) { return quad_entries<uint2>(make_quad_dims<uint2>(d)) * sizeof(uint2); }

// bfloat16 corner-quad copy (BORDER) of a float grid
hipError_t quad_copy_bf16_launch(hipStream_t s, const float* lin, void* out, Dims3 d) {
  const QuadDims q = make_quad_dims<uint2>(d);
  const size_t n = quad_entries<uint2>(q);
  KernelTimer t(s, "quad_copy");
  hipLaunchKernelGGL((k_quad_copy<B_BORDER, uint2>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lin,
                     nullptr, reinterpret_cast<uint2*>(out), d, q);
  return hipGetLastError();
}

// BORDER forward from the bfloat16 corner-quad copy
hipError_t render_density_fwd_bf16_launch(hipStream_t s, const void* quad, float* img, const RayDims& rd,
                                          const RayLaunch& L, const RenderParams& rp, const RenderSaved& sv) {
  const dim3 grid = tile_grid(rd, L.num_views);
  const QuadDims q = make_quad_dims<uint2>(rd.in);
  KernelTimer t(s, "render_fwd");
  hipLaunchKernelGGL((k_render_density_fwd<B_BORDER, uint2>), grid, dim3(256), 0, s,
                     reinterpret_cast<const float*>(quad), img, rd, L, rp, q, sv);
  return hipGetLastError();
}

// bm: B_BORDER or B_CLAMP
// fac: null, or a per-voxel factor the copy multiplies in (lin * fac)
hipError_t quad_copy_launch(hipStream_t s, int bm, const float* lin, float* out, Dims3 d, const float* fac) {
  const QuadDims q = make_quad_dims(d);
  const size_t n = quad_entries(q);
  KernelTimer t(s, "quad_copy");
  if (bm == B_CLAMP)
    hipLaunchKernelGGL(k_quad_copy<B_CLAMP>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lin, fac,
                       reinterpret_cast<float4*>(out), d, q);
  else
    hipLaunchKernelGGL(k_quad_copy<B_BORDER>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, lin, fac,
                       reinterpret_cast<float4*>(out), d, q);
  return hipGetLastError();
}

// spot-lit forward (BORDER): dq, eq the corner-quad copies of d and d * lambda
hipError_t render_density_lit_fwd_launch(hipStream_t s, const float* dq, const float* eq, float* img,
                                         const RayDims& rd, const RayLaunch& L, const RenderParams& rp,
                                         const RenderSaved& sv) {
  const dim3 grid = tile_grid(rd, L.num_views);
  const QuadDims q = make_quad_dims(rd.in);
  KernelTimer t(s, "render_lit_fwd");
  hipLaunchKernelGGL(k_render_density_lit_fwd, grid, dim3(256), 0, s, reinterpret_cast<const float4*>(dq),
                     reinterpret_cast<const float4*>(eq), img, rd, L, rp, q, sv);
  return hipGetLastError();
}

// BORDER, CLAMP: dens is the corner-quad copy of that boundary.
hipError_t render_density_fwd_launch(int bm, hipStream_t s, const float* dens, float* img, const RayDims& rd,
                                     const RayLaunch& L, const RenderParams& rp, const RenderSaved& sv) {
  const dim3 grid = tile_grid(rd, L.num_views);
  const QuadDims q = make_quad_dims(rd.in);
  const RenderSaved none{nullptr, nullptr, nullptr};
  KernelTimer t(s, "render_fwd");
  switch (bm) {
    case B_BORDER: hipLaunchKernelGGL(k_render_density_fwd<B_BORDER>, grid, dim3(256), 0, s, dens, img, rd, L, rp, q, sv); break;
    case B_CLAMP: hipLaunchKernelGGL(k_render_density_fwd<B_CLAMP>, grid, dim3(256), 0, s, dens, img, rd, L, rp, q, none); break;
    default: hipLaunchKernelGGL(k_render_density_fwd<B_WRAP>, grid, dim3(256), 0, s, dens, img, rd, L, rp, q, none); break;
  }
  return hipGetLastError();
}

size_t render_tiles_per_view(const RayDims& rd) {
  const dim3 g = tile_grid(rd, 1);
  return (size_t)g.x * g.y;
}

// BORDER pixel pass (per-tile loss partials and per-pixel gradient coefficients)
// for one chunk of views; with the loss folded in, the last block of the last
// chunk writes the loss
hipError_t render_density_pixel_launch(bool l2, hipStream_t s, const float* img, const float* second,
                                       const RenderSaved& sv, const RayDims& rd, const RayLaunch& L,
                                       const RenderParams& rp) {
  const dim3 grid = tile_grid(rd, L.num_views);
  KernelTimer t(s, "render_bwd_pixel");
  if (l2) hipLaunchKernelGGL(k_render_density_bwd_pixel<true>, grid, dim3(256), 0, s, img, second, sv, rd, L, rp);
  else hipLaunchKernelGGL(k_render_density_bwd_pixel<false>, grid, dim3(256), 0, s, img, second, sv, rd, L, rp);
  return hipGetLastError();
}

constexpr int kGatherWX = 16, kGatherWY = 16, kGatherWZ = 2;

// loss: null, or where workgroup 0 writes loss_scale * the sum of the loss partials
hipError_t render_density_gather_launch(hipStream_t s, const RenderSaved& sv, float* grad, const RayDims& rd,
                                        const ProjLaunch& PL, float* loss, float loss_scale) {
  const int nb = ((rd.in.x + kGatherWX - 1) / kGatherWX) * ((rd.in.y + kGatherWY - 1) / kGatherWY) *
                 ((rd.in.z + kGatherWZ - 1) / kGatherWZ);
  KernelTimer t(s, "render_bwd_gather");
  auto kern = k_render_density_bwd_gather<kGatherWX, kGatherWY, kGatherWZ>;
  hipLaunchKernelGGL(kern, dim3((nb + kRGatherWaves - 1) / kRGatherWaves), dim3(64 * kRGatherWaves), 0, s, sv, grad, rd,
                     PL, nullptr, nullptr, nullptr, loss, loss_scale);
  return hipGetLastError();
}

// spot-lit gather: grad and lam_grad from the saved c_i and u_i arrays
hipError_t render_density_lit_gather_launch(hipStream_t s, const RenderSaved& sv, const float* dens,
                                            const float* lam, float* grad, float* lam_grad, const RayDims& rd,
                                            const ProjLaunch& PL, float* loss, float loss_scale) {
  const int nb = ((rd.in.x + kGatherWX - 1) / kGatherWX) * ((rd.in.y + kGatherWY - 1) / kGatherWY) *
                 ((rd.in.z + kGatherWZ - 1) / kGatherWZ);
  KernelTimer t(s, "render_lit_bwd_gather");
  auto kern = k_render_density_bwd_gather<kGatherWX, kGatherWY, kGatherWZ, true>;
  hipLaunchKernelGGL(kern, dim3((nb + kRGatherWaves - 1) / kRGatherWaves), dim3(64 * kRGatherWaves), 0, s, sv, grad, rd,
                     PL, dens, lam, lam_grad, loss, loss_scale);
  return hipGetLastError();
}

// scatter backward (grad must be zero on entry): BORDER and CLAMP walk the
// corner-quad copy of that boundary in dens; WRAP reads the x-fastest grid.
hipError_t render_density_scatter_launch(int bm, bool l2, hipStream_t s, const float* dens, const float* img,
                                         const float* second, float* grad, float* loss, const RayDims& rd,
                                         const RayLaunch& L, const RenderParams& rp) {
  const dim3 grid = tile_grid(rd, L.num_views);
  KernelTimer t(s, "render_bwd_scatter");
  if (bm == B_CLAMP) {
    const float4* quad = reinterpret_cast<const float4*>(dens);
    const QuadDims q = make_quad_dims(rd.in);
    if (l2) hipLaunchKernelGGL(k_render_density_bwd_clamp<true>, grid, dim3(256), 0, s, quad, q, img, second, grad, loss, rd, L, rp);
    else hipLaunchKernelGGL(k_render_density_bwd_clamp<false>, grid, dim3(256), 0, s, quad, q, img, second, grad, loss, rd, L, rp);
  } else if (bm == B_BORDER) {
    const float4* quad = reinterpret_cast<const float4*>(dens);
    const QuadDims q = make_quad_dims(rd.in);
    if (l2) hipLaunchKernelGGL((k_render_density_bwd_clamp<true, B_BORDER>), grid, dim3(256), 0, s, quad, q, img, second, grad, loss, rd, L, rp);
    else hipLaunchKernelGGL((k_render_density_bwd_clamp<false, B_BORDER>), grid, dim3(256), 0, s, quad, q, img, second, grad, loss, rd, L, rp);
  } else {
    if (l2) hipLaunchKernelGGL(k_render_density_bwd_wrap<true>, grid, dim3(256), 0, s, dens, img, second, grad, loss, rd, L, rp);
    else hipLaunchKernelGGL(k_render_density_bwd_wrap<false>, grid, dim3(256), 0, s, dens, img, second, grad, loss, rd, L, rp);
  }
  return hipGetLastError();
}

}  // namespace gft
