// Fused resample + blend raymarch (RaymarchGridTransform) for gfx950.
//
// Forward (raymarch_grid.cu.cc:107-143): one lane per pixel; the ray runs
// from IDXtoOS(x+.5, y+.5, .5) to IDXtoOS(x+.5, y+.5, zMax+.5) with
// zMax = max(D-1, 1), taking D equally spaced samples, each sampled from the
// NDHWC grid and blended front to back. All views of a launch go in one grid
// (blockIdx.z = view), so the camera table is a kernel argument instead of
// per-camera constant uploads.
//
// Backward (raymarch_grid.cu.cc:343-372): the ray is walked in reverse from
// the forward output; transmittance is recovered by subtracting each cell's
// optical depth. CLAMP scatters each sample's gradient to its 8 trilinear
// corners with float atomics (k_raymarch_bwd); BORDER stores every sample's
// gradient and sums them per voxel block without atomics in global memory
// (k_raymarch_bwd_walk + k_raymarch_bwd_gather, below).
//   EXACT     : true adjoint, all views summed into one gradient, sample
//               positions identical to the forward.
//   REFERENCE : per-camera normalisation by the accumulated trilinear weight
//               (kNormalize3DGradients, raymarch_grid.cu.cc:319-341), the
//               BORDER scatter band (-1, dims) and the far-to-near positions.
#include "gft_device.hpp"

namespace gft {

constexpr int kTileX = 16;
constexpr int kTileY = 16;
// samples read per round trip by the LINEAR ray walks (8 float4 corners per
// C=4 sample: two samples already hold 64 values in registers)
template <int C> constexpr int kRayBatch = C == 4 ? 2 : (C == 2 ? 4 : 8);

template <int C, int FM, int BM>
__device__ __forceinline__ Vec<C> sample_ray(const float* __restrict__ buf, Dims3 d, float3 p) {
  float px = p.x - 0.5f, py = p.y - 0.5f, pz = p.z - 0.5f;
  if constexpr (FM == F_LINEAR) return read_linear<C, BM>(buf, d, px, py, pz);
  else if constexpr (FM == F_NEAREST) return read_nearest<C, BM>(buf, d, px, py, pz);
  else return read_minmax_v2<C, BM>(buf, d, px, py, pz, FM == F_MAX);
}

__device__ __forceinline__ void ray_endpoints(const RayView& v, int x, int y, const RayDims& rd, float3& nearp,
                                              float3& farp, float3& step) {
  float dinvx = 1.f / (float)rd.out_w, dinvy = 1.f / (float)rd.out_h, dinvz = 1.f / (float)rd.out_d;
  float zmax = (float)max(rd.out_d - 1, 1);
  nearp = idx_to_os(v, x + 0.5f, y + 0.5f, 0.5f, dinvx, dinvy, dinvz);
  farp = idx_to_os(v, x + 0.5f, y + 0.5f, zmax + 0.5f, dinvx, dinvy, dinvz);
  float iz = 1.f / zmax;
  step = make_float3((farp.x - nearp.x) * iz, (farp.y - nearp.y) * iz, (farp.z - nearp.z) * iz);
}

template <int C, int FM, int BM, int BL>
__global__ __launch_bounds__(256) void k_raymarch_fwd(const float* __restrict__ input, float* __restrict__ output,
                                                      RayDims rd, RayLaunch L) {
  const int vi = blockIdx.z;
  const RayView& v = L.view[vi];
  const int x = blockIdx.x * kTileX + (threadIdx.x % kTileX);
  const int y = blockIdx.y * kTileY + (threadIdx.x / kTileX);
  if (x >= rd.out_w || y >= rd.out_h) return;
  float3 nearp, farp, step;
  ray_endpoints(v, x, y, rd, nearp, farp, step);
  const float* buf = input + rd.in_batch_stride * v.in_batch;
  Vec<C> acc = vzero<C>();
  // BORDER LINEAR / NEAREST: samples outside the box that any corner or the
  // truncated nearest index can reach read 0, and blending a 0 sample leaves the
  // accumulator unchanged in every mode, so only the clipped range is read
  // (MIN/MAX return +-FLT_MAX outside and are not clipped)
  int lo = 0, hi = rd.out_d - 1;
  if (BM == B_BORDER && (FM == F_LINEAR || FM == F_NEAREST))
    clip_ray(nearp, step, rd.in, rd.out_d, FM == F_LINEAR ? 0.f : 1.f, lo, hi);
  if constexpr (FM == F_LINEAR) {
    // kRayBatch<C> samples per round trip: all their corner loads in flight
    constexpr int U = kRayBatch<C>;
    for (int i = lo; i <= hi; i += U) {
      float px[U], py[U], pz[U];
      Vec<C> s[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float fi = (float)min(i + u, hi);  // past the range: a repeat, not blended
        px[u] = fmaf(fi, step.x, nearp.x) - 0.5f;
        py[u] = fmaf(fi, step.y, nearp.y) - 0.5f;
        pz[u] = fmaf(fi, step.z, nearp.z) - 0.5f;
      }
      read_linear_batch<C, BM, U>(buf, rd.in, px, py, pz, s);
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (i + u <= hi) blend<C, BL>(acc, s[u]);
    }
  } else {
    for (int i = lo; i <= hi; ++i) {
      float fi = (float)i;
      float3 p = make_float3(fmaf(fi, step.x, nearp.x), fmaf(fi, step.y, nearp.y), fmaf(fi, step.z, nearp.z));
      Vec<C> s = sample_ray<C, FM, BM>(buf, rd.in, p);
      blend<C, BL>(acc, s);
    }
  }
  vstore<C>(output + rd.out_view_stride * v.out_slot, size_t(y) * rd.out_w + x, acc);
}

// Scatter of one sample's gradient to the trilinear corners.
// EXACT: only samples whose forward read was inside (BORDER), valid corners.
// REFERENCE: BORDER accepts p in (-1, dims) and checks every corner; weights
// are also accumulated into wsum.
template <int C, int BM, bool EXACT>
__device__ __forceinline__ void scatter_sample(float* __restrict__ grad, float* __restrict__ wsum, Dims3 d, float px,
                                               float py, float pz, const Vec<C>& g) {
  if constexpr (BM == B_BORDER && !EXACT) {
    if (!((-1.f < px) & (px < (float)d.x) & (-1.f < py) & (py < (float)d.y) & (-1.f < pz) & (pz < (float)d.z))) return;
  }
  Taps t = make_taps<BM>(px, py, pz, d);
  if constexpr (BM == B_BORDER && EXACT) {
    if (!t.inside) return;
  }
  const float wx[2] = {1.f - t.wx, t.wx}, wy[2] = {1.f - t.wy, t.wy}, wz[2] = {1.f - t.wz, t.wz};
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        if (BM == B_BORDER && !(t.vx[i] & t.vy[j] & t.vz[k])) continue;
        float w = wx[i] * wy[j] * wz[k];
        // a zero weight adds nothing; this also covers the floor+1 corner at
        // integer positions, where the reference's ceil corner is the floor one
        if (w == 0.f) continue;
        size_t f = flat3(t.ix[i], t.iy[j], t.iz[k], d);
        vatomic_add<C>(grad, f, g, w);
        if (!EXACT) atomic_addf(wsum + f, w);
      }
}

template <int C, int BM, int BL, bool EXACT>
__global__ __launch_bounds__(256) void k_raymarch_bwd(const float* __restrict__ input,
                                                      const float* __restrict__ output,
                                                      const float* __restrict__ output_grad, float* __restrict__ grad,
                                                      float* __restrict__ wsum, RayDims rd, RayLaunch L) {
  const int vi = blockIdx.z;
  const RayView& v = L.view[vi];
  const int x = blockIdx.x * kTileX + (threadIdx.x % kTileX);
  const int y = blockIdx.y * kTileY + (threadIdx.x / kTileX);
  if (x >= rd.out_w || y >= rd.out_h) return;
  float3 nearp, farp, step;
  ray_endpoints(v, x, y, rd, nearp, farp, step);
  const float* buf = input + rd.in_batch_stride * v.in_batch;
  // EXACT sums all views into one gradient per input batch; REFERENCE gets a
  // per-view buffer that is normalised and added afterwards.
  float* gout = EXACT ? grad + rd.in_batch_stride * v.in_batch : grad;
  const size_t pix = size_t(y) * rd.out_w + x;
  Vec<C> acc = vload<C>(output + rd.out_view_stride * v.out_slot, pix);
  Vec<C> g = vload<C>(output_grad + rd.out_view_stride * v.out_slot, pix);
  const int S = rd.out_d;
  for (int i = S - 1; i >= 0; --i) {
    float3 p;
    if (EXACT) {
      float fi = (float)i;
      p = make_float3(fmaf(fi, step.x, nearp.x), fmaf(fi, step.y, nearp.y), fmaf(fi, step.z, nearp.z));
    } else {
      float fi = (float)(S - 1 - i);
      p = make_float3(fmaf(-fi, step.x, farp.x), fmaf(-fi, step.y, farp.y), fmaf(-fi, step.z, farp.z));
    }
    Vec<C> s = sample_ray<C, F_LINEAR, BM>(buf, rd.in, p);
    Vec<C> gs = blend_grad<C, BL, EXACT>(g, s, acc);
    scatter_sample<C, BM, EXACT>(gout, wsum, rd.in, p.x - 0.5f, p.y - 0.5f, p.z - 0.5f, gs);
  }
}

// total += buf / w where w > 0; buf and w are cleared for the next camera.
template <int C>
__global__ __launch_bounds__(256) void k_normalize_add(float* __restrict__ total, float* __restrict__ buf,
                                                       float* __restrict__ wsum, size_t nvox) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nvox) return;
  float w = wsum[i];
  if (w > 0.f) {
    float inv = 1.0f / w;
    Vec<C> b = vload<C>(buf, i);
    Vec<C> t = vload<C>(total, i);
#pragma unroll
    for (int c = 0; c < C; ++c) t.v[c] += b.v[c] * inv;
    vstore<C>(total, i, t);
  }
  vstore<C>(buf, i, vzero<C>());
  wsum[i] = 0.f;
}

// ------------------------------------------------ BORDER: walk + gather
// The BORDER backward without atomics: a walk stores every sample's gradient,
// a voxel-block gather sums them (the transpose of the trilinear scatter).
//   walk   : one lane per pixel, all lanes walk the whole ray in lock-step
//            (the blend recurrences see every sample, as the scatter walk does)
//            and store, for the samples of the pixel's clipped range, the
//            gradient vector that the scatter would spread, or a NaN mark for
//            a sample the scatter skips (EXACT: outside the grid; REFERENCE:
//            outside the band (-1, dims)). Layout [view][S][H][W][C] plus the
//            range (lo, hi) per pixel.
//   gather : one wave per block of voxels with a wave-private FP64 LDS
//            accumulator; per view, the candidate pixels of the block's
//            support box, each ray cut by the box, hat weights per sample.
//            REFERENCE also sums the weights and divides per view, as
//            kNormalize3DGradients does per camera.
template <int C, int BL, bool EXACT>
__global__ __launch_bounds__(256) void k_raymarch_bwd_walk(const float* __restrict__ input,
                                                           const float* __restrict__ output,
                                                           const float* __restrict__ output_grad,
                                                           float* __restrict__ gsamp, int2* __restrict__ ranges,
                                                           RayDims rd, RayLaunch L) {
  const int vi = blockIdx.z;
  const RayView& v = L.view[vi];
  const int x = blockIdx.x * kTileX + (threadIdx.x % kTileX);
  const int y = blockIdx.y * kTileY + (threadIdx.x / kTileX);
  if (x >= rd.out_w || y >= rd.out_h) return;
  float3 nearp, farp, step;
  ray_endpoints(v, x, y, rd, nearp, farp, step);
  int lo, hi;
  clip_ray(nearp, step, rd.in, rd.out_d, EXACT ? 0.f : 0.5f, lo, hi);
  const size_t plane = size_t(rd.out_w) * rd.out_h;
  const size_t pix = size_t(y) * rd.out_w + x;
  ranges[(size_t)vi * plane + pix] = make_int2(lo, hi);
  const float* buf = input + rd.in_batch_stride * v.in_batch;
  Vec<C> acc = vload<C>(output + rd.out_view_stride * v.out_slot, pix);
  Vec<C> g = vload<C>(output_grad + rd.out_view_stride * v.out_slot, pix);
  const int S = rd.out_d;
  float* gview = gsamp + (size_t)vi * S * plane * C;
  constexpr int U = kRayBatch<C>;
  auto position = [&](int i) {
    if (EXACT) {
      const float fi = (float)i;
      return make_float3(fmaf(fi, step.x, nearp.x), fmaf(fi, step.y, nearp.y), fmaf(fi, step.z, nearp.z));
    }
    const float fi = (float)(S - 1 - i);
    return make_float3(fmaf(-fi, step.x, farp.x), fmaf(-fi, step.y, farp.y), fmaf(-fi, step.z, farp.z));
  };
  for (int ib = S - 1; ib >= 0; ib -= U) {
    // the batch's samples ib, ib-1, ... are read together when it meets the
    // clipped range; outside the range the BORDER sample reads 0 (the blend
    // still runs)
    Vec<C> smps[U];
    if ((ib - U + 1 <= hi) & (ib >= lo)) {
      float px[U], py[U], pz[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float3 q = position(max(ib - u, 0));
        px[u] = q.x - 0.5f;
        py[u] = q.y - 0.5f;
        pz[u] = q.z - 0.5f;
      }
      read_linear_batch<C, B_BORDER, U>(buf, rd.in, px, py, pz, smps);
    } else {
#pragma unroll
      for (int u = 0; u < U; ++u) smps[u] = vzero<C>();
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
    const int i = ib - u;
    if (i < 0) break;
    const float3 p = position(i);
    const bool in_range = (lo <= i) & (i <= hi);
    const Vec<C> smp = in_range ? smps[u] : vzero<C>();
    const Vec<C> gs = blend_grad<C, BL, EXACT>(g, smp, acc);
    if (in_range) {
      const float px = p.x - 0.5f, py = p.y - 0.5f, pz = p.z - 0.5f;
      bool counted;
      if (EXACT) {
        const float qx = px + 0.5f, qy = py + 0.5f, qz = pz + 0.5f;
        counted = (0.f <= qx) & (qx < (float)rd.in.x) & (0.f <= qy) & (qy < (float)rd.in.y) & (0.f <= qz) &
                  (qz < (float)rd.in.z);
      } else {
        counted = (-1.f < px) & (px < (float)rd.in.x) & (-1.f < py) & (py < (float)rd.in.y) & (-1.f < pz) &
                  (pz < (float)rd.in.z);
      }
      Vec<C> out = gs;
      if (!counted) out.v[0] = __builtin_nanf("");
      vstore<C>(gview, (size_t)i * plane + pix, out);
    }
    }
  }
}

template <int C, int WX, int WY, int WZ, bool EXACT>
__global__ __launch_bounds__(256) void k_raymarch_bwd_gather(const float* __restrict__ gsamp,
                                                             const int2* __restrict__ ranges,
                                                             float* __restrict__ grad, RayDims rd, ProjLaunch PL) {
  constexpr int kVox = WX * WY * WZ;
  constexpr int kCh = EXACT ? C : C + 1;  // REFERENCE keeps the weight sum as the last channel
  constexpr int kOwn = kVox / 64;         // voxels per lane in the per-view normalisation
  static_assert(kVox % 64 == 0, "block must be a multiple of the wave");
  __shared__ double region[4][kVox * kCh];
  typedef __attribute__((address_space(3))) double lds_double;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  lds_double* acc = (lds_double*)region[wave];
  const Dims3 d = rd.in;
  const int nbx = (d.x + WX - 1) / WX, nby = (d.y + WY - 1) / WY, nbz = (d.z + WZ - 1) / WZ;
  const int blk = blockIdx.x * 4 + wave;
  const bool live = blk < nbx * nby * nbz;
  const int x0 = (blk % nbx) * WX, y0 = ((blk / nbx) % nby) * WY, z0 = (blk / (nbx * nby)) * WZ;
  for (int k = lane; k < kVox * kCh; k += 64) acc[k] = 0.0;
  float tot[kOwn][C];
#pragma unroll
  for (int j = 0; j < kOwn; ++j)
#pragma unroll
    for (int c = 0; c < C; ++c) tot[j][c] = 0.f;
  const size_t plane = size_t(rd.out_w) * rd.out_h;
  const float fx0 = (float)x0, fy0 = (float)y0, fz0 = (float)z0;
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
    const float* gview = gsamp + (size_t)vi * rd.out_d * plane * C;
    const int2* rview = ranges + (size_t)vi * plane;
    if (xs <= xe && ys <= ye) {
      const int nx = xe - xs + 1, cnt = nx * (ye - ys + 1);
      for (int k0 = 0; k0 < cnt; k0 += 64) {
        const int k = k0 + lane;
        if (k >= cnt) break;
        const int px = xs + k % nx, py = ys + k / nx;
        const size_t pix = size_t(py) * rd.out_w + px;
        const int2 range = rview[pix];
        const int lo = range.x, hi = range.y;
        if (lo > hi) continue;
        const float fpx = (float)px, fpy = (float)py;
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
        // slab ends as products: a zero step component keeps a ray inside its slab
        const float ax = (bx0 - nx_) * irx, bx = (bx1 - nx_) * irx;
        const float ay = (by0 - ny_) * iry, by = (by1 - ny_) * iry;
        const float az = (bz0 - nz_) * irz, bz = (bz1 - nz_) * irz;
        const float tlo = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
        const float thi = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
        const int i0 = max(lo, (int)ceilf(fmaxf(tlo, -1.f) - 1e-3f));
        const int i1 = min(hi, (int)floorf(fminf(thi, 1e9f) + 1e-3f));
        if (i0 > i1) continue;
        const float fi0 = (float)i0;
        float qx = fmaf(fi0, sx, nx_) - (fx0 + 0.5f), qy = fmaf(fi0, sy, ny_) - (fy0 + 0.5f),
              qz = fmaf(fi0, sz, nz_) - (fz0 + 0.5f);
        const float* gp = gview + ((size_t)i0 * plane + pix) * C;
        for (int i = i0; i <= i1; ++i) {
          const Vec<C> gs = vload<C>(gp, 0);
          gp += plane * C;
          if (gs.v[0] == gs.v[0]) {  // NaN mark: a sample the scatter skips
            const float flx = floorf(qx), fly = floorf(qy), flz = floorf(qz);
            const float tx = qx - flx, ty = qy - fly, tz = qz - flz;
            const int lx = (int)flx, ly = (int)fly, lz = (int)flz;
            const float wx[2] = {1.f - tx, tx}, wy[2] = {1.f - ty, ty}, wz[2] = {1.f - tz, tz};
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
                  const float w = wx[kx] * wyz;
                  lds_double* a = acc + ((cz * WY + cy) * WX + cx) * kCh;
#pragma unroll
                  for (int c = 0; c < C; ++c)
                    __hip_atomic_fetch_add(a + c, (double)(gs.v[c] * w), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                  if (!EXACT)
                    __hip_atomic_fetch_add(a + C, (double)w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
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
    if (!EXACT) {
      // per-view normalisation by the weight sum (where positive), then clear
#pragma unroll
      for (int j = 0; j < kOwn; ++j) {
        lds_double* a = acc + (lane + 64 * j) * kCh;
        const float w = (float)a[C];
        if (w > 0.f) {
          const float inv = 1.0f / w;
#pragma unroll
          for (int c = 0; c < C; ++c) tot[j][c] += (float)a[c] * inv;
        }
#pragma unroll
        for (int c = 0; c < kCh; ++c) a[c] = 0.0;
      }
    }
  }
  __syncthreads();
  if (!live) return;
#pragma unroll
  for (int j = 0; j < kOwn; ++j) {
    const int k = lane + 64 * j;
    const int x = x0 + k % WX, y = y0 + (k / WX) % WY, z = z0 + k / (WX * WY);
    if (x < d.x && y < d.y && z < d.z) {
      Vec<C> r;
#pragma unroll
      for (int c = 0; c < C; ++c) r.v[c] = EXACT ? (float)acc[k * kCh + c] : tot[j][c];
      // the caller cleared grad; chunks of views add up
      Vec<C> o = vload<C>(grad, flat3(x, y, z, d));
#pragma unroll
      for (int c = 0; c < C; ++c) o.v[c] += r.v[c];
      vstore<C>(grad, flat3(x, y, z, d), o);
    }
  }
}

// ------------------------------------------------------------- dispatch
template <int C, int FM, int BM>
static hipError_t launch_fwd_bl(int bl, dim3 grid, hipStream_t s, const float* in, float* out, const RayDims& rd,
                                const RayLaunch& L) {
  switch (bl) {
    case BL_BEER_LAMBERT: hipLaunchKernelGGL((k_raymarch_fwd<C, FM, BM, BL_BEER_LAMBERT>), grid, dim3(256), 0, s, in, out, rd, L); break;
    case BL_ALPHA: hipLaunchKernelGGL((k_raymarch_fwd<C, FM, BM, BL_ALPHA>), grid, dim3(256), 0, s, in, out, rd, L); break;
    case BL_ALPHA_ADDITIVE: hipLaunchKernelGGL((k_raymarch_fwd<C, FM, BM, BL_ALPHA_ADDITIVE>), grid, dim3(256), 0, s, in, out, rd, L); break;
    default: hipLaunchKernelGGL((k_raymarch_fwd<C, FM, BM, BL_ADDITIVE>), grid, dim3(256), 0, s, in, out, rd, L); break;
  }
  return hipGetLastError();
}
template <int C, int FM>
static hipError_t launch_fwd_bm(int bm, int bl, dim3 grid, hipStream_t s, const float* in, float* out,
                                const RayDims& rd, const RayLaunch& L) {
  switch (bm) {
    case B_BORDER: return launch_fwd_bl<C, FM, B_BORDER>(bl, grid, s, in, out, rd, L);
    case B_CLAMP: return launch_fwd_bl<C, FM, B_CLAMP>(bl, grid, s, in, out, rd, L);
    default: return launch_fwd_bl<C, FM, B_WRAP>(bl, grid, s, in, out, rd, L);
  }
}
template <int C>
static hipError_t launch_fwd_fm(int fm, int bm, int bl, dim3 grid, hipStream_t s, const float* in, float* out,
                                const RayDims& rd, const RayLaunch& L) {
  switch (fm) {
    case F_LINEAR: return launch_fwd_bm<C, F_LINEAR>(bm, bl, grid, s, in, out, rd, L);
    case F_NEAREST: return launch_fwd_bm<C, F_NEAREST>(bm, bl, grid, s, in, out, rd, L);
    case F_MIN: return launch_fwd_bm<C, F_MIN>(bm, bl, grid, s, in, out, rd, L);
    default: return launch_fwd_bm<C, F_MAX>(bm, bl, grid, s, in, out, rd, L);
  }
}

hipError_t raymarch_fwd_launch(int C, int fm, int bm, int bl, hipStream_t s, const float* in, float* out,
                               const RayDims& rd, const RayLaunch& L) {
  dim3 grid((rd.out_w + kTileX - 1) / kTileX, (rd.out_h + kTileY - 1) / kTileY, L.num_views);
  switch (C) {
    case 1: return launch_fwd_fm<1>(fm, bm, bl, grid, s, in, out, rd, L);
    case 2: return launch_fwd_fm<2>(fm, bm, bl, grid, s, in, out, rd, L);
    default: return launch_fwd_fm<4>(fm, bm, bl, grid, s, in, out, rd, L);
  }
}

template <int C, int BM, bool EX>
static hipError_t launch_bwd_bl(int bl, dim3 grid, hipStream_t s, const float* in, const float* o, const float* og,
                                float* g, float* w, const RayDims& rd, const RayLaunch& L) {
  switch (bl) {
    case BL_BEER_LAMBERT: hipLaunchKernelGGL((k_raymarch_bwd<C, BM, BL_BEER_LAMBERT, EX>), grid, dim3(256), 0, s, in, o, og, g, w, rd, L); break;
    case BL_ALPHA: hipLaunchKernelGGL((k_raymarch_bwd<C, BM, BL_ALPHA, EX>), grid, dim3(256), 0, s, in, o, og, g, w, rd, L); break;
    case BL_ALPHA_ADDITIVE: hipLaunchKernelGGL((k_raymarch_bwd<C, BM, BL_ALPHA_ADDITIVE, EX>), grid, dim3(256), 0, s, in, o, og, g, w, rd, L); break;
    default: hipLaunchKernelGGL((k_raymarch_bwd<C, BM, BL_ADDITIVE, EX>), grid, dim3(256), 0, s, in, o, og, g, w, rd, L); break;
  }
  return hipGetLastError();
}
template <int C, bool EX>
static hipError_t launch_bwd_bm(int bm, int bl, dim3 grid, hipStream_t s, const float* in, const float* o,
                                const float* og, float* g, float* w, const RayDims& rd, const RayLaunch& L) {
  switch (bm) {
    case B_BORDER: return launch_bwd_bl<C, B_BORDER, EX>(bl, grid, s, in, o, og, g, w, rd, L);
    case B_CLAMP: return launch_bwd_bl<C, B_CLAMP, EX>(bl, grid, s, in, o, og, g, w, rd, L);
    default: return launch_bwd_bl<C, B_WRAP, EX>(bl, grid, s, in, o, og, g, w, rd, L);
  }
}
template <int C>
static hipError_t launch_bwd_c(bool exact, int bm, int bl, dim3 grid, hipStream_t s, const float* in, const float* o,
                               const float* og, float* g, float* w, const RayDims& rd, const RayLaunch& L) {
  return exact ? launch_bwd_bm<C, true>(bm, bl, grid, s, in, o, og, g, w, rd, L)
               : launch_bwd_bm<C, false>(bm, bl, grid, s, in, o, og, g, w, rd, L);
}

hipError_t raymarch_bwd_launch(int C, bool exact, int bm, int bl, hipStream_t s, const float* in, const float* o,
                               const float* og, float* g, float* w, const RayDims& rd, const RayLaunch& L) {
  dim3 grid((rd.out_w + kTileX - 1) / kTileX, (rd.out_h + kTileY - 1) / kTileY, L.num_views);
  switch (C) {
    case 1: return launch_bwd_c<1>(exact, bm, bl, grid, s, in, o, og, g, w, rd, L);
    case 2: return launch_bwd_c<2>(exact, bm, bl, grid, s, in, o, og, g, w, rd, L);
    default: return launch_bwd_c<4>(exact, bm, bl, grid, s, in, o, og, g, w, rd, L);
  }
}

hipError_t normalize_add_launch(int C, hipStream_t s, float* total, float* buf, float* wsum, size_t nvox) {
  dim3 grid((unsigned)((nvox + 255) / 256));
  switch (C) {
    case 1: hipLaunchKernelGGL((k_normalize_add<1>), grid, dim3(256), 0, s, total, buf, wsum, nvox); break;
    case 2: hipLaunchKernelGGL((k_normalize_add<2>), grid, dim3(256), 0, s, total, buf, wsum, nvox); break;
    default: hipLaunchKernelGGL((k_normalize_add<4>), grid, dim3(256), 0, s, total, buf, wsum, nvox); break;
  }
  return hipGetLastError();
}

// BORDER walk for one chunk of views (launch-local view index vi)
template <int C, bool EX>
static hipError_t launch_walk_bl(int bl, dim3 grid, hipStream_t s, const float* in, const float* o, const float* og,
                                 float* gs, int2* rg, const RayDims& rd, const RayLaunch& L) {
  switch (bl) {
    case BL_BEER_LAMBERT: hipLaunchKernelGGL((k_raymarch_bwd_walk<C, BL_BEER_LAMBERT, EX>), grid, dim3(256), 0, s, in, o, og, gs, rg, rd, L); break;
    case BL_ALPHA: hipLaunchKernelGGL((k_raymarch_bwd_walk<C, BL_ALPHA, EX>), grid, dim3(256), 0, s, in, o, og, gs, rg, rd, L); break;
    case BL_ALPHA_ADDITIVE: hipLaunchKernelGGL((k_raymarch_bwd_walk<C, BL_ALPHA_ADDITIVE, EX>), grid, dim3(256), 0, s, in, o, og, gs, rg, rd, L); break;
    default: hipLaunchKernelGGL((k_raymarch_bwd_walk<C, BL_ADDITIVE, EX>), grid, dim3(256), 0, s, in, o, og, gs, rg, rd, L); break;
  }
  return hipGetLastError();
}

hipError_t raymarch_bwd_walk_launch(int C, bool exact, int bl, hipStream_t s, const float* in, const float* o,
                                    const float* og, float* gsamp, int2* ranges, const RayDims& rd,
                                    const RayLaunch& L) {
  dim3 grid((rd.out_w + kTileX - 1) / kTileX, (rd.out_h + kTileY - 1) / kTileY, L.num_views);
  KernelTimer t(s, "raymarch_bwd_walk");
  switch (C) {
    case 1: return exact ? launch_walk_bl<1, true>(bl, grid, s, in, o, og, gsamp, ranges, rd, L)
                         : launch_walk_bl<1, false>(bl, grid, s, in, o, og, gsamp, ranges, rd, L);
    case 2: return exact ? launch_walk_bl<2, true>(bl, grid, s, in, o, og, gsamp, ranges, rd, L)
                         : launch_walk_bl<2, false>(bl, grid, s, in, o, og, gsamp, ranges, rd, L);
    default: return exact ? launch_walk_bl<4, true>(bl, grid, s, in, o, og, gsamp, ranges, rd, L)
                          : launch_walk_bl<4, false>(bl, grid, s, in, o, og, gsamp, ranges, rd, L);
  }
}

// voxel blocks per channel count: 512, 256, 128 voxels (FP64 LDS per wave <= 5 KB)
template <int C, int WX, int WY, int WZ>
static hipError_t launch_gather(bool exact, hipStream_t s, const float* gsamp, const int2* ranges, float* grad,
                                const RayDims& rd, const ProjLaunch& PL) {
  const int nb = ((rd.in.x + WX - 1) / WX) * ((rd.in.y + WY - 1) / WY) * ((rd.in.z + WZ - 1) / WZ);
  const dim3 grid((nb + 3) / 4);
  if (exact) hipLaunchKernelGGL((k_raymarch_bwd_gather<C, WX, WY, WZ, true>), grid, dim3(256), 0, s, gsamp, ranges, grad, rd, PL);
  else hipLaunchKernelGGL((k_raymarch_bwd_gather<C, WX, WY, WZ, false>), grid, dim3(256), 0, s, gsamp, ranges, grad, rd, PL);
  return hipGetLastError();
}

hipError_t raymarch_bwd_gather_launch(int C, bool exact, hipStream_t s, const float* gsamp, const int2* ranges,
                                      float* grad, const RayDims& rd, const ProjLaunch& PL) {
  KernelTimer t(s, "raymarch_bwd_gather");
  switch (C) {
    case 1: return launch_gather<1, 16, 16, 2>(exact, s, gsamp, ranges, grad, rd, PL);
    case 2: return launch_gather<2, 16, 8, 2>(exact, s, gsamp, ranges, grad, rd, PL);
    default: return launch_gather<4, 8, 8, 2>(exact, s, gsamp, ranges, grad, rd, PL);
  }
}

}  // namespace gft
