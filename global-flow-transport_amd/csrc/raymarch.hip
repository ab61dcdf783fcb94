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
// optical depth (no per-sample state), and the per-sample gradient is
// scattered to the 8 trilinear corners with float atomics.
//   EXACT     : true adjoint, all views summed into one gradient, sample
//               positions identical to the forward.
//   REFERENCE : per-camera normalisation by the accumulated trilinear weight
//               (kNormalize3DGradients, raymarch_grid.cu.cc:319-341), the
//               BORDER scatter band (-1, dims) and the far-to-near positions.
#include "gft_device.hpp"

namespace gft {

constexpr int kTileX = 16;
constexpr int kTileY = 16;

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
  for (int i = 0; i < rd.out_d; ++i) {
    float fi = (float)i;
    float3 p = make_float3(fmaf(fi, step.x, nearp.x), fmaf(fi, step.y, nearp.y), fmaf(fi, step.z, nearp.z));
    Vec<C> s = sample_ray<C, FM, BM>(buf, rd.in, p);
    blend<C, BL>(acc, s);
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

}  // namespace gft
