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
// Backward: either from an image gradient, or with the L2 loss
// loss = scale * sum (rgb - target)^2 over all views and pixels folded in,
// in which case each lane computes g_rgb from its own pixel. The lane walks the ray from far to near
// recovering D_{i-1} = max(D_i - sigma_i, 0) (no per-sample storage), forms
// the density gradient of the sample
//   g_d = s * (exp(-D_i) * G + g_D),  G = sum_c (c*i)_c g_rgb_c,
//   g_D <- g_D - exp(-D_i) * sigma_i * G
// and scatters it to the 8 trilinear corners with float atomics (exact
// adjoint, BORDER samples outside the grid carry no gradient).
#include "gft_device.hpp"

namespace gft {

constexpr int kRTileX = 16;
constexpr int kRTileY = 16;


template <int BM>
__global__ __launch_bounds__(256) void k_render_density_fwd(const float* __restrict__ dens, float* __restrict__ img,
                                                            RayDims rd, RayLaunch L, RenderParams rp) {
  const RayView& v = L.view[blockIdx.z];
  const int x = blockIdx.x * kRTileX + (threadIdx.x % kRTileX);
  const int y = blockIdx.y * kRTileY + (threadIdx.x / kRTileX);
  if (x >= rd.out_w || y >= rd.out_h) return;
  float dinvx = 1.f / (float)rd.out_w, dinvy = 1.f / (float)rd.out_h, dinvz = 1.f / (float)rd.out_d;
  float zmax = (float)max(rd.out_d - 1, 1);
  float3 nearp = idx_to_os(v, x + 0.5f, y + 0.5f, 0.5f, dinvx, dinvy, dinvz);
  float3 farp = idx_to_os(v, x + 0.5f, y + 0.5f, zmax + 0.5f, dinvx, dinvy, dinvz);
  float iz = 1.f / zmax;
  float3 step = make_float3((farp.x - nearp.x) * iz, (farp.y - nearp.y) * iz, (farp.z - nearp.z) * iz);
  const float* buf = dens + rd.in_batch_stride * v.in_batch;
  float D = 0.f, E = 0.f;
  for (int i = 0; i < rd.out_d; ++i) {
    float fi = (float)i;
    float px = fmaf(fi, step.x, nearp.x) - 0.5f, py = fmaf(fi, step.y, nearp.y) - 0.5f,
          pz = fmaf(fi, step.z, nearp.z) - 0.5f;
    float sigma = rp.step * read_linear<1, BM>(buf, rd.in, px, py, pz).v[0];
    D = fmaxf(D + sigma, 0.f);
    E = fmaf(expf(-D), sigma, E);
  }
  float4 o = make_float4(rp.light[0] * E, rp.light[1] * E, rp.light[2] * E, D);
  reinterpret_cast<float4*>(img + rd.out_view_stride * v.out_slot)[size_t(y) * rd.out_w + x] = o;
}

// L2 = true: `target` holds rgb targets and the loss is accumulated;
// L2 = false: `target` holds the image gradient (float4 per pixel).
template <int BM, bool L2>
__global__ __launch_bounds__(256) void k_render_density_bwd(const float* __restrict__ dens,
                                                            const float* __restrict__ img,
                                                            const float* __restrict__ target,
                                                            float* __restrict__ grad, float* __restrict__ loss,
                                                            RayDims rd, RayLaunch L, RenderParams rp) {
  __shared__ float red[256 / 64];
  const RayView& v = L.view[blockIdx.z];
  const int x = blockIdx.x * kRTileX + (threadIdx.x % kRTileX);
  const int y = blockIdx.y * kRTileY + (threadIdx.x / kRTileX);
  const bool active = x < rd.out_w && y < rd.out_h;
  float sq = 0.f;
  if (active) {
    const size_t pix = size_t(y) * rd.out_w + x;
    float4 o = reinterpret_cast<const float4*>(img + rd.out_view_stride * v.out_slot)[pix];
    float G, gD;
    if (L2) {
      const float* t = target + (size_t)v.out_slot * rd.out_w * rd.out_h * 3 + pix * 3;
      float e0 = o.x - t[0], e1 = o.y - t[1], e2 = o.z - t[2];
      sq = e0 * e0 + e1 * e1 + e2 * e2;
      G = 2.f * rp.loss_scale * (rp.light[0] * e0 + rp.light[1] * e1 + rp.light[2] * e2);
      gD = 0.f;  // no loss term on the optical depth channel
    } else {
      float4 go = reinterpret_cast<const float4*>(target + rd.out_view_stride * v.out_slot)[pix];
      G = rp.light[0] * go.x + rp.light[1] * go.y + rp.light[2] * go.z;
      gD = go.w;
    }
    float dinvx = 1.f / (float)rd.out_w, dinvy = 1.f / (float)rd.out_h, dinvz = 1.f / (float)rd.out_d;
    float zmax = (float)max(rd.out_d - 1, 1);
    float3 nearp = idx_to_os(v, x + 0.5f, y + 0.5f, 0.5f, dinvx, dinvy, dinvz);
    float3 farp = idx_to_os(v, x + 0.5f, y + 0.5f, zmax + 0.5f, dinvx, dinvy, dinvz);
    float iz = 1.f / zmax;
    float3 step = make_float3((farp.x - nearp.x) * iz, (farp.y - nearp.y) * iz, (farp.z - nearp.z) * iz);
    const float* buf = dens + rd.in_batch_stride * v.in_batch;
    float* gout = grad + rd.in_batch_stride * v.in_batch;
    float D = o.w;
    if (G != 0.f || gD != 0.f) {
      for (int i = rd.out_d - 1; i >= 0; --i) {
        float fi = (float)i;
        float px = fmaf(fi, step.x, nearp.x) - 0.5f, py = fmaf(fi, step.y, nearp.y) - 0.5f,
              pz = fmaf(fi, step.z, nearp.z) - 0.5f;
        Taps tp = make_taps<BM>(px, py, pz, rd.in);
        // a BORDER sample outside the grid read 0: D is unchanged and the
        // sample does not depend on the density
        if (BM == B_BORDER && !tp.inside) continue;
        float c[2][2][2];
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) {
              float val = buf[flat3(tp.ix[ii], tp.iy[j], tp.iz[k], rd.in)];
              c[k][j][ii] = (BM == B_BORDER && !(tp.vx[ii] & tp.vy[j] & tp.vz[k])) ? 0.f : val;
            }
        float v00 = lerpf(c[0][0][0], c[0][0][1], tp.wx), v01 = lerpf(c[0][1][0], c[0][1][1], tp.wx);
        float v10 = lerpf(c[1][0][0], c[1][0][1], tp.wx), v11 = lerpf(c[1][1][0], c[1][1][1], tp.wx);
        float sigma = rp.step * lerpf(lerpf(v00, v01, tp.wy), lerpf(v10, v11, tp.wy), tp.wz);
        float e = expf(-D);
        D = fmaxf(D - sigma, 0.f);
        float gd_sample = rp.step * (e * G + (gD - e * sigma * G));
        gD = gD - e * sigma * G;
        const float wx[2] = {1.f - tp.wx, tp.wx}, wy[2] = {1.f - tp.wy, tp.wy}, wz[2] = {1.f - tp.wz, tp.wz};
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int ii = 0; ii < 2; ++ii) {
              if (BM == B_BORDER && !(tp.vx[ii] & tp.vy[j] & tp.vz[k])) continue;
              float w = wx[ii] * wy[j] * wz[k];
              if (w != 0.f) atomic_addf(gout + flat3(tp.ix[ii], tp.iy[j], tp.iz[k], rd.in), gd_sample * w);
            }
      }
    }
  }
  if (!L2) return;
  // block reduction of the squared error, one atomic per block
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int w = 0; w < 256 / 64; ++w) s += red[w];
    atomic_addf(loss, s * rp.loss_scale);
  }
}

hipError_t render_density_fwd_launch(int bm, hipStream_t s, const float* dens, float* img, const RayDims& rd,
                                     const RayLaunch& L, const RenderParams& rp) {
  dim3 grid((rd.out_w + kRTileX - 1) / kRTileX, (rd.out_h + kRTileY - 1) / kRTileY, L.num_views);
  switch (bm) {
    case B_BORDER: hipLaunchKernelGGL(k_render_density_fwd<B_BORDER>, grid, dim3(256), 0, s, dens, img, rd, L, rp); break;
    case B_CLAMP: hipLaunchKernelGGL(k_render_density_fwd<B_CLAMP>, grid, dim3(256), 0, s, dens, img, rd, L, rp); break;
    default: hipLaunchKernelGGL(k_render_density_fwd<B_WRAP>, grid, dim3(256), 0, s, dens, img, rd, L, rp); break;
  }
  return hipGetLastError();
}

template <bool L2>
static void launch_bwd(int bm, dim3 grid, hipStream_t s, const float* dens, const float* img, const float* target,
                       float* grad, float* loss, const RayDims& rd, const RayLaunch& L, const RenderParams& rp) {
  switch (bm) {
    case B_BORDER: hipLaunchKernelGGL((k_render_density_bwd<B_BORDER, L2>), grid, dim3(256), 0, s, dens, img, target, grad, loss, rd, L, rp); break;
    case B_CLAMP: hipLaunchKernelGGL((k_render_density_bwd<B_CLAMP, L2>), grid, dim3(256), 0, s, dens, img, target, grad, loss, rd, L, rp); break;
    default: hipLaunchKernelGGL((k_render_density_bwd<B_WRAP, L2>), grid, dim3(256), 0, s, dens, img, target, grad, loss, rd, L, rp); break;
  }
}

hipError_t render_density_bwd_launch(int bm, bool l2, hipStream_t s, const float* dens, const float* img,
                                     const float* target, float* grad, float* loss, const RayDims& rd,
                                     const RayLaunch& L, const RenderParams& rp) {
  dim3 grid((rd.out_w + kRTileX - 1) / kRTileX, (rd.out_h + kRTileY - 1) / kRTileY, L.num_views);
  if (l2) launch_bwd<true>(bm, grid, s, dens, img, target, grad, loss, rd, L, rp);
  else launch_bwd<false>(bm, grid, s, dens, img, target, grad, loss, rd, L, rp);
  return hipGetLastError();
}

}  // namespace gft
