// ReduceGridBlend / ReduceGridBlendGrad on gfx950: depth-prefix blending of a
// materialised frustum grid [N,D,H,W,C] along D, and its reverse-walk
// gradient (phitest/render/cuda/src/reduce_blend.cu.cc:34-227,
// reduce_blend.cc:23-242).
//
// One thread per (batch, y, x) column with a serial loop over depth: at every
// depth step a wave reads 64 consecutive voxels, so the walk streams the grid
// at HBM rate. No atomics; the gradient is exact up to the reference's ALPHA
// quirk, which the REFERENCE flavour keeps.
#include "gft_host.hpp"

namespace gft {

template <int C, int BL, bool KEEP>
__global__ __launch_bounds__(256) void k_reduce_blend(const float* __restrict__ in, float* __restrict__ out, int D,
                                                      int H, int W, int N) {
  size_t plane = (size_t)H * W;
  size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= plane * N) return;
  size_t n = col / plane, p = col % plane;
  const float* src = in + n * (size_t)D * plane * C;
  Vec<C> acc = vzero<C>();
  for (int z = 0; z < D; ++z) {
    size_t v = (size_t)z * plane + p;
    Vec<C> c = vload<C>(src, v);
    blend<C, BL>(acc, c);
    if (KEEP) vstore<C>(out + n * (size_t)D * plane * C, v, acc);
  }
  if (!KEEP) vstore<C>(out + n * plane * C, p, acc);
}

template <int C, int BL, bool KEEP, bool EXACT>
__global__ __launch_bounds__(256) void k_reduce_blend_grad(const float* __restrict__ in, const float* __restrict__ out,
                                                           const float* __restrict__ og, float* __restrict__ gin,
                                                           int D, int H, int W, int N) {
  size_t plane = (size_t)H * W;
  size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= plane * N) return;
  size_t n = col / plane, p = col % plane;
  const size_t base = n * (size_t)D * plane * C;
  Vec<C> g = vzero<C>(), acc = vzero<C>();
  if (!KEEP) {
    g = vload<C>(og + n * plane * C, p);
    acc = vload<C>(out + n * plane * C, p);
  }
  for (int z = D - 1; z >= 0; --z) {
    size_t v = (size_t)z * plane + p;
    if (KEEP) {
      Vec<C> gz = vload<C>(og + base, v);
#pragma unroll
      for (int i = 0; i < C; ++i) g.v[i] += gz.v[i];
      acc = vload<C>(out + base, v);
    }
    Vec<C> c = vload<C>(in + base, v);
    Vec<C> gc = blend_grad<C, BL, EXACT>(g, c, acc);
    vstore<C>(gin + base, v, gc);
  }
}

#define GFT_BL_SWITCH(bl, ...)                                                         \
  switch (bl) {                                                                        \
    case BL_BEER_LAMBERT: { constexpr int BB = BL_BEER_LAMBERT; __VA_ARGS__; } break;  \
    case BL_ALPHA: { constexpr int BB = BL_ALPHA; __VA_ARGS__; } break;                \
    case BL_ALPHA_ADDITIVE: { constexpr int BB = BL_ALPHA_ADDITIVE; __VA_ARGS__; } break; \
    default: { constexpr int BB = BL_ADDITIVE; __VA_ARGS__; } break;                   \
  }
#define GFT_C_SWITCH(C, ...)                                  \
  switch (C) {                                                \
    case 1: { constexpr int CC = 1; __VA_ARGS__; } break;     \
    case 2: { constexpr int CC = 2; __VA_ARGS__; } break;     \
    default: { constexpr int CC = 4; __VA_ARGS__; } break;    \
  }

static hipError_t reduce_blend_launch(int C, int bl, bool keep, hipStream_t s, const float* in, float* out, int N,
                                      int D, int H, int W) {
  size_t cols = (size_t)N * H * W;
  dim3 grid((unsigned)((cols + 255) / 256));
  GFT_C_SWITCH(C, GFT_BL_SWITCH(bl, {
    if (keep) hipLaunchKernelGGL((k_reduce_blend<CC, BB, true>), grid, dim3(256), 0, s, in, out, D, H, W, N);
    else hipLaunchKernelGGL((k_reduce_blend<CC, BB, false>), grid, dim3(256), 0, s, in, out, D, H, W, N);
  }));
  return hipGetLastError();
}

static hipError_t reduce_blend_grad_launch(int C, int bl, bool keep, bool exact, hipStream_t s, const float* in,
                                           const float* out, const float* og, float* gin, int N, int D, int H,
                                           int W) {
  size_t cols = (size_t)N * H * W;
  dim3 grid((unsigned)((cols + 255) / 256));
  GFT_C_SWITCH(C, GFT_BL_SWITCH(bl, {
    if (keep && exact)
      hipLaunchKernelGGL((k_reduce_blend_grad<CC, BB, true, true>), grid, dim3(256), 0, s, in, out, og, gin, D, H, W, N);
    else if (keep)
      hipLaunchKernelGGL((k_reduce_blend_grad<CC, BB, true, false>), grid, dim3(256), 0, s, in, out, og, gin, D, H, W,
                         N);
    else if (exact)
      hipLaunchKernelGGL((k_reduce_blend_grad<CC, BB, false, true>), grid, dim3(256), 0, s, in, out, og, gin, D, H, W,
                         N);
    else
      hipLaunchKernelGGL((k_reduce_blend_grad<CC, BB, false, false>), grid, dim3(256), 0, s, in, out, og, gin, D, H,
                         W, N);
  }));
  return hipGetLastError();
}

static int check_blend_input(const int64_t* shape, int32_t blending_mode) {
  for (int i = 0; i < 5; ++i)
    if (shape[i] < 1)
      return fail(GFT_ERR_INVALID_ARGUMENT, "Invalid input shape: dimension %d is %lld", i, (long long)shape[i]);
  if (shape[4] > 4) return fail(GFT_ERR_INVALID_ARGUMENT, "Invalid input shape: channel %lld > 4", (long long)shape[4]);
  if (shape[4] == 3) return fail(GFT_ERR_UNIMPLEMENTED, "Only 1,2 and 4 Channel supported.");
  if (blending_mode < 0 || blending_mode > 3) return fail(GFT_ERR_INVALID_ARGUMENT, "invalid blending mode");
  return GFT_OK;
}

}  // namespace gft

using namespace gft;

extern "C" {

int gft_reduce_grid_blend_fwd(void* stream, const float* input, const int64_t shape[5], int32_t blending_mode,
                              int32_t keep_dims, float* output) {
  int rc = check_blend_input(shape, blending_mode);
  if (rc) return rc;
  if (!input || !output) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  return check_hip(reduce_blend_launch((int)shape[4], blending_mode, keep_dims != 0, (hipStream_t)stream, input,
                                       output, (int)shape[0], (int)shape[1], (int)shape[2], (int)shape[3]),
                   "reduce blend launch");
}

int gft_reduce_grid_blend_bwd(void* stream, const float* input, const int64_t shape[5], const float* output,
                              const float* output_grad, int32_t blending_mode, int32_t keep_dims, int32_t grad_mode,
                              float* input_grad) {
  int rc = check_blend_input(shape, blending_mode);
  if (rc) return rc;
  if (grad_mode != GFT_GRAD_REFERENCE && grad_mode != GFT_GRAD_EXACT)
    return fail(GFT_ERR_INVALID_ARGUMENT, "invalid grad_mode");
  // The reference's gradient launcher has no ALPHA_ADDITIVE case and raises
  // "Unkown blend_mode" (reduce_blend.cu.cc:194-211).
  if (grad_mode == GFT_GRAD_REFERENCE && blending_mode == GFT_ALPHA_ADDITIVE)
    return fail(GFT_ERR_UNIMPLEMENTED, "Unkown blend_mode");
  if (!input || !output || !output_grad || !input_grad) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  return check_hip(reduce_blend_grad_launch((int)shape[4], blending_mode, keep_dims != 0,
                                            grad_mode == GFT_GRAD_EXACT, (hipStream_t)stream, input, output,
                                            output_grad, input_grad, (int)shape[0], (int)shape[1], (int)shape[2],
                                            (int)shape[3]),
                   "reduce blend gradient launch");
}

}  // extern "C"
