// Loss and optimizer epilogue of the reconstruction step on gfx950: the
// divergence loss of a staggered velocity with its gradient, and one Adam
// update over several parameter tensors in a single launch.
//
// Divergence (VelocityGrid.divergence, data_structures.py:929-936, with the
// SE loss of loss_vel_divergence, optimization.py:824-838):
//   div(c) = wx (x[i+1] - x[i]) + wy (y[j+1] - y[j]) + wz (z[k+1] - z[k])
//   loss   = scale * sum_c div(c)^2
// A cell pass writes div and per-block loss partials; a face pass gathers
// each face's gradient 2 scale w (div(left cell) - div(right cell)) from the
// two cells it separates (no atomics); one block sums the partials in a fixed
// order, so the loss and the gradients are deterministic.
//
// Adam follows the update that tf.train.AdamOptimizer documents (the
// reference's optimizer, reconstruct_sequence.py:1899-1905):
//   lr_t = lr sqrt(1 - b2^t) / (1 - b1^t)   (formed on the host)
//   m = b1 m + (1 - b1) g;  v = b2 v + (1 - b2) g^2
//   p = p - lr_t m / (sqrt(v) + eps)
// One thread per 4 consecutive elements of one tensor; the launch walks up to
// kMaxAdamTensors tensors laid end to end.
#include <cmath>

#include "gft_host.hpp"

namespace gft {

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// ------------------------------------------------------------- divergence
struct DivDims {
  int W, H, D;  // centred cells
};

__global__ __launch_bounds__(256) void k_divergence(const float* __restrict__ x, const float* __restrict__ y,
                                                    const float* __restrict__ z, DivDims d, float wx, float wy,
                                                    float wz, float* __restrict__ div, float* __restrict__ partials) {
  const size_t n = (size_t)d.W * d.H * d.D;
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float sq = 0.f;
  if (c < n) {
    const int i = (int)(c % d.W);
    const size_t r = c / d.W;
    const int j = (int)(r % d.H), k = (int)(r / d.H);
    const size_t xi = ((size_t)k * d.H + j) * (d.W + 1) + i;
    const size_t yi = ((size_t)k * (d.H + 1) + j) * d.W + i;
    const size_t zi = ((size_t)k * d.H + j) * d.W + i;
    const float dv = (x[xi + 1] - x[xi]) * wx + (y[yi + d.W] - y[yi]) * wy + (z[zi + (size_t)d.W * d.H] - z[zi]) * wz;
    div[c] = dv;
    sq = dv * dv;
  }
  __shared__ float red[256 / 64];
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sq;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// blockIdx.y: 0 = x faces [D,H,W+1], 1 = y faces [D,H+1,W], 2 = z faces [D+1,H,W]
__global__ __launch_bounds__(256) void k_divergence_grad(const float* __restrict__ div, DivDims d, float gx_scale,
                                                         float gy_scale, float gz_scale, int accumulate,
                                                         float* __restrict__ gx, float* __restrict__ gy,
                                                         float* __restrict__ gz) {
  const int a = blockIdx.y;
  const int FW = d.W + (a == 0), FH = d.H + (a == 1), FD = d.D + (a == 2);
  const size_t n = (size_t)FW * FH * FD;
  const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const int i = (int)(f % FW);
  const size_t r = f / FW;
  const int j = (int)(r % FH), k = (int)(r / FH);
  // the face separates cell (i,j,k) - e_a (its "left" cell) and cell (i,j,k)
  const int pos = a == 0 ? i : (a == 1 ? j : k);
  const int ext = a == 0 ? d.W : (a == 1 ? d.H : d.D);
  const size_t stride = a == 0 ? 1 : (a == 1 ? (size_t)d.W : (size_t)d.W * d.H);
  const size_t cell = ((size_t)min(k, d.D - 1) * d.H + min(j, d.H - 1)) * d.W + min(i, d.W - 1);
  float left = 0.f, right = 0.f;
  if (pos >= 1) left = div[pos < ext ? cell - stride : cell];
  if (pos < ext) right = div[cell];
  const float sc = a == 0 ? gx_scale : (a == 1 ? gy_scale : gz_scale);
  float* g = a == 0 ? gx : (a == 1 ? gy : gz);
  const float v = sc * (left - right);
  g[f] = accumulate ? g[f] + v : v;
}

__global__ __launch_bounds__(256) void k_partials_sum(const float* __restrict__ partials, int n, float scale,
                                                      float* __restrict__ loss) {
  __shared__ float red[256];
  float t = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) t += partials[i];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = red[0] * scale;
}

// ------------------------------------------------------------------ Adam
constexpr int kMaxAdamTensors = 8;
struct AdamLaunch {
  float* p[kMaxAdamTensors];
  const float* g[kMaxAdamTensors];
  float* m[kMaxAdamTensors];
  float* v[kMaxAdamTensors];
  size_t n[kMaxAdamTensors];
  size_t first_quad[kMaxAdamTensors + 1];  // prefix sums of ceil(n / 4)
  int count;
  float lr_t, b1, b2, eps;
};

__device__ __forceinline__ void adam_one(float& p, float g, float& m, float& v, const AdamLaunch& L, float lr_t) {
  m = L.b1 * m + (1.f - L.b1) * g;
  v = L.b2 * v + (1.f - L.b2) * (g * g);
  p = p - lr_t * m / (sqrtf(v) + L.eps);
}

// lr_t of update t = *counter + 1 (device counter; the host form passes lr_t in L)
__device__ __forceinline__ float adam_lr_t(const AdamLaunch& L, const long long* counter, float lr) {
  const double t = (double)(*counter + 1);
  return (float)(lr * sqrt(1.0 - pow((double)L.b2, t)) / (1.0 - pow((double)L.b1, t)));
}

__global__ void k_adam_count(long long* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] += 1;
}

template <bool VEC>
__global__ __launch_bounds__(256) void k_adam(AdamLaunch L, const long long* __restrict__ counter, float lr) {
  // one bias correction per block from the device counter
  __shared__ float s_lr;
  if (counter) {
    if (threadIdx.x == 0) s_lr = adam_lr_t(L, counter, lr);
    __syncthreads();
  }
  const float lr_t = counter ? s_lr : L.lr_t;
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= L.first_quad[L.count]) return;
  int t = 0;
  while (t + 1 < L.count && q >= L.first_quad[t + 1]) ++t;
  const size_t e = (q - L.first_quad[t]) * 4;
  const size_t n = L.n[t];
  float* p = L.p[t];
  const float* g = L.g[t];
  float* m = L.m[t];
  float* v = L.v[t];
  if (VEC && e + 4 <= n) {
    float4 pp = *reinterpret_cast<float4*>(p + e);
    const float4 gg = *reinterpret_cast<const float4*>(g + e);
    float4 mm = *reinterpret_cast<float4*>(m + e);
    float4 vv = *reinterpret_cast<float4*>(v + e);
    adam_one(pp.x, gg.x, mm.x, vv.x, L, lr_t);
    adam_one(pp.y, gg.y, mm.y, vv.y, L, lr_t);
    adam_one(pp.z, gg.z, mm.z, vv.z, L, lr_t);
    adam_one(pp.w, gg.w, mm.w, vv.w, L, lr_t);
    *reinterpret_cast<float4*>(p + e) = pp;
    *reinterpret_cast<float4*>(m + e) = mm;
    *reinterpret_cast<float4*>(v + e) = vv;
  } else {
    for (size_t k = e; k < e + 4 && k < n; ++k) adam_one(p[k], g[k], m[k], v[k], L, lr_t);
  }
}

}  // namespace gft

using namespace gft;

extern "C" {

size_t gft_divergence_loss_workspace(const int32_t centered_shape[3]) {
  if (!centered_shape || centered_shape[0] < 1 || centered_shape[1] < 1 || centered_shape[2] < 1) return 0;
  const size_t n = (size_t)centered_shape[0] * centered_shape[1] * centered_shape[2];
  return align256(n * sizeof(float)) + align256(((n + 255) / 256) * sizeof(float));
}

int gft_divergence_loss(void* stream, const float* x, const float* y, const float* z,
                        const int32_t centered_shape[3], const float world_scale[3], float loss_scale,
                        int32_t accumulate, void* workspace, size_t workspace_bytes, float* grad_x, float* grad_y,
                        float* grad_z, float* loss) {
  if (!centered_shape || !world_scale) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  for (int i = 0; i < 3; ++i)
    if (centered_shape[i] < 1)
      return fail(GFT_ERR_INVALID_ARGUMENT, "All shape dimensions must be positive, dimension %d is %d", i,
                  centered_shape[i]);
  if (!x || !y || !z || !loss || ((grad_x || grad_y || grad_z) && !(grad_x && grad_y && grad_z)))
    return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  const size_t need = gft_divergence_loss_workspace(centered_shape);
  if (workspace_bytes < need || !workspace)
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                need);
  const DivDims d{centered_shape[2], centered_shape[1], centered_shape[0]};
  const size_t n = (size_t)d.W * d.H * d.D;
  const unsigned nb = (unsigned)((n + 255) / 256);
  float* div = (float*)workspace;
  float* partials = (float*)((char*)workspace + align256(n * sizeof(float)));
  hipStream_t s = (hipStream_t)stream;
  {
    KernelTimer t(s, "divergence");
    hipLaunchKernelGGL(k_divergence, dim3(nb), dim3(256), 0, s, x, y, z, d, world_scale[0], world_scale[1],
                       world_scale[2], div, partials);
  }
  int rc;
  if ((rc = check_hip(hipGetLastError(), "divergence launch"))) return rc;
  if (grad_x) {
    const size_t nmax = (size_t)(d.W + 1) * (d.H + 1) * (d.D + 1);
    KernelTimer t(s, "divergence_grad");
    hipLaunchKernelGGL(k_divergence_grad, dim3((unsigned)((nmax + 255) / 256), 3), dim3(256), 0, s, div, d,
                       2.f * loss_scale * world_scale[0], 2.f * loss_scale * world_scale[1],
                       2.f * loss_scale * world_scale[2], accumulate, grad_x, grad_y, grad_z);
    if ((rc = check_hip(hipGetLastError(), "divergence gradient launch"))) return rc;
  }
  hipLaunchKernelGGL(k_partials_sum, dim3(1), dim3(256), 0, s, partials, (int)nb, loss_scale, loss);
  return check_hip(hipGetLastError(), "divergence loss sum");
}

static int adam_common(void* stream, int32_t num_tensors, float* const* params, const float* const* grads,
                       float* const* exp_avg, float* const* exp_avg_sq, const int64_t* sizes, float learning_rate,
                       float beta1, float beta2, float epsilon, int64_t step, int64_t* step_counter) {
  if (num_tensors < 0 || num_tensors > kMaxAdamTensors)
    return fail(GFT_ERR_INVALID_ARGUMENT, "num_tensors must be in [0, %d], is %d", kMaxAdamTensors, num_tensors);
  if (!step_counter && step < 1)
    return fail(GFT_ERR_INVALID_ARGUMENT, "step must be >= 1 (the update count including this one)");
  if (num_tensors && (!params || !grads || !exp_avg || !exp_avg_sq || !sizes))
    return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  if (!(beta1 < 1.f)) return fail(GFT_ERR_INVALID_ARGUMENT, "beta1 must be < 1");
  AdamLaunch L{};
  L.count = num_tensors;
  L.b1 = beta1;
  L.b2 = beta2;
  L.eps = epsilon;
  if (!step_counter) {
    const double bc1 = 1.0 - std::pow((double)beta1, (double)step), bc2 = 1.0 - std::pow((double)beta2, (double)step);
    L.lr_t = (float)(learning_rate * std::sqrt(bc2) / bc1);
  }
  bool vec = true;
  L.first_quad[0] = 0;
  for (int t = 0; t < num_tensors; ++t) {
    if (sizes[t] < 0) return fail(GFT_ERR_INVALID_ARGUMENT, "negative size of tensor %d", t);
    if (sizes[t] && (!params[t] || !grads[t] || !exp_avg[t] || !exp_avg_sq[t]))
      return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer for tensor %d", t);
    L.p[t] = params[t];
    L.g[t] = grads[t];
    L.m[t] = exp_avg[t];
    L.v[t] = exp_avg_sq[t];
    L.n[t] = (size_t)sizes[t];
    L.first_quad[t + 1] = L.first_quad[t] + (L.n[t] + 3) / 4;
    for (const void* ptr : {(const void*)params[t], (const void*)grads[t], (const void*)exp_avg[t],
                            (const void*)exp_avg_sq[t]})
      if (reinterpret_cast<uintptr_t>(ptr) & 15) vec = false;
  }
  const size_t quads = L.first_quad[num_tensors];
  hipStream_t s = (hipStream_t)stream;
  const long long* counter = reinterpret_cast<const long long*>(step_counter);
  if (quads) {
    KernelTimer t(s, "adam");
    dim3 grid((unsigned)((quads + 255) / 256));
    if (vec) hipLaunchKernelGGL(k_adam<true>, grid, dim3(256), 0, s, L, counter, learning_rate);
    else hipLaunchKernelGGL(k_adam<false>, grid, dim3(256), 0, s, L, counter, learning_rate);
    int rc = check_hip(hipGetLastError(), "adam launch");
    if (rc) return rc;
  }
  if (step_counter) {
    hipLaunchKernelGGL(k_adam_count, dim3(1), dim3(64), 0, s, reinterpret_cast<long long*>(step_counter));
    return check_hip(hipGetLastError(), "adam step count");
  }
  return GFT_OK;
}

int gft_adam_step(void* stream, int32_t num_tensors, float* const* params, const float* const* grads,
                  float* const* exp_avg, float* const* exp_avg_sq, const int64_t* sizes, float learning_rate,
                  float beta1, float beta2, float epsilon, int64_t step) {
  return adam_common(stream, num_tensors, params, grads, exp_avg, exp_avg_sq, sizes, learning_rate, beta1, beta2,
                     epsilon, step, nullptr);
}

int gft_adam_step_device(void* stream, int32_t num_tensors, float* const* params, const float* const* grads,
                         float* const* exp_avg, float* const* exp_avg_sq, const int64_t* sizes, float learning_rate,
                         float beta1, float beta2, float epsilon, int64_t* step_counter) {
  if (!step_counter) return fail(GFT_ERR_INVALID_ARGUMENT, "null step counter");
  return adam_common(stream, num_tensors, params, grads, exp_avg, exp_avg_sq, sizes, learning_rate, beta1, beta2,
                     epsilon, 0, step_counter);
}

}  // extern "C"
