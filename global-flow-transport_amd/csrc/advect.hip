// Fused semi-Lagrangian / MacCormack advection of a scalar grid with the
// MC_SMOOTH revert clamp, and its exact adjoint w.r.t. the scalar and the
// centred velocity.
//
// Restates VelocityGrid.warp (phitest/render/data_structures.py:870-905) over
// SampleGridLut in relative, CLAMP, cell-centre-offset-0 mode
// (resample_grid.cu.cc:465-484, sampling.hpp:158-491):
//   p(x)   = x - v(x)*dt                       (lut = -v*dt, relative)
//   f(x)   = trilinear(d, p(x))                (forward warp)
//   b(x)   = trilinear(f, x + v(x)*dt)         (back warp with -lut)
//   c(x)   = f(x) + 0.5*(d(x) - b(x))
//   mn,mx  = min/max of the 8 corners of d around p(x)
//   out(x) = (c < mn || c > mx) ? f(x) : c     (MC_SMOOTH)
// The composition is two launches: the first writes f, the second computes
// b, the corner extrema and the selection, and stores the clamp mask for the
// backward. The backward is two tiled scatter launches (LDS accumulation,
// then float atomics into the gradient):
//   A: gb = mask ? 0 : -g/2 is scattered into gf at x+v*dt; gd(x) = mask?0:g/2;
//      gv(x) = gb * dt * grad f(x+v*dt)
//   B: gf = g + scattered gb; gf is scattered into gd at x-v*dt;
//      gv(x) -= gf * dt * grad d(x-v*dt)
// Trilinear derivatives use the floor/floor+1 corners (right derivative at
// integer positions) with CLAMP indices, so the derivative vanishes where the
// clamped read is constant.
//
// Storage: the density and the velocity are read as float or as bfloat16
// (uint16_t bit patterns, the upper half of a float); every value is widened
// to float on load and all arithmetic, the intermediate f, the output and the
// gradients are float.
#include "gft_device.hpp"

namespace gft {


enum { CLAMP_NONE = 0, CLAMP_MC_SMOOTH = 1, CLAMP_NEGATIVE = 2 };

__device__ __forceinline__ float ld(const float* __restrict__ p, size_t i) { return p[i]; }
__device__ __forceinline__ float ld(const uint16_t* __restrict__ p, size_t i) {
  return __uint_as_float((unsigned)p[i] << 16);
}

template <typename T>
__device__ __forceinline__ float3 load_vel(const T* __restrict__ vel, size_t i) {
  return make_float3(ld(vel, 3 * i), ld(vel, 3 * i + 1), ld(vel, 3 * i + 2));
}

// Clamped trilinear read of a scalar grid plus its 8 corner values.
struct Tri {
  float v[2][2][2];  // [z][y][x], corners floor / floor+1 (clamped)
  float wx, wy, wz;
  int ix[2], iy[2], iz[2];
};

template <typename T>
__device__ __forceinline__ Tri tri_clamp(const T* __restrict__ buf, Dims3 d, float px, float py, float pz) {
  Tri t;
  float fxf = floorf(px), fyf = floorf(py), fzf = floorf(pz);
  t.wx = px - fxf; t.wy = py - fyf; t.wz = pz - fzf;
  int fx = (int)fxf, fy = (int)fyf, fz = (int)fzf;
  t.ix[0] = clampi(fx, 0, d.x - 1); t.ix[1] = clampi(fx + 1, 0, d.x - 1);
  t.iy[0] = clampi(fy, 0, d.y - 1); t.iy[1] = clampi(fy + 1, 0, d.y - 1);
  t.iz[0] = clampi(fz, 0, d.z - 1); t.iz[1] = clampi(fz + 1, 0, d.z - 1);
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) t.v[k][j][i] = ld(buf, flat3(t.ix[i], t.iy[j], t.iz[k], d));
  return t;
}

__device__ __forceinline__ float tri_value(const Tri& t) {
  float v00 = lerpf(t.v[0][0][0], t.v[0][0][1], t.wx);
  float v01 = lerpf(t.v[0][1][0], t.v[0][1][1], t.wx);
  float v0 = lerpf(v00, v01, t.wy);
  float v10 = lerpf(t.v[1][0][0], t.v[1][0][1], t.wx);
  float v11 = lerpf(t.v[1][1][0], t.v[1][1][1], t.wx);
  float v1 = lerpf(v10, v11, t.wy);
  return lerpf(v0, v1, t.wz);
}

__device__ __forceinline__ float3 tri_grad(const Tri& t) {
  const float (*v)[2][2] = t.v;
  float gx = lerpf(lerpf(v[0][0][1] - v[0][0][0], v[0][1][1] - v[0][1][0], t.wy),
                   lerpf(v[1][0][1] - v[1][0][0], v[1][1][1] - v[1][1][0], t.wy), t.wz);
  float gy = lerpf(lerpf(v[0][1][0] - v[0][0][0], v[0][1][1] - v[0][0][1], t.wx),
                   lerpf(v[1][1][0] - v[1][0][0], v[1][1][1] - v[1][0][1], t.wx), t.wz);
  float gz = lerpf(lerpf(v[1][0][0] - v[0][0][0], v[1][0][1] - v[0][0][1], t.wx),
                   lerpf(v[1][1][0] - v[0][1][0], v[1][1][1] - v[0][1][1], t.wx), t.wy);
  return make_float3(gx, gy, gz);
}

__device__ __forceinline__ void voxel_of(size_t i, Dims3 d, int& x, int& y, int& z) {
  x = (int)(i % d.x);
  size_t r = i / d.x;
  y = (int)(r % d.y);
  z = (int)(r / d.y);
}

// pass 1: f = trilinear(d, x - v dt)
template <typename T>
__global__ __launch_bounds__(256) void k_advect_sl(const T* __restrict__ dens, const T* __restrict__ vel,
                                                   float* __restrict__ f, AdvectDims a, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  voxel_of(i, a.d, x, y, z);
  float3 v = load_vel(vel, i);
  Tri t = tri_clamp(dens, a.d, x - v.x * a.dt, y - v.y * a.dt, z - v.z * a.dt);
  f[i] = tri_value(t);
}

// order 1 output: out = f, or max(f, 0) with NEGATIVE (mask marks clamped cells;
// tf.maximum passes the gradient at equality)
__global__ __launch_bounds__(256) void k_advect_o1_out(const float* __restrict__ f, float* __restrict__ out,
                                                       uint8_t* __restrict__ mask, AdvectDims a, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float r = f[i];
  uint8_t m = 0;
  if (a.clamp == CLAMP_NEGATIVE) {
    m = r < 0.f;
    r = fmaxf(r, 0.f);
  }
  out[i] = r;
  mask[i] = m;
}

// pass 2: MacCormack correction with the MC_SMOOTH revert clamp
template <typename T>
__global__ __launch_bounds__(256) void k_advect_mc(const T* __restrict__ dens, const T* __restrict__ vel,
                                                   const float* __restrict__ f, float* __restrict__ out,
                                                   uint8_t* __restrict__ mask, AdvectDims a, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  voxel_of(i, a.d, x, y, z);
  float3 v = load_vel(vel, i);
  float fx = f[i];
  Tri tb = tri_clamp(f, a.d, x + v.x * a.dt, y + v.y * a.dt, z + v.z * a.dt);
  float b = tri_value(tb);
  float c = fx + 0.5f * (ld(dens, i) - b);
  uint8_t m = 0;
  if (a.clamp == CLAMP_MC_SMOOTH) {
    // corner extrema with the reference's ceil corners (ceil == floor at
    // integer positions), CLAMP indices (sampling.hpp:372-491)
    float px = x - v.x * a.dt, py = y - v.y * a.dt, pz = z - v.z * a.dt;
    int fxi = (int)floorf(px), fyi = (int)floorf(py), fzi = (int)floorf(pz);
    int cxi = (int)ceilf(px), cyi = (int)ceilf(py), czi = (int)ceilf(pz);
    int ix[2] = {clampi(fxi, 0, a.d.x - 1), clampi(cxi, 0, a.d.x - 1)};
    int iy[2] = {clampi(fyi, 0, a.d.y - 1), clampi(cyi, 0, a.d.y - 1)};
    int iz[2] = {clampi(fzi, 0, a.d.z - 1), clampi(czi, 0, a.d.z - 1)};
    float mn = kFloatMax, mx = -kFloatMax;
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int ii = 0; ii < 2; ++ii) {
          float s = ld(dens, flat3(ix[ii], iy[j], iz[k], a.d));
          mn = fminf(mn, s);
          mx = fmaxf(mx, s);
        }
    m = (c < mn) | (c > mx);
    if (m) c = fx;
  }
  if (a.clamp == CLAMP_NEGATIVE) {
    m = c < 0.f;  // max(c, 0): zero gradient where clamped
    c = fmaxf(c, 0.f);
  }
  out[i] = c;
  mask[i] = m;
}

// ----------------------------------------------------------- backward
// The adjoint scatters are accumulated per tile of kTX x kTY x kTZ voxels in
// an LDS region covering the tile plus a kHalo-cell margin (a position moved
// by s with -kHalo <= s < kHalo - 1 in every axis has all 8 corners inside
// it), with LDS atomics; the region is then added to the global gradient with one coalesced
// global atomic per nonzero cell. Corners outside the region (faster motion)
// go straight to a global atomic, so any velocity is handled exactly.
// kHalo = 2 keeps the corners of displacements in [-2, 1) in the region; the
// rest (faster motion along an axis) take global atomics. At c2 (|v| <= 1.5)
// the two passes take 0.051 + 0.052 ms with kHalo = 2 against 0.060 + 0.060 ms
// with kHalo = 3, whose larger region costs more in flushes than the atomics
// it saves; at c5 (|v| <= 2) 0.188 + 0.203 against 0.228 + 0.232 ms.
constexpr int kTX = 32, kTY = 4, kTZ = 4, kHalo = 2;
constexpr int kRX = kTX + 2 * kHalo, kRY = kTY + 2 * kHalo, kRZ = kTZ + 2 * kHalo;
constexpr int kRegion = kRX * kRY * kRZ;

// The region accumulates in FP64: ds_add_f64 runs at about 20x the rate of
// ds_add_f32 on gfx950 (measured 7 vs 0.33 lane-ops per clock per CU), and the
// FP64 sum of the FP32 addends is rounded once when it is flushed. The pointer
// carries the LDS address space (a generic pointer held in the struct would
// make the compiler emit flat atomics).
typedef __attribute__((address_space(3))) double lds_double;

struct TileScatter {
  lds_double* lds;
  int x0, y0, z0;  // tile origin

  __device__ __forceinline__ void clear() {
    for (int k = threadIdx.x; k < kRegion; k += blockDim.x) lds[k] = 0.0;
  }
  __device__ __forceinline__ void add(float* __restrict__ g, Dims3 d, const Tri& t, float val) {
    const float wx[2] = {1.f - t.wx, t.wx}, wy[2] = {1.f - t.wy, t.wy}, wz[2] = {1.f - t.wz, t.wz};
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const float w = wx[i] * wy[j] * wz[k];
          if (w == 0.f) continue;
          const int lx = t.ix[i] - x0 + kHalo, ly = t.iy[j] - y0 + kHalo, lz = t.iz[k] - z0 + kHalo;
          if (((unsigned)lx < (unsigned)kRX) & ((unsigned)ly < (unsigned)kRY) & ((unsigned)lz < (unsigned)kRZ))
            __hip_atomic_fetch_add(lds + (lz * kRY + ly) * kRX + lx, (double)(val * w), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
          else
            atomic_addf(g + flat3(t.ix[i], t.iy[j], t.iz[k], d), val * w);
        }
  }
  __device__ __forceinline__ void flush(float* __restrict__ g, Dims3 d) {
    for (int k = threadIdx.x; k < kRegion; k += blockDim.x) {
      const float v = (float)lds[k];
      if (v == 0.f) continue;
      const int x = x0 - kHalo + k % kRX, y = y0 - kHalo + (k / kRX) % kRY, z = z0 - kHalo + k / (kRX * kRY);
      // cells outside the grid never receive a corner (indices are clamped)
      if (((unsigned)x < (unsigned)d.x) & ((unsigned)y < (unsigned)d.y) & ((unsigned)z < (unsigned)d.z))
        atomic_addf(g + flat3(x, y, z, d), v);
    }
  }
};

__device__ __forceinline__ bool tile_voxel(int k, Dims3 d, int x0, int y0, int z0, int& x, int& y, int& z) {
  x = x0 + k % kTX;
  y = y0 + (k / kTX) % kTY;
  z = z0 + k / (kTX * kTY);
  return (x < d.x) & (y < d.y) & (z < d.z);
}

// backward A (MacCormack only): scatter of the back-warp gradient into gf,
// direct density term, velocity term of the back warp
template <typename T>
__global__ __launch_bounds__(256) void k_advect_bwd_a(const T* __restrict__ vel, const float* __restrict__ f,
                                                      const uint8_t* __restrict__ mask,
                                                      const float* __restrict__ gout, float* __restrict__ gf_acc,
                                                      float* __restrict__ gd, float* __restrict__ gv, AdvectDims a) {
  __shared__ double region[kRegion];
  TileScatter ts{(lds_double*)region, (int)blockIdx.x * kTX, (int)blockIdx.y * kTY, (int)blockIdx.z * kTZ};
  ts.clear();
  __syncthreads();
  for (int k = threadIdx.x; k < kTX * kTY * kTZ; k += blockDim.x) {
    int x, y, z;
    if (!tile_voxel(k, a.d, ts.x0, ts.y0, ts.z0, x, y, z)) continue;
    const size_t i = flat3(x, y, z, a.d);
    float3 v = load_vel(vel, i);
    float g = gout[i];
    uint8_t m = mask[i];
    // gradient of the corrected value c: zero where MC_SMOOTH reverted to f or
    // NEGATIVE clamped to 0
    float gc = m ? 0.f : g;
    float gb = -0.5f * gc;
    gd[i] = 0.5f * gc;
    Tri tb = tri_clamp(f, a.d, x + v.x * a.dt, y + v.y * a.dt, z + v.z * a.dt);
    float3 df = tri_grad(tb);
    gv[3 * i] = gb * a.dt * df.x;
    gv[3 * i + 1] = gb * a.dt * df.y;
    gv[3 * i + 2] = gb * a.dt * df.z;
    if (gb != 0.f) ts.add(gf_acc, a.d, tb, gb);
  }
  __syncthreads();
  ts.flush(gf_acc, a.d);
}

// backward B: total gf scattered into gd at x - v dt, velocity term of the
// forward warp. FIRST = true when pass A did not run (order 1): gv is
// initialised here and gd must have been cleared.
template <typename T, bool FIRST>
__global__ __launch_bounds__(256) void k_advect_bwd_b(const T* __restrict__ dens, const T* __restrict__ vel,
                                                      const float* __restrict__ gout,
                                                      const float* __restrict__ gf_acc,
                                                      const uint8_t* __restrict__ mask, float* __restrict__ gd,
                                                      float* __restrict__ gv, AdvectDims a) {
  __shared__ double region[kRegion];
  TileScatter ts{(lds_double*)region, (int)blockIdx.x * kTX, (int)blockIdx.y * kTY, (int)blockIdx.z * kTZ};
  ts.clear();
  __syncthreads();
  for (int k = threadIdx.x; k < kTX * kTY * kTZ; k += blockDim.x) {
    int x, y, z;
    if (!tile_voxel(k, a.d, ts.x0, ts.y0, ts.z0, x, y, z)) continue;
    const size_t i = flat3(x, y, z, a.d);
    float3 v = load_vel(vel, i);
    float g = gout[i];
    float gf;
    if (FIRST) {
      // order 1: out = f (or max(f,0) with NEGATIVE)
      gf = (a.clamp == CLAMP_NEGATIVE && mask[i]) ? 0.f : g;
    } else {
      // f(x) enters out directly in both clamp branches, plus the back-warp scatter
      uint8_t m = mask[i];
      float gc = m ? 0.f : g;
      gf = (a.clamp == CLAMP_MC_SMOOTH ? g : gc) + gf_acc[i];
    }
    Tri tf = tri_clamp(dens, a.d, x - v.x * a.dt, y - v.y * a.dt, z - v.z * a.dt);
    float3 dd = tri_grad(tf);
    float3 gvv = make_float3(-gf * a.dt * dd.x, -gf * a.dt * dd.y, -gf * a.dt * dd.z);
    if (FIRST) {
      gv[3 * i] = gvv.x;
      gv[3 * i + 1] = gvv.y;
      gv[3 * i + 2] = gvv.z;
    } else {
      gv[3 * i] += gvv.x;
      gv[3 * i + 1] += gvv.y;
      gv[3 * i + 2] += gvv.z;
    }
    if (gf != 0.f) ts.add(gd, a.d, tf, gf);
  }
  __syncthreads();
  ts.flush(gd, a.d);
}

static dim3 grid1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

template <typename T>
static hipError_t advect_fwd_typed(hipStream_t s, const T* dens, const T* vel, float* f, float* out, uint8_t* mask,
                                   const AdvectDims& a) {
  size_t n = size_t(a.d.x) * a.d.y * a.d.z;
  {
    KernelTimer t(s, "advect_sl");
    hipLaunchKernelGGL(k_advect_sl<T>, grid1(n), dim3(256), 0, s, dens, vel, f, a, n);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (a.order == 2) {
    KernelTimer t(s, "advect_mc");
    hipLaunchKernelGGL(k_advect_mc<T>, grid1(n), dim3(256), 0, s, dens, vel, f, out, mask, a, n);
  } else {
    KernelTimer t(s, "advect_o1_out");
    hipLaunchKernelGGL(k_advect_o1_out, grid1(n), dim3(256), 0, s, f, out, mask, a, n);
  }
  return hipGetLastError();
}

template <typename T>
static hipError_t advect_bwd_typed(hipStream_t s, const T* dens, const T* vel, const float* f, const uint8_t* mask,
                                   const float* gout, float* gf_acc, float* gd, float* gv, const AdvectDims& a) {
  const dim3 tiles((unsigned)((a.d.x + kTX - 1) / kTX), (unsigned)((a.d.y + kTY - 1) / kTY),
                   (unsigned)((a.d.z + kTZ - 1) / kTZ));
  if (a.order == 2) {
    {
      KernelTimer t(s, "advect_bwd_a");
      hipLaunchKernelGGL(k_advect_bwd_a<T>, tiles, dim3(256), 0, s, vel, f, mask, gout, gf_acc, gd, gv, a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    KernelTimer t(s, "advect_bwd_b");
    hipLaunchKernelGGL((k_advect_bwd_b<T, false>), tiles, dim3(256), 0, s, dens, vel, gout, gf_acc, mask, gd, gv, a);
  } else {
    // gd must be zero on entry (the caller clears it)
    KernelTimer t(s, "advect_bwd_b");
    hipLaunchKernelGGL((k_advect_bwd_b<T, true>), tiles, dim3(256), 0, s, dens, vel, gout, gf_acc, mask, gd, gv, a);
  }
  return hipGetLastError();
}

// dens and vel are float, or bfloat16 bit patterns with `bf16`
hipError_t advect_fwd_launch(hipStream_t s, const void* dens, const void* vel, bool bf16, float* f, float* out,
                             uint8_t* mask, const AdvectDims& a) {
  if (bf16)
    return advect_fwd_typed(s, (const uint16_t*)dens, (const uint16_t*)vel, f, out, mask, a);
  return advect_fwd_typed(s, (const float*)dens, (const float*)vel, f, out, mask, a);
}

hipError_t advect_bwd_launch(hipStream_t s, const void* dens, const void* vel, bool bf16, const float* f,
                             const uint8_t* mask, const float* gout, float* gf_acc, float* gd, float* gv,
                             const AdvectDims& a) {
  if (bf16)
    return advect_bwd_typed(s, (const uint16_t*)dens, (const uint16_t*)vel, f, mask, gout, gf_acc, gd, gv, a);
  return advect_bwd_typed(s, (const float*)dens, (const float*)vel, f, mask, gout, gf_acc, gd, gv, a);
}

}  // namespace gft
