// Device-side helpers shared by the gfx950 kernels: channel vectors, grid
// indexing, trilinear/nearest taps with the reference's boundary semantics,
// blend recurrences, and the camera records passed as kernel arguments.
//
// Layout: NDHWC float32, x fastest; a voxel holds C in {1,2,4} floats, so a
// C=4 voxel is one 16-byte load. Cell centres sit at +0.5 in index
// coordinates; samplers subtract the cell-centre offset before the tap
// (sampling_v2.hpp:358-372, sampling.hpp:521-548).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gft {

enum Filter { F_NEAREST = 0, F_LINEAR = 1, F_MIN = 2, F_MAX = 3 };
enum Boundary { B_BORDER = 0, B_CLAMP = 1, B_WRAP = 2 };
enum Blend { BL_BEER_LAMBERT = 0, BL_ALPHA = 1, BL_ALPHA_ADDITIVE = 2, BL_ADDITIVE = 3 };
enum Mip { MIP_NONE = 0, MIP_NEAREST = 1, MIP_LINEAR = 2 };
enum Coord { CM_LINDEPTH = 0, CM_LINDEPTH_REVERSE = 1, CM_TRANSFORM = 2, CM_TRANSFORM_REVERSE = 3 };
enum GradMode { GM_REFERENCE = 0, GM_EXACT = 1 };

constexpr int kMaxViewsPerLaunch = 16;
// FLOAT_MAX of the reference (vectormath.hpp:12-14)
constexpr float kFloatMax = 3.40282e+38f;

// ---------------------------------------------------------------- vectors
template <int C> struct Vec { float v[C]; };

template <int C> __device__ __forceinline__ Vec<C> vzero() {
  Vec<C> r;
#pragma unroll
  for (int i = 0; i < C; ++i) r.v[i] = 0.f;
  return r;
}

template <int C> __device__ __forceinline__ Vec<C> vload(const float* __restrict__ p, size_t voxel) {
  Vec<C> r;
  if constexpr (C == 4) {
    float4 t = *reinterpret_cast<const float4*>(p + voxel * 4);
    r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
  } else if constexpr (C == 2) {
    float2 t = *reinterpret_cast<const float2*>(p + voxel * 2);
    r.v[0] = t.x; r.v[1] = t.y;
  } else {
    r.v[0] = p[voxel];
  }
  return r;
}

template <int C> __device__ __forceinline__ void vstore(float* __restrict__ p, size_t voxel, const Vec<C>& a) {
  if constexpr (C == 4) {
    *reinterpret_cast<float4*>(p + voxel * 4) = make_float4(a.v[0], a.v[1], a.v[2], a.v[3]);
  } else if constexpr (C == 2) {
    *reinterpret_cast<float2*>(p + voxel * 2) = make_float2(a.v[0], a.v[1]);
  } else {
    p[voxel] = a.v[0];
  }
}

__device__ __forceinline__ void atomic_addf(float* p, float v) {
  // hardware float atomic on coarse-grained device memory
  unsafeAtomicAdd(p, v);
}

template <int C> __device__ __forceinline__ void vatomic_add(float* p, size_t voxel, const Vec<C>& a, float w) {
#pragma unroll
  for (int i = 0; i < C; ++i) atomic_addf(p + voxel * C + i, a.v[i] * w);
}

// ------------------------------------------------------------------- grid
struct Dims3 {
  int x, y, z;
};

__device__ __forceinline__ size_t flat3(int x, int y, int z, Dims3 d) {
  return (size_t(z) * d.y + y) * d.x + x;
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }
__device__ __forceinline__ int pmodi(int v, int m) {
  int r = v % m;
  return r < 0 ? r + m : r;
}
__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + t * (b - a); }

// Trilinear corner set for a position with the cell-centre offset already
// removed. Corner 0/1 per axis are floor and floor+1 (at exact integers the
// weight of corner 1 is 0, which reproduces the reference's ceil==floor case).
// valid[] marks corners that a BORDER read treats as present; indices are
// always in range so reads never leave the buffer.
struct Taps {
  int ix[2], iy[2], iz[2];
  bool vx[2], vy[2], vz[2];
  float wx, wy, wz;  // weight of corner 1
  bool inside;       // BORDER: the sample itself lies in [0,dims)
};

template <int BM> __device__ __forceinline__ Taps make_taps(float px, float py, float pz, Dims3 d) {
  Taps t;
  float fxf = floorf(px), fyf = floorf(py), fzf = floorf(pz);
  t.wx = px - fxf; t.wy = py - fyf; t.wz = pz - fzf;
  int fx = (int)fxf, fy = (int)fyf, fz = (int)fzf;
  int cx = fx + 1, cy = fy + 1, cz = fz + 1;
  if constexpr (BM == B_BORDER) {
    float qx = px + 0.5f, qy = py + 0.5f, qz = pz + 0.5f;
    t.inside = (0.f <= qx) & (qx < (float)d.x) & (0.f <= qy) & (qy < (float)d.y) & (0.f <= qz) & (qz < (float)d.z);
    t.vx[0] = (fx >= 0) & (fx < d.x); t.vx[1] = (cx >= 0) & (cx < d.x);
    t.vy[0] = (fy >= 0) & (fy < d.y); t.vy[1] = (cy >= 0) & (cy < d.y);
    t.vz[0] = (fz >= 0) & (fz < d.z); t.vz[1] = (cz >= 0) & (cz < d.z);
    fx = clampi(fx, 0, d.x - 1); cx = clampi(cx, 0, d.x - 1);
    fy = clampi(fy, 0, d.y - 1); cy = clampi(cy, 0, d.y - 1);
    fz = clampi(fz, 0, d.z - 1); cz = clampi(cz, 0, d.z - 1);
  } else if constexpr (BM == B_CLAMP) {
    t.inside = true;
    t.vx[0] = t.vx[1] = t.vy[0] = t.vy[1] = t.vz[0] = t.vz[1] = true;
    fx = clampi(fx, 0, d.x - 1); cx = clampi(cx, 0, d.x - 1);
    fy = clampi(fy, 0, d.y - 1); cy = clampi(cy, 0, d.y - 1);
    fz = clampi(fz, 0, d.z - 1); cz = clampi(cz, 0, d.z - 1);
  } else {
    t.inside = true;
    t.vx[0] = t.vx[1] = t.vy[0] = t.vy[1] = t.vz[0] = t.vz[1] = true;
    fx = pmodi(fx, d.x); cx = pmodi(cx, d.x);
    fy = pmodi(fy, d.y); cy = pmodi(cy, d.y);
    fz = pmodi(fz, d.z); cz = pmodi(cz, d.z);
  }
  t.ix[0] = fx; t.ix[1] = cx; t.iy[0] = fy; t.iy[1] = cy; t.iz[0] = fz; t.iz[1] = cz;
  return t;
}

// Sample range [lo, hi] of the ray n + i st, i in [0, S), that can hold
// positions in the box [-m, dims + m] (ray coordinates, which carry the +0.5
// cell-centre offset: m = 0 is the grid box, inside meaning 0 <= p < dims);
// one sample of slack on each end, so per-sample tests stay exact.
// lo > hi: no sample.
__device__ __forceinline__ void clip_ray(float3 n, float3 st, Dims3 d, int S, float m, int& lo, int& hi) {
  float tlo = 0.f, thi = (float)(S - 1);
  const float nn[3] = {n.x, n.y, n.z}, ss[3] = {st.x, st.y, st.z};
  const float dd[3] = {(float)d.x, (float)d.y, (float)d.z};
  bool miss = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (ss[a] != 0.f) {
      float t0 = (-m - nn[a]) / ss[a], t1 = (dd[a] + m - nn[a]) / ss[a];
      tlo = fmaxf(tlo, fminf(t0, t1));
      thi = fminf(thi, fmaxf(t0, t1));
    } else if (!(-m <= nn[a] && nn[a] < dd[a] + m)) {
      miss = true;
    }
  }
  lo = max(0, (int)floorf(tlo) - 1);
  hi = min(S - 1, (int)ceilf(thi) + 1);
  if (miss || thi < tlo - 2.f) {
    lo = 1;
    hi = 0;
  }
}

// Trilinear value, lerp order x, y, z as in read3DInterpolated.
template <int C, int BM>
__device__ __forceinline__ Vec<C> read_linear(const float* __restrict__ buf, Dims3 d, float px, float py, float pz) {
  Taps t = make_taps<BM>(px, py, pz, d);
  Vec<C> r = vzero<C>();
  if (BM == B_BORDER && !t.inside) return r;
  Vec<C> c[2][2][2];
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        c[k][j][i] = vload<C>(buf, flat3(t.ix[i], t.iy[j], t.iz[k], d));
        if (BM == B_BORDER && !(t.vx[i] & t.vy[j] & t.vz[k])) c[k][j][i] = vzero<C>();
      }
#pragma unroll
  for (int ch = 0; ch < C; ++ch) {
    float v00 = lerpf(c[0][0][0].v[ch], c[0][0][1].v[ch], t.wx);
    float v01 = lerpf(c[0][1][0].v[ch], c[0][1][1].v[ch], t.wx);
    float v0 = lerpf(v00, v01, t.wy);
    float v10 = lerpf(c[1][0][0].v[ch], c[1][0][1].v[ch], t.wx);
    float v11 = lerpf(c[1][1][0].v[ch], c[1][1][1].v[ch], t.wx);
    float v1 = lerpf(v10, v11, t.wy);
    r.v[ch] = lerpf(v0, v1, t.wz);
  }
  return r;
}

// U trilinear reads with all 8U loads issued before any value is used (the
// indices make_taps returns are always inside the grid, so the loads need no
// guard): the same corners, zeroing and lerp order as read_linear, so every
// value equals read_linear's, and a BORDER sample outside the grid reads 0.
template <int C, int BM, int U>
__device__ __forceinline__ void read_linear_batch(const float* __restrict__ buf, Dims3 d, const float px[U],
                                                  const float py[U], const float pz[U], Vec<C> out[U]) {
  Taps t[U];
  Vec<C> c[U][2][2][2];
#pragma unroll
  for (int u = 0; u < U; ++u) t[u] = make_taps<BM>(px[u], py[u], pz[u], d);
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 2; ++i) c[u][k][j][i] = vload<C>(buf, flat3(t[u].ix[i], t[u].iy[j], t[u].iz[k], d));
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (BM == B_BORDER) {
#pragma unroll
      for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int i = 0; i < 2; ++i)
            if (!(t[u].vx[i] & t[u].vy[j] & t[u].vz[k])) c[u][k][j][i] = vzero<C>();
    }
    Vec<C> r;
#pragma unroll
    for (int ch = 0; ch < C; ++ch) {
      const float v00 = lerpf(c[u][0][0][0].v[ch], c[u][0][0][1].v[ch], t[u].wx);
      const float v01 = lerpf(c[u][0][1][0].v[ch], c[u][0][1][1].v[ch], t[u].wx);
      const float v0 = lerpf(v00, v01, t[u].wy);
      const float v10 = lerpf(c[u][1][0][0].v[ch], c[u][1][0][1].v[ch], t[u].wx);
      const float v11 = lerpf(c[u][1][1][0].v[ch], c[u][1][1][1].v[ch], t[u].wx);
      const float v1 = lerpf(v10, v11, t[u].wy);
      r.v[ch] = lerpf(v0, v1, t[u].wz);
    }
    out[u] = (BM == B_BORDER && !t[u].inside) ? vzero<C>() : r;
  }
}

// Nearest tap: index = trunc(p + 0.5) as the reference's int cast.
template <int BM> __device__ __forceinline__ bool nearest_index(float px, float py, float pz, Dims3 d, size_t& f) {
  int ix = (int)(px + 0.5f), iy = (int)(py + 0.5f), iz = (int)(pz + 0.5f);
  if constexpr (BM == B_BORDER) {
    bool ok = (ix >= 0) & (ix < d.x) & (iy >= 0) & (iy < d.y) & (iz >= 0) & (iz < d.z);
    f = ok ? flat3(ix, iy, iz, d) : 0;
    return ok;
  } else if constexpr (BM == B_CLAMP) {
    f = flat3(clampi(ix, 0, d.x - 1), clampi(iy, 0, d.y - 1), clampi(iz, 0, d.z - 1), d);
  } else {
    f = flat3(pmodi(ix, d.x), pmodi(iy, d.y), pmodi(iz, d.z), d);
  }
  return true;
}
template <int C, int BM>
__device__ __forceinline__ Vec<C> read_nearest(const float* __restrict__ buf, Dims3 d, float px, float py, float pz) {
  size_t f;
  bool ok = nearest_index<BM>(px, py, pz, d, f);
  return ok ? vload<C>(buf, f) : vzero<C>();
}

// Raymarch-sampler MIN/MAX (sampling_v2.hpp:149-253): extrema start at
// +-FLOAT_MAX and are updated by lerp(ext, v, v<ext); an out-of-domain BORDER
// sample returns the initial extrema. Restated literally.
template <int C, int BM>
__device__ __forceinline__ Vec<C> read_minmax_v2(const float* __restrict__ buf, Dims3 d, float px, float py, float pz,
                                                 bool is_max) {
  Taps t = make_taps<BM>(px, py, pz, d);
  Vec<C> m;
#pragma unroll
  for (int ch = 0; ch < C; ++ch) m.v[ch] = is_max ? -kFloatMax : kFloatMax;
  if (BM == B_BORDER && !t.inside) return m;
  // the floor==ceil case must visit the same cell twice like the reference
  int ex[2] = {t.ix[0], t.wx == 0.f ? t.ix[0] : t.ix[1]};
  int ey[2] = {t.iy[0], t.wy == 0.f ? t.iy[0] : t.iy[1]};
  int ez[2] = {t.iz[0], t.wz == 0.f ? t.iz[0] : t.iz[1]};
  bool evx[2] = {t.vx[0], t.wx == 0.f ? t.vx[0] : t.vx[1]};
  bool evy[2] = {t.vy[0], t.wy == 0.f ? t.vy[0] : t.vy[1]};
  bool evz[2] = {t.vz[0], t.wz == 0.f ? t.vz[0] : t.vz[1]};
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        if (BM == B_BORDER && !(evx[i] & evy[j] & evz[k])) continue;
        Vec<C> v = vload<C>(buf, flat3(ex[i], ey[j], ez[k], d));
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
          float s = is_max ? (float)(m.v[ch] < v.v[ch]) : (float)(v.v[ch] < m.v[ch]);
          m.v[ch] = m.v[ch] + s * (v.v[ch] - m.v[ch]);
        }
      }
  return m;
}

// ------------------------------------------------------------------ blend
// Recurrences of blending.hpp:9-227. C==1 and ADDITIVE are plain sums.
template <int C, int BL> __device__ __forceinline__ void blend(Vec<C>& acc, const Vec<C>& c) {
  if constexpr (C == 1 || BL == BL_ADDITIVE) {
#pragma unroll
    for (int i = 0; i < C; ++i) acc.v[i] += c.v[i];
  } else {
    constexpr int A = C - 1;
    if constexpr (BL == BL_BEER_LAMBERT) {
      float dd = fmaxf(acc.v[A] + c.v[A], 0.f);
      float t = expf(-dd);
#pragma unroll
      for (int i = 0; i < A; ++i) acc.v[i] += t * c.v[i];
      acc.v[A] = dd;
    } else if constexpr (BL == BL_ALPHA) {
      float t = 1.f - acc.v[A];
#pragma unroll
      for (int i = 0; i < A; ++i) acc.v[i] += t * c.v[i];
      acc.v[A] += t * c.v[A];
    } else {
      float t = 1.f - c.v[A];
#pragma unroll
      for (int i = 0; i < A; ++i) acc.v[i] += t * c.v[i];
      acc.v[A] += c.v[A];
    }
  }
}

// Reverse step: g is the running output gradient, acc the accumulator after
// this cell; both are updated to the state before the cell. Returns the
// gradient w.r.t. the cell. EXACT corrects the ALPHA derivative; REFERENCE
// keeps blending.hpp:128-142.
template <int C, int BL, bool EXACT>
__device__ __forceinline__ Vec<C> blend_grad(Vec<C>& g, const Vec<C>& c, Vec<C>& acc) {
  Vec<C> gc;
  if constexpr (C == 1 || BL == BL_ADDITIVE) {
    gc = g;
  } else {
    constexpr int A = C - 1;
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < A; ++i) s += c.v[i] * g.v[i];
    if constexpr (BL == BL_BEER_LAMBERT) {
      float e = expf(-acc.v[A]);
      acc.v[A] = fmaxf(acc.v[A] - c.v[A], 0.f);
      float d_dacc = g.v[A] - e * s;
#pragma unroll
      for (int i = 0; i < A; ++i) gc.v[i] = e * g.v[i];
      gc.v[A] = d_dacc;
      g.v[A] = d_dacc;
    } else if constexpr (BL == BL_ALPHA) {
      float tprev = (1.f - acc.v[A]) / (1.f - c.v[A]);
      if constexpr (EXACT) {
        acc.v[A] = 1.f - tprev;
#pragma unroll
        for (int i = 0; i < A; ++i) gc.v[i] = tprev * g.v[i];
        gc.v[A] = tprev * g.v[A];
        g.v[A] = (1.f - c.v[A]) * g.v[A] - s;
      } else {
        acc.v[A] = tprev;
        float t = 1.f - acc.v[A];
        float d_dacc = (1.f - c.v[A]) * g.v[A] - s;
#pragma unroll
        for (int i = 0; i < A; ++i) gc.v[i] = t * g.v[i];
        gc.v[A] = t * g.v[A];
        g.v[A] = d_dacc;
      }
    } else {
      acc.v[A] = acc.v[A] - c.v[A];
      float t = 1.f - c.v[A];
#pragma unroll
      for (int i = 0; i < A; ++i) gc.v[i] = t * g.v[i];
      gc.v[A] = g.v[A] - s;
    }
  }
  return gc;
}

// --------------------------------------------------------- camera records
// Column-major 4x4: m[4*col + row].
__device__ __forceinline__ float4 mat_vec(const float* m, float x, float y, float z, float w) {
  return make_float4(fmaf(m[12], w, fmaf(m[8], z, fmaf(m[4], y, m[0] * x))),
                     fmaf(m[13], w, fmaf(m[9], z, fmaf(m[5], y, m[1] * x))),
                     fmaf(m[14], w, fmaf(m[10], z, fmaf(m[6], y, m[2] * x))),
                     fmaf(m[15], w, fmaf(m[11], z, fmaf(m[7], y, m[3] * x))));
}

// One (batch, camera) view of the fused raymarch: the inverse projection and
// inverse model-view, and where its input and output live.
struct RayView {
  float pinv[16];
  float mvinv[16];
  int in_batch;
  int out_slot;
};
struct RayLaunch {
  RayView view[kMaxViewsPerLaunch];
  int num_views;
};

// Forward projection of one view for the voxel-parallel gather: object space
// -> clip space (P * V * M) and the output slot of the view.
struct ProjView {
  float pmv[16];
  // pixel rays without a per-pixel record: with g(px, py) = g0 + px gx + py gy
  // (homogeneous, object space), the ray of pixel (px, py) starts at
  // near = gn.xyz / gn.w (g0 = near0) and ends at far = gf.xyz / gf.w (g0 =
  // far0), both the index->object map of the raymarch at depth 0.5 and
  // S - 1 + 0.5 (for S > 1)
  float near0[4], far0[4], gx[4], gy[4];
  int out_slot;
};
struct ProjLaunch {
  ProjView view[kMaxViewsPerLaunch];
  int num_views;
  float inv_zmax;  // 1 / max(S - 1, 1): step = (far - near) * inv_zmax
  int accumulate;  // 0: write the sum, 1: add it to the existing gradient
};

// Per-sample values of the BORDER density render are stored wave-major: each
// 16 x 4-pixel wave of a 16 x 16-pixel tile owns S consecutive runs of 64
// floats, one run per sample (lane = 16 (y % 4) + x % 16), so that the
// forward's one-sample store of a wave fills two whole 128-byte lines.
constexpr int kSampleRun = 64;
__host__ __device__ inline size_t render_waves_per_view(int out_w, int out_h) {
  return (size_t)((out_w + 15) / 16) * (size_t)((out_h + 15) / 16) * 4;
}
// offset of sample 0 of pixel (x, y) of output slot `slot`; sample i is at + i * kSampleRun
__host__ __device__ inline size_t sample_base(int out_w, int out_h, int S, int slot, int x, int y) {
  const size_t tiles_x = (size_t)((out_w + 15) / 16);
  const size_t wave = ((size_t)(y >> 4) * tiles_x + (size_t)(x >> 4)) * 4 + (size_t)((y & 15) >> 2);
  return ((size_t)slot * render_waves_per_view(out_w, out_h) + wave) * (size_t)S * kSampleRun +
         (size_t)(((y & 3) << 4) | (x & 15));
}

// What the BORDER density render forward saves for its backward (device
// pointers, indexed by output view slot): per-sample values (wave-major, above),
// per-pixel records [views][H][W] (sample range lo, hi as int bits in x, y;
// coefficients in z, w) and per-tile partial sums of the L2 loss
// [views][tiles].
struct RenderSaved {
  float* samples;
  float4* records;
  float* partials;
  float* samples2;   // the lit render's second per-sample array (u_i), else null
  int n_partials;    // loss partials: tiles of every view
};

// Fused advection launch parameters.
struct AdvectDims {
  Dims3 d;
  float dt;
  int order;   // 1 = semi-Lagrangian, 2 = MacCormack
  int clamp;   // 0 none, 1 MC_SMOOTH, 2 NEGATIVE
};

// Fused density render parameters.
enum LossKind { LOSS_SE = 0, LOSS_AE = 1 };
struct RenderParams {
  float light[3];  // light colour * intensity
  float step;      // camera depth step (renderer.py:420-437)
  float loss_scale;
  int loss_kind;       // LossKind of the folded image loss
  const float* bkg;    // [V,H,W,3] background composited as rgb + bkg * exp(-depth), or null
  float ambient;       // lit render: ambient intensity relative to `light` (the colour)
};

// Shape of one raymarch launch.
struct RayDims {
  Dims3 in;       // W, H, D of the sampled grid
  int C;
  int out_w, out_h, out_d;  // pixels and depth samples
  size_t in_batch_stride;   // floats per input batch
  size_t out_view_stride;   // floats per output view slot
};


// ------------------------------------------------ resampling (v1 sampler)
// Mip pyramid of a batch of grids. Level 0 is the grid itself; levels 1..L
// live in an atlas, one block of levels per batch (MipAtlas3D,
// resample_grid.cu.cc:104-236, without its pinned-host pointer table).
constexpr int kMaxMips = 15;
struct MipTable {
  int levels;                    // levels above the base (0 = no mips)
  int C;
  Dims3 dims[kMaxMips + 1];      // W, H, D per level
  size_t offset[kMaxMips + 1];   // float offset of level m >= 1 in a batch's atlas block
  size_t atlas_stride;           // floats per batch in the atlas
  size_t base_stride;            // floats per batch of the base grid
};

// Sampler settings of the resample/LuT ops (SamplerSettings, sampling.hpp:496-537).
struct SampleCfg {
  int fm;        // Filter
  int mm;        // Mip
  int levels;    // usable mip levels
  float bias;    // mip bias
  float offset;  // cell-centre offset subtracted before the tap
};

// One (batch, camera) view of a transform resample. Forward coordinate modes
// (idx -> object space) use A = P^-1, B = (V M)^-1; reverse modes (object
// space -> idx) use A = V M, B = P. p22/p32 are P's entries (col 2, row 2)
// and (col 3, row 2) for the linear-depth remap.
struct SampleView {
  float A[16];
  float B[16];
  float p22, p32, nearp, farp;
  int in_batch;
  int out_slot;
};
constexpr int kMaxSampleViews = 8;
struct SampleLaunch {
  SampleView view[kMaxSampleViews];
  int num_views;
  int cm;          // Coord
  int accumulate;  // sum-over-views launch: add to the existing output
};

// One SampleGridLut launch: lookup-table shape and flags; blockIdx.y walks
// the (batch, table) pairs starting at batch0.
struct LutLaunch {
  Dims3 dout;     // lookup table / output cells
  int num_cams;   // lookup tables
  int global;     // separate_camera_batch
  int relative;
  int normalized;
  int batch0;
};

// ------------------------------------------------------ kernel timers
// Optional HIP-event timing of named launches (gft_timer_* in the C ABI).
// A scope records an event before and after the launches it encloses when
// its name is being timed on the calling thread; otherwise it does nothing.
bool timer_active(const char* name);
void timer_record(hipStream_t s, const char* name, bool begin);
struct KernelTimer {
  hipStream_t s;
  const char* name;
  bool on;
  KernelTimer(hipStream_t s_, const char* n) : s(s_), name(n), on(timer_active(n)) {
    if (on) timer_record(s, name, true);
  }
  ~KernelTimer() {
    if (on) timer_record(s, name, false);
  }
};

// idx -> ndc -> view (divide by w) -> object space (transformations_v2.hpp:166-179)
__device__ __forceinline__ float3 idx_to_os(const RayView& v, float ix, float iy, float iz, float dinvx, float dinvy,
                                            float dinvz) {
  float4 vs = mat_vec(v.pinv, ix * dinvx * 2.f - 1.f, iy * dinvy * 2.f - 1.f, iz * dinvz * 2.f - 1.f, 1.f);
  float iw = 1.f / vs.w;
  float4 os = mat_vec(v.mvinv, vs.x * iw, vs.y * iw, vs.z * iw, 1.f);
  return make_float3(os.x, os.y, os.z);
}

}  // namespace gft
