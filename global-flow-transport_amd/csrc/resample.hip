// Frustum and lookup-table resampling on gfx950: SampleGridTransform and
// SampleGridLut forward, their gradients (the reference's count-normalised
// scatter and the exact adjoint), mip pyramids and the LoD transform.
//
// One thread per output cell, x fastest so that a wave writes 64 consecutive
// cells; views of one launch are the grid's y dimension and their camera
// records travel as kernel arguments (no __constant__ state, re-entrant).
// Sampler semantics follow the v1 sampler of the reference
// (phitest/render/cuda/src/sampling.hpp:158-252,356-491,539-764).
#include "gft_device.hpp"
#include "sampler_v1.hpp"

namespace gft {

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ void unflat(size_t f, Dims3 d, int& x, int& y, int& z) {
  x = (int)(f % (size_t)d.x);
  size_t r = f / (size_t)d.x;
  y = (int)(r % (size_t)d.y);
  z = (int)(r / (size_t)d.y);
}

// v1 MIN/MAX over the 8 corners (read3DMin/read3DMax, sampling.hpp:372-491):
// BORDER outside -> 0; if any corner is invalid the extremum starts at 0;
// floor and ceil corners (ceil == floor at integers). arg receives the flat
// index of the winning corner per channel (-1: the start value won).
template <int C, int BM>
__device__ __forceinline__ Vec<C> read_minmax_v1(const float* __restrict__ buf, Dims3 d, float px, float py, float pz,
                                                 bool is_max, long long* arg) {
  Taps t = make_taps<BM>(px, py, pz, d);
  Vec<C> m;
  if (arg)
#pragma unroll
    for (int ch = 0; ch < C; ++ch) arg[ch] = -1;
  if (BM == B_BORDER && !t.inside) return vzero<C>();
  int ex[2] = {t.ix[0], t.wx == 0.f ? t.ix[0] : t.ix[1]};
  int ey[2] = {t.iy[0], t.wy == 0.f ? t.iy[0] : t.iy[1]};
  int ez[2] = {t.iz[0], t.wz == 0.f ? t.iz[0] : t.iz[1]};
  bool evx[2] = {t.vx[0], t.wx == 0.f ? t.vx[0] : t.vx[1]};
  bool evy[2] = {t.vy[0], t.wy == 0.f ? t.vy[0] : t.vy[1]};
  bool evz[2] = {t.vz[0], t.wz == 0.f ? t.vz[0] : t.vz[1]};
  float init = is_max ? -kFloatMax : kFloatMax;
  if (BM == B_BORDER && !(evx[0] & evx[1] & evy[0] & evy[1] & evz[0] & evz[1])) init = 0.f;
#pragma unroll
  for (int ch = 0; ch < C; ++ch) m.v[ch] = init;
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        if (BM == B_BORDER && !(evx[i] & evy[j] & evz[k])) continue;
        size_t f = flat3(ex[i], ey[j], ez[k], d);
        Vec<C> v = vload<C>(buf, f);
#pragma unroll
        for (int ch = 0; ch < C; ++ch) {
          bool better = is_max ? (v.v[ch] > m.v[ch]) : (v.v[ch] < m.v[ch]);
          if (better) {
            m.v[ch] = v.v[ch];
            if (arg) arg[ch] = (long long)f;
          }
        }
      }
  return m;
}

// One tap of the v1 sampler at a position with the offset removed.
template <int C, int BM>
__device__ __forceinline__ Vec<C> fetch_v1(const float* __restrict__ buf, Dims3 d, int fm, float px, float py,
                                           float pz) {
  if (fm == F_LINEAR) return read_linear<C, BM>(buf, d, px, py, pz);
  if (fm == F_NEAREST) return read_nearest<C, BM>(buf, d, px, py, pz);
  return read_minmax_v1<C, BM>(buf, d, px, py, pz, fm == F_MAX, nullptr);
}

template <int C, int BM>
__device__ __forceinline__ Vec<C> sample_v1(const float* base, const float* atlas, const MipTable& mt, int b,
                                            const SampleCfg& s, float3 p, float lod) {
  int lv[2];
  float w[2];
  int n = level_taps(s, lod, lv, w);
  float sc = 1.f / (float)(1 << lv[0]);
  Vec<C> a = fetch_v1<C, BM>(level_ptr(base, atlas, mt, b, lv[0]), mt.dims[lv[0]], s.fm, p.x * sc - s.offset,
                             p.y * sc - s.offset, p.z * sc - s.offset);
  if (n == 2) {
    float sc1 = 1.f / (float)(1 << lv[1]);
    Vec<C> c = fetch_v1<C, BM>(level_ptr(base, atlas, mt, b, lv[1]), mt.dims[lv[1]], s.fm, p.x * sc1 - s.offset,
                               p.y * sc1 - s.offset, p.z * sc1 - s.offset);
#pragma unroll
    for (int ch = 0; ch < C; ++ch) a.v[ch] = lerpf(a.v[ch], c.v[ch], w[1]);
  }
  return a;
}

// ---------------------------------------------------------------- mips
// kGenerateMip3D (resample_grid.cu.cc:46-72): pairwise means in z, then y,
// then x. One thread per coarse voxel; blockIdx.y = batch. Every source value
// is multiplied by `scale` first (1 leaves it unchanged): the pyramid of a
// scaled grid without materialising the scaled grid.
template <int C>
__global__ __launch_bounds__(256) void k_mip_down(const float* __restrict__ src, size_t src_stride, Dims3 s,
                                                  float* __restrict__ dst, size_t dst_stride, Dims3 o, float scale) {
#pragma clang fp contract(off)
  size_t n = (size_t)o.x * o.y * o.z;
  size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const float* in = src + (size_t)blockIdx.y * src_stride;
  float* out = dst + (size_t)blockIdx.y * dst_stride;
  int x, y, z;
  unflat(f, o, x, y, z);
  Vec<C> ya[2];
#pragma unroll
  for (int xi = 0; xi < 2; ++xi) {
    Vec<C> za[2];
#pragma unroll
    for (int yi = 0; yi < 2; ++yi) {
      Vec<C> a = vload<C>(in, flat3(2 * x + xi, 2 * y + yi, 2 * z, s));
      Vec<C> b = vload<C>(in, flat3(2 * x + xi, 2 * y + yi, 2 * z + 1, s));
#pragma unroll
      for (int ch = 0; ch < C; ++ch) za[yi].v[ch] = (a.v[ch] * scale + b.v[ch] * scale) * 0.5f;
    }
#pragma unroll
    for (int ch = 0; ch < C; ++ch) ya[xi].v[ch] = (za[0].v[ch] + za[1].v[ch]) * 0.5f;
  }
  Vec<C> r;
#pragma unroll
  for (int ch = 0; ch < C; ++ch) r.v[ch] = (ya[0].v[ch] + ya[1].v[ch]) * 0.5f;
  vstore<C>(out, f, r);
}

// Adjoint of one mip step: every fine voxel under a coarse voxel receives 1/8
// of its gradient (gather form, no atomics).
template <int C>
__global__ __launch_bounds__(256) void k_mip_collapse(const float* __restrict__ gparent, size_t parent_stride, Dims3 o,
                                                      float* __restrict__ gchild, size_t child_stride, Dims3 s) {
  size_t n = (size_t)s.x * s.y * s.z;
  size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  int x, y, z;
  unflat(f, s, x, y, z);
  int px = x >> 1, py = y >> 1, pz = z >> 1;
  if (px >= o.x || py >= o.y || pz >= o.z) return;
  const float* gp = gparent + (size_t)blockIdx.y * parent_stride;
  float* gc = gchild + (size_t)blockIdx.y * child_stride;
  Vec<C> g = vload<C>(gp, flat3(px, py, pz, o));
  Vec<C> c = vload<C>(gc, f);
#pragma unroll
  for (int ch = 0; ch < C; ++ch) c.v[ch] += 0.125f * g.v[ch];
  vstore<C>(gc, f, c);
}

// ------------------------------------------------------ transform forward
// kSample3D (resample_grid.cu.cc:240-252). blockIdx.y = view of the launch.
// SUM: one thread loops over the launch's views and writes (or adds) the sum
// into slot 0, the view-summed gather of the production sampling gradient.
template <int C, int BM, bool SUM>
__global__ __launch_bounds__(256) void k_sample_transform(const float* __restrict__ in,
                                                          const float* __restrict__ atlas, MipTable mt, SampleCfg s,
                                                          Dims3 dout, SampleLaunch L, float* __restrict__ out) {
  size_t n = (size_t)dout.x * dout.y * dout.z;
  size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  int x, y, z;
  unflat(f, dout, x, y, z);
  const float3 oinv = make_float3(1.f / dout.x, 1.f / dout.y, 1.f / dout.z);
  const float3 idims = make_float3((float)mt.dims[0].x, (float)mt.dims[0].y, (float)mt.dims[0].z);
  const bool lod = s.mm != MIP_NONE;
  if constexpr (SUM) {
    Vec<C> acc = vzero<C>();
    for (int vi = 0; vi < L.num_views; ++vi) {
      const SampleView& v = L.view[vi];
      float l;
      float3 p = sample_pos(v, L.cm, lod, x + 0.5f, y + 0.5f, z + 0.5f, oinv, idims, l);
      Vec<C> r = sample_v1<C, BM>(in, atlas, mt, v.in_batch, s, p, l);
#pragma unroll
      for (int ch = 0; ch < C; ++ch) acc.v[ch] += r.v[ch];
    }
    if (L.accumulate) {
      Vec<C> o = vload<C>(out, f);
#pragma unroll
      for (int ch = 0; ch < C; ++ch) acc.v[ch] += o.v[ch];
    }
    vstore<C>(out, f, acc);
  } else {
    const SampleView& v = L.view[blockIdx.y];
    float l;
    float3 p = sample_pos(v, L.cm, lod, x + 0.5f, y + 0.5f, z + 0.5f, oinv, idims, l);
    Vec<C> r = sample_v1<C, BM>(in, atlas, mt, v.in_batch, s, p, l);
    vstore<C>(out + (size_t)v.out_slot * n * C, f, r);
  }
}

// ------------------------------------------------------ gradient scatters
// Reference scatter (scatterGradInterpolated, resample_grid.cu.cc:761-812,
// sampling_v2.hpp:459-486): BORDER accepts p in (-1, dims), floor and ceil
// corners, each valid corner counted once per visit.
template <int C, int BM>
__device__ __forceinline__ void scatter_reference(float* __restrict__ grad, unsigned* __restrict__ cnt, Dims3 d,
                                                  float px, float py, float pz, const Vec<C>& g) {
  if (BM == B_BORDER) {
    if (!((-1.f < px) & (px < (float)d.x) & (-1.f < py) & (py < (float)d.y) & (-1.f < pz) & (pz < (float)d.z))) return;
  }
  float fxf = floorf(px), fyf = floorf(py), fzf = floorf(pz);
  float cwx = px - fxf, cwy = py - fyf, cwz = pz - fzf;
  int fx = (int)fxf, fy = (int)fyf, fz = (int)fzf;
  int cx = (int)ceilf(px), cy = (int)ceilf(py), cz = (int)ceilf(pz);
  if (BM == B_CLAMP) {
    fx = clampi(fx, 0, d.x - 1); cx = clampi(cx, 0, d.x - 1);
    fy = clampi(fy, 0, d.y - 1); cy = clampi(cy, 0, d.y - 1);
    fz = clampi(fz, 0, d.z - 1); cz = clampi(cz, 0, d.z - 1);
  } else if (BM == B_WRAP) {
    fx = pmodi(fx, d.x); cx = pmodi(cx, d.x);
    fy = pmodi(fy, d.y); cy = pmodi(cy, d.y);
    fz = pmodi(fz, d.z); cz = pmodi(cz, d.z);
  }
  int ix[2] = {fx, cx}, iy[2] = {fy, cy}, iz[2] = {fz, cz};
  float wx[2] = {1.f - cwx, cwx}, wy[2] = {1.f - cwy, cwy}, wz[2] = {1.f - cwz, cwz};
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        if (BM == B_BORDER &&
            !((ix[i] >= 0) & (ix[i] < d.x) & (iy[j] >= 0) & (iy[j] < d.y) & (iz[k] >= 0) & (iz[k] < d.z)))
          continue;
        size_t f = flat3(ix[i], iy[j], iz[k], d);
        vatomic_add<C>(grad, f, g, wx[i] * wy[j] * wz[k]);
        if (cnt) atomicAdd(cnt + f, 1u);
      }
}

// Exact adjoint of one v1 tap: trilinear with the forward read's domain and
// corners, nearest, or the winning corner of MIN/MAX.
template <int C, int BM>
__device__ __forceinline__ void fetch_v1_adjoint(const float* __restrict__ buf, float* __restrict__ grad, Dims3 d,
                                                 int fm, float px, float py, float pz, const Vec<C>& g, float scale) {
  if (fm == F_LINEAR) {
    Taps t = make_taps<BM>(px, py, pz, d);
    if (BM == B_BORDER && !t.inside) return;
    float wx[2] = {1.f - t.wx, t.wx}, wy[2] = {1.f - t.wy, t.wy}, wz[2] = {1.f - t.wz, t.wz};
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          if (BM == B_BORDER && !(t.vx[i] & t.vy[j] & t.vz[k])) continue;
          float w = wx[i] * wy[j] * wz[k];
          if (w == 0.f) continue;
          vatomic_add<C>(grad, flat3(t.ix[i], t.iy[j], t.iz[k], d), g, w * scale);
        }
  } else if (fm == F_NEAREST) {
    size_t f;
    if (nearest_index<BM>(px, py, pz, d, f)) vatomic_add<C>(grad, f, g, scale);
  } else {
    long long arg[C];
    (void)read_minmax_v1<C, BM>(buf, d, px, py, pz, fm == F_MAX, arg);
#pragma unroll
    for (int ch = 0; ch < C; ++ch)
      if (arg[ch] >= 0) atomic_addf(grad + (size_t)arg[ch] * C + ch, g.v[ch] * scale);
  }
}

// SampleGridTransformGrad, reference flavour: scatter of every view of one
// batch into its gradient and per-voxel counter (kScatter3DGradients,
// resample_grid.cu.cc:841-861). grad/cnt point at the batch.
template <int C, int BM>
__global__ __launch_bounds__(256) void k_sample_transform_scatter_ref(const float* __restrict__ og, Dims3 din,
                                                                      Dims3 dout, SampleLaunch L,
                                                                      float* __restrict__ grad,
                                                                      unsigned* __restrict__ cnt) {
  size_t n = (size_t)dout.x * dout.y * dout.z;
  size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const SampleView& v = L.view[blockIdx.y];
  int x, y, z;
  unflat(f, dout, x, y, z);
  const float3 oinv = make_float3(1.f / dout.x, 1.f / dout.y, 1.f / dout.z);
  const float3 idims = make_float3((float)din.x, (float)din.y, (float)din.z);
  float l;
  float3 p = sample_pos(v, L.cm, false, x + 0.5f, y + 0.5f, z + 0.5f, oinv, idims, l);
  Vec<C> g = vload<C>(og + (size_t)v.out_slot * n * C, f);
  scatter_reference<C, BM>(grad, cnt, din, p.x - 0.5f, p.y - 0.5f, p.z - 0.5f, g);
}

// Exact adjoint of the transform forward, mips included: level m taps land
// in the level-m gradient (level 0 = the input gradient).
template <int C, int BM>
__global__ __launch_bounds__(256) void k_sample_transform_adjoint(const float* __restrict__ in,
                                                                  const float* __restrict__ atlas, MipTable mt,
                                                                  SampleCfg s, Dims3 dout, SampleLaunch L,
                                                                  const float* __restrict__ og,
                                                                  float* __restrict__ grad,
                                                                  float* __restrict__ gatlas) {
  size_t n = (size_t)dout.x * dout.y * dout.z;
  size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const SampleView& v = L.view[blockIdx.y];
  int x, y, z;
  unflat(f, dout, x, y, z);
  const float3 oinv = make_float3(1.f / dout.x, 1.f / dout.y, 1.f / dout.z);
  const float3 idims = make_float3((float)mt.dims[0].x, (float)mt.dims[0].y, (float)mt.dims[0].z);
  float l;
  float3 p = sample_pos(v, L.cm, s.mm != MIP_NONE, x + 0.5f, y + 0.5f, z + 0.5f, oinv, idims, l);
  Vec<C> g = vload<C>(og + (size_t)v.out_slot * n * C, f);
  int lv[2];
  float w[2];
  int nt = level_taps(s, l, lv, w);
  for (int t = 0; t < nt; ++t) {
    int m = lv[t];
    float sc = 1.f / (float)(1 << m);
    fetch_v1_adjoint<C, BM>(level_ptr(in, atlas, mt, v.in_batch, m), level_ptr_mut(grad, gatlas, mt, v.in_batch, m),
                            mt.dims[m], s.fm, p.x * sc - s.offset, p.y * sc - s.offset, p.z * sc - s.offset, g, w[t]);
  }
}

// ----------------------------------------- gradient gather (BORDER, no mips)
// SampleGridTransformGrad in gather form: one block per kGBX x kGBY x kGBZ
// voxel block of one batch. For each view the block bounds the frustum cells
// whose sample can reach its voxels: the corners of the position box
// [x0-1, x0+kGBX] x ... (every tap of a sample inside it, floor/ceil or
// floor/floor+1, lies in the block only if the sample does) are mapped back
// to frustum index space with the inverse coordinate mode (a projective map,
// so the image of the box lies in the box of its corner images; one cell of
// margin covers rounding), and the block walks those cells, recomputes each
// sample position exactly as the forward op does and adds the contributions
// of the taps that fall in the block to FP64 LDS accumulators (plus the
// per-voxel visit count of the REFERENCE flavour). A corner image behind the
// camera widens the walk to the whole frustum. No global atomics; one store
// per voxel: the REFERENCE flavour stores sum / count
// (kNormalize3DGradients), EXACT stores or adds the sum.
constexpr int kGBX = 16, kGBY = 8, kGBZ = 4, kGBVox = kGBX * kGBY * kGBZ;

// frustum-space bounds [lo, hi] (cell indices) of the cells whose sample can
// land in the position box b0..b1 (vinv: the view in the reverse coordinate
// mode, grid position -> frustum position, as map_pos computes it). The x and
// y images are projective in the position (extremes at the corners while
// every corner is in front of the camera, w > 0); the linear-depth image
// -p32 / (ndz + p22) is monotonic on each side of its pole, and frustum cells
// only produce depths on the negative side: all corners there -> the corner
// hull, all on the positive side -> no cell, straddling the pole (possible
// for orthographic projections) -> every depth. Out-of-front corners or NaN
// widen to every cell.
__device__ __forceinline__ void gather_cell_bounds(const SampleView& vinv, int cm_inv, float3 b0, float3 b1,
                                                   float3 idims_out, Dims3 dout, int lo[3], int hi[3]) {
  float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
  bool ok = true;
  int neg = 0, pos = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    // positions are in sample units; the coordinate maps take cell-centre units (+0.5)
    const float qx = ((k & 1) ? b1.x : b0.x) + 0.5f, qy = ((k & 2) ? b1.y : b0.y) + 0.5f,
                qz = ((k & 4) ? b1.z : b0.z) + 0.5f;
    const float4 vs = mat_vec(vinv.A, qx, qy, qz, 1.f);
    const float4 ndc = mat_vec(vinv.B, vs.x, vs.y, vs.z, vs.w);
    ok &= ndc.w > 0.f;
    const float ndx = ndc.x / ndc.w, ndy = ndc.y / ndc.w, ndz = ndc.z / ndc.w;
    float r[3] = {(ndx * 0.5f + 0.5f) * idims_out.x, (ndy * 0.5f + 0.5f) * idims_out.y,
                  (ndz * 0.5f + 0.5f) * idims_out.z};
    if (cm_inv == CM_LINDEPTH_REVERSE) {
      const float z = -vinv.p32 / (ndz + vinv.p22);
      neg += z < 0.f;
      pos += z > 0.f;
      r[2] = -(z + vinv.nearp) / (vinv.farp - vinv.nearp) * idims_out.z;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      ok &= r[a] == r[a];
      mn[a] = fminf(mn[a], r[a]);
      mx[a] = fmaxf(mx[a], r[a]);
    }
  }
  const int dd[3] = {dout.x, dout.y, dout.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    // cell c has its centre at c + 0.5
    lo[a] = (int)fmaxf(ceilf(fmaxf(mn[a], -4.f) - 0.5f) - 1.f, 0.f);
    hi[a] = (int)fminf(floorf(fminf(mx[a], (float)dd[a] + 4.f) - 0.5f) + 1.f, (float)(dd[a] - 1));
  }
  if (cm_inv == CM_LINDEPTH_REVERSE && neg + pos < 8) ok = false;  // a corner on the pole
  if (cm_inv == CM_LINDEPTH_REVERSE && ok && pos == 8) {
    lo[2] = 1;  // no frustum cell has a sample there
    hi[2] = 0;
  } else if (cm_inv == CM_LINDEPTH_REVERSE && ok && pos > 0) {
    lo[2] = 0;
    hi[2] = dd[2] - 1;
  }
  if (!ok) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = 0;
      hi[a] = dd[a] - 1;
    }
  }
}

template <int C, bool REF, int FM>
__global__ __launch_bounds__(256) void k_sample_transform_gather_adj(const float* __restrict__ og, Dims3 din,
                                                                     Dims3 dout, SampleLaunch L, SampleLaunch Linv,
                                                                     int accumulate, float* __restrict__ grad) {
  __shared__ double acc[kGBVox * C];
  __shared__ unsigned cnt[REF ? kGBVox : 1];
  typedef __attribute__((address_space(3))) double lds_double;
  typedef __attribute__((address_space(3))) unsigned lds_uint;
  lds_double* a = (lds_double*)acc;
  lds_uint* cn = (lds_uint*)cnt;
  const int nbx = (din.x + kGBX - 1) / kGBX, nby = (din.y + kGBY - 1) / kGBY;
  const int x0 = (blockIdx.x % nbx) * kGBX, y0 = ((blockIdx.x / nbx) % nby) * kGBY,
            z0 = (blockIdx.x / (nbx * nby)) * kGBZ;
  for (int k = threadIdx.x; k < kGBVox * C; k += 256) acc[k] = 0.0;
  if (REF)
    for (int k = threadIdx.x; k < kGBVox; k += 256) cnt[k] = 0u;
  __syncthreads();
  const size_t n = (size_t)dout.x * dout.y * dout.z;
  const float3 oinv = make_float3(1.f / dout.x, 1.f / dout.y, 1.f / dout.z);
  const float3 idims = make_float3((float)din.x, (float)din.y, (float)din.z);
  const float3 idims_out = make_float3((float)dout.x, (float)dout.y, (float)dout.z);
  // nearest truncates p + 0.5 toward zero, so p down to x0 - 1.5 can reach x0 = 0
  const float3 b0 = make_float3((float)x0 - 2.f, (float)y0 - 2.f, (float)z0 - 2.f);
  const float3 b1 = make_float3((float)(x0 + kGBX), (float)(y0 + kGBY), (float)(z0 + kGBZ));
  for (int vi = 0; vi < L.num_views; ++vi) {
    const SampleView& v = L.view[vi];
    int lo[3], hi[3];
    gather_cell_bounds(Linv.view[vi], Linv.cm, b0, b1, idims_out, dout, lo, hi);
    if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) continue;
    const unsigned nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
    const unsigned cells = nx * ny * (unsigned)(hi[2] - lo[2] + 1);  // the caller keeps n below 2^31
    const float* ogv = og + (size_t)v.out_slot * n * C;
    for (unsigned k = threadIdx.x; k < cells; k += 256) {
      const unsigned r = k / nx;
      const int cx = lo[0] + (int)(k - r * nx), cy = lo[1] + (int)(r % ny), cz = lo[2] + (int)(r / ny);
      float l;
      const float3 q = sample_pos(v, L.cm, false, cx + 0.5f, cy + 0.5f, cz + 0.5f, oinv, idims, l);
      const float px = q.x - 0.5f, py = q.y - 0.5f, pz = q.z - 0.5f;
      if (!((px >= b0.x) & (px <= b1.x) & (py >= b0.y) & (py <= b1.y) & (pz >= b0.z) & (pz <= b1.z))) continue;
      const Vec<C> g = vload<C>(ogv, ((size_t)cz * dout.y + cy) * dout.x + cx);
      int ix[2], iy[2], iz[2];
      float wx[2], wy[2], wz[2];
      bool take = true;
      if (REF) {
        // scatterGradInterpolated: p in (-1, dims), floor and ceil corners
        take = (-1.f < px) & (px < (float)din.x) & (-1.f < py) & (py < (float)din.y) & (-1.f < pz) &
               (pz < (float)din.z);
        const float fxf = floorf(px), fyf = floorf(py), fzf = floorf(pz);
        ix[0] = (int)fxf; ix[1] = (int)ceilf(px);
        iy[0] = (int)fyf; iy[1] = (int)ceilf(py);
        iz[0] = (int)fzf; iz[1] = (int)ceilf(pz);
        wx[1] = px - fxf; wx[0] = 1.f - wx[1];
        wy[1] = py - fyf; wy[0] = 1.f - wy[1];
        wz[1] = pz - fzf; wz[0] = 1.f - wz[1];
      } else if (FM == F_LINEAR) {
        const Taps t = make_taps<B_BORDER>(px, py, pz, din);
        take = t.inside;
        ix[0] = t.vx[0] ? t.ix[0] : -1000; ix[1] = t.vx[1] ? t.ix[1] : -1000;
        iy[0] = t.vy[0] ? t.iy[0] : -1000; iy[1] = t.vy[1] ? t.iy[1] : -1000;
        iz[0] = t.vz[0] ? t.iz[0] : -1000; iz[1] = t.vz[1] ? t.iz[1] : -1000;
        wx[0] = 1.f - t.wx; wx[1] = t.wx;
        wy[0] = 1.f - t.wy; wy[1] = t.wy;
        wz[0] = 1.f - t.wz; wz[1] = t.wz;
      } else {
        size_t f;
        take = nearest_index<B_BORDER>(px, py, pz, din, f);
        ix[0] = ix[1] = (int)(px + 0.5f);
        iy[0] = iy[1] = (int)(py + 0.5f);
        iz[0] = iz[1] = (int)(pz + 0.5f);
        wx[0] = wy[0] = wz[0] = 1.f;
        wx[1] = wy[1] = wz[1] = 0.f;
      }
      if (!take) continue;
#pragma unroll
      for (int kk = 0; kk < 2; ++kk)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            if (FM == F_NEAREST && !REF && (i | j | kk)) continue;
            const int lx = ix[i] - x0, ly = iy[j] - y0, lz = iz[kk] - z0;
            if (!(((unsigned)lx < (unsigned)kGBX) & ((unsigned)ly < (unsigned)kGBY) & ((unsigned)lz < (unsigned)kGBZ)))
              continue;
            // the reference also drops corners outside the grid (BORDER)
            if (REF && !((ix[i] < din.x) & (iy[j] < din.y) & (iz[kk] < din.z))) continue;
            if (!REF && FM == F_LINEAR && !((ix[i] < din.x) & (iy[j] < din.y) & (iz[kk] < din.z))) continue;
            const float w = wx[i] * wy[j] * wz[kk];
            if (!REF && FM == F_LINEAR && w == 0.f) continue;
            const int vox = (lz * kGBY + ly) * kGBX + lx;
#pragma unroll
            for (int ch = 0; ch < C; ++ch)
              __hip_atomic_fetch_add(a + vox * C + ch, (double)(g.v[ch] * w), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP);
            if (REF) __hip_atomic_fetch_add(cn + vox, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kGBVox; k += 256) {
    const int x = x0 + k % kGBX, y = y0 + (k / kGBX) % kGBY, z = z0 + k / (kGBX * kGBY);
    if (x >= din.x || y >= din.y || z >= din.z) continue;
    const size_t f = flat3(x, y, z, din);
    Vec<C> r;
    if (REF) {
      const unsigned c = cnt[k];
      const float inv = c ? 1.f / (float)c : 0.f;
#pragma unroll
      for (int ch = 0; ch < C; ++ch) r.v[ch] = c ? (float)acc[k * C + ch] * inv : 0.f;
    } else {
      Vec<C> old = accumulate ? vload<C>(grad, f) : vzero<C>();
#pragma unroll
      for (int ch = 0; ch < C; ++ch) r.v[ch] = old.v[ch] + (float)acc[k * C + ch];
    }
    vstore<C>(grad, f, r);
  }
}

// kNormalize3DGradients (resample_grid.cu.cc:814-832): divide by the count
// where it is positive.
template <int C>
__global__ __launch_bounds__(256) void k_normalize_count(float* __restrict__ grad, const unsigned* __restrict__ cnt,
                                                         size_t nvox) {
  size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nvox) return;
  unsigned c = cnt[f];
  if (c == 0) return;
  Vec<C> g = vload<C>(grad, f);
  float inv = 1.f / (float)c;
#pragma unroll
  for (int ch = 0; ch < C; ++ch) g.v[ch] = g.v[ch] * inv;
  vstore<C>(grad, f, g);
}

// ------------------------------------------------------------------- LuT

// kSampleGrid3DLuT (resample_grid.cu.cc:465-484): position = lut.xyz
// (* dims if normalised) (+ idx if relative), lod = lut.w.
__device__ __forceinline__ float3 lut_pos(const float* __restrict__ lut, const LutLaunch& L, int cam, size_t f, int x,
                                          int y, int z, float& lod) {
  size_t n = (size_t)L.dout.x * L.dout.y * L.dout.z;
  float4 e = *reinterpret_cast<const float4*>(lut + ((size_t)cam * n + f) * 4);
  float3 p = make_float3(e.x, e.y, e.z);
  if (L.normalized) p = make_float3(p.x * L.dout.x, p.y * L.dout.y, p.z * L.dout.z);
  if (L.relative) p = make_float3(p.x + x, p.y + y, p.z + z);
  lod = e.w;
  return p;
}

// blockIdx.y = (batch - batch0) * views_per_batch + view.
template <int C, int BM>
__global__ __launch_bounds__(256) void k_sample_lut(const float* __restrict__ in, const float* __restrict__ atlas,
                                                    MipTable mt, SampleCfg s, LutLaunch L,
                                                    const float* __restrict__ lut, float* __restrict__ out) {
  const Dims3 dout = L.dout;
  size_t n = (size_t)dout.x * dout.y * dout.z;
  size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  int per = L.global ? L.num_cams : 1;
  int b = L.batch0 + (int)blockIdx.y / per;
  int cam = L.global ? (int)blockIdx.y % per : b;
  int slot = b * per + (L.global ? cam : 0);
  int x, y, z;
  unflat(f, dout, x, y, z);
  float lod;
  float3 p = lut_pos(lut, L, cam, f, x, y, z, lod);
  Vec<C> r = sample_v1<C, BM>(in, atlas, mt, b, s, p, lod);
  vstore<C>(out + (size_t)slot * n * C, f, r);
}

// Spatial derivative of the trilinear read (read3DGrad, sampling.hpp:257-352).
// EXACT = false: floor/ceil corners (zero derivative at integer positions);
// EXACT = true: floor/floor+1, the right derivative. Same boundary rules.
template <int C, int BM, bool EXACT>
__device__ __forceinline__ void read_grad(const float* __restrict__ buf, Dims3 d, float px, float py, float pz,
                                          Vec<C>& gx, Vec<C>& gy, Vec<C>& gz) {
  gx = vzero<C>();
  gy = vzero<C>();
  gz = vzero<C>();
  float fxf = floorf(px), fyf = floorf(py), fzf = floorf(pz);
  float wx = px - fxf, wy = py - fyf, wz = pz - fzf;
  int fx = (int)fxf, fy = (int)fyf, fz = (int)fzf;
  int cx = EXACT ? fx + 1 : (int)ceilf(px), cy = EXACT ? fy + 1 : (int)ceilf(py), cz = EXACT ? fz + 1 : (int)ceilf(pz);
  bool vx[2] = {true, true}, vy[2] = {true, true}, vz[2] = {true, true};
  if (BM == B_BORDER) {
    float qx = px + 0.5f, qy = py + 0.5f, qz = pz + 0.5f;
    if (!((0.f <= qx) & (qx < (float)d.x) & (0.f <= qy) & (qy < (float)d.y) & (0.f <= qz) & (qz < (float)d.z))) return;
    vx[0] = (fx > -1) & (fx < d.x); vx[1] = (cx > -1) & (cx < d.x);
    vy[0] = (fy > -1) & (fy < d.y); vy[1] = (cy > -1) & (cy < d.y);
    vz[0] = (fz > -1) & (fz < d.z); vz[1] = (cz > -1) & (cz < d.z);
    fx = clampi(fx, 0, d.x - 1); cx = clampi(cx, 0, d.x - 1);
    fy = clampi(fy, 0, d.y - 1); cy = clampi(cy, 0, d.y - 1);
    fz = clampi(fz, 0, d.z - 1); cz = clampi(cz, 0, d.z - 1);
  } else if (BM == B_CLAMP) {
    fx = clampi(fx, 0, d.x - 1); cx = clampi(cx, 0, d.x - 1);
    fy = clampi(fy, 0, d.y - 1); cy = clampi(cy, 0, d.y - 1);
    fz = clampi(fz, 0, d.z - 1); cz = clampi(cz, 0, d.z - 1);
  } else {
    fx = pmodi(fx, d.x); cx = pmodi(cx, d.x);
    fy = pmodi(fy, d.y); cy = pmodi(cy, d.y);
    fz = pmodi(fz, d.z); cz = pmodi(cz, d.z);
  }
  int ix[2] = {fx, cx}, iy[2] = {fy, cy}, iz[2] = {fz, cz};
  Vec<C> v[2][2][2];
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        v[k][j][i] = vload<C>(buf, flat3(ix[i], iy[j], iz[k], d));
        if (BM == B_BORDER && !(vx[i] & vy[j] & vz[k])) v[k][j][i] = vzero<C>();
      }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    gx.v[c] = lerpf(lerpf(v[0][0][1].v[c] - v[0][0][0].v[c], v[0][1][1].v[c] - v[0][1][0].v[c], wy),
                    lerpf(v[1][0][1].v[c] - v[1][0][0].v[c], v[1][1][1].v[c] - v[1][1][0].v[c], wy), wz);
    gy.v[c] = lerpf(lerpf(v[0][1][0].v[c] - v[0][0][0].v[c], v[0][1][1].v[c] - v[0][0][1].v[c], wx),
                    lerpf(v[1][1][0].v[c] - v[1][0][0].v[c], v[1][1][1].v[c] - v[1][0][1].v[c], wx), wz);
    gz.v[c] = lerpf(lerpf(v[1][0][0].v[c] - v[0][0][0].v[c], v[1][0][1].v[c] - v[0][0][1].v[c], wx),
                    lerpf(v[1][1][0].v[c] - v[0][1][0].v[c], v[1][1][1].v[c] - v[0][1][1].v[c], wx), wy);
  }
}

// SampleGridLutGrad for the views of one data batch (blockIdx.y = view).
// EXACT = false: kScatterLuT3DGradients (resample_grid.cu.cc:1051-1081): lut
// gradient written per view (not scaled for normalised coordinates), input
// gradient as a counted trilinear scatter whatever the filter.
// EXACT = true: the lut gradient is added (one thread per lut cell, so no
// atomics), scaled by the output dims for normalised coordinates; the input
// gradient is the adjoint of the filter.
template <int C, int BM, bool EXACT>
__global__ __launch_bounds__(256) void k_sample_lut_bwd(const float* __restrict__ in_b, int fm, float offset,
                                                        LutLaunch L, const float* __restrict__ lut,
                                                        const float* __restrict__ og, Dims3 din,
                                                        float* __restrict__ grad_b, unsigned* __restrict__ cnt,
                                                        float* __restrict__ lut_grad) {
  const Dims3 dout = L.dout;
  size_t n = (size_t)dout.x * dout.y * dout.z;
  size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  int per = L.global ? L.num_cams : 1;
  int b = L.batch0;
  int cam = L.global ? (int)blockIdx.y : b;
  int slot = b * per + (L.global ? cam : 0);
  int x, y, z;
  unflat(f, dout, x, y, z);
  float lod;
  float3 p = lut_pos(lut, L, cam, f, x, y, z, lod);
  Vec<C> g = vload<C>(og + (size_t)slot * n * C, f);
  float px = p.x - offset, py = p.y - offset, pz = p.z - offset;
  if (fm == F_LINEAR) {
    Vec<C> gx, gy, gz;
    read_grad<C, BM, EXACT>(in_b, din, px, py, pz, gx, gy, gz);
    float lx = 0.f, ly = 0.f, lz = 0.f;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      lx += g.v[c] * gx.v[c];
      ly += g.v[c] * gy.v[c];
      lz += g.v[c] * gz.v[c];
    }
    float4* dst = reinterpret_cast<float4*>(lut_grad + ((size_t)cam * n + f) * 4);
    if (EXACT) {
      if (L.normalized) {
        lx *= dout.x;
        ly *= dout.y;
        lz *= dout.z;
      }
      float4 o = *dst;
      *dst = make_float4(o.x + lx, o.y + ly, o.z + lz, o.w);
    } else {
      *dst = make_float4(lx, ly, lz, 0.f);
    }
  }
  if (EXACT) fetch_v1_adjoint<C, BM>(in_b, grad_b, din, fm, px, py, pz, g, 1.f);
  else scatter_reference<C, BM>(grad_b, cnt, din, px, py, pz, g);
}

// Tiled form of k_sample_lut_bwd for one-channel grids and relative lookups
// (the advection LuT, whose positions stay near their cells): a block takes a
// kLTX x kLTY x kLTZ tile of lookup cells and accumulates the input-gradient
// taps (and the REFERENCE visit counts) that fall in the tile plus a
// kLHalo-cell margin in FP64 / u32 LDS, then adds each nonzero region cell to
// the global gradient (and counter) with one atomic. Taps outside the region
// go to global atomics directly, so any lookup is handled exactly. Tap
// positions, weights, validity and counts are those of scatter_reference and
// fetch_v1_adjoint; the lookup gradient is written per cell as in
// k_sample_lut_bwd. MIN/MAX taps keep the untiled kernel.
// (a 2-cell margin: 0.122 / 0.080 ms REFERENCE / EXACT at c2 against 0.138 / 0.084
// with 3 cells, whose larger region costs more in flushes than it saves)
constexpr int kLTX = 32, kLTY = 4, kLTZ = 4, kLHalo = 2;
constexpr int kLRX = kLTX + 2 * kLHalo, kLRY = kLTY + 2 * kLHalo, kLRZ = kLTZ + 2 * kLHalo;
constexpr int kLRegion = kLRX * kLRY * kLRZ;

template <int BM, bool EXACT>
__global__ __launch_bounds__(256) void k_sample_lut_bwd_tiled(const float* __restrict__ in_b, int fm, float offset,
                                                              LutLaunch L, const float* __restrict__ lut,
                                                              const float* __restrict__ og, Dims3 din,
                                                              float* __restrict__ grad_b, unsigned* __restrict__ cnt,
                                                              float* __restrict__ lut_grad) {
  __shared__ double acc[kLRegion];
  __shared__ unsigned cts[EXACT ? 1 : kLRegion];
  typedef __attribute__((address_space(3))) double lds_double;
  typedef __attribute__((address_space(3))) unsigned lds_uint;
  lds_double* a = (lds_double*)acc;
  lds_uint* cn = (lds_uint*)cts;
  const Dims3 dout = L.dout;
  const size_t n = (size_t)dout.x * dout.y * dout.z;
  const int ntx = (dout.x + kLTX - 1) / kLTX, nty = (dout.y + kLTY - 1) / kLTY;
  const int x0 = (blockIdx.x % ntx) * kLTX, y0 = ((blockIdx.x / ntx) % nty) * kLTY,
            z0 = (blockIdx.x / (ntx * nty)) * kLTZ;
  const int rx0 = x0 - kLHalo, ry0 = y0 - kLHalo, rz0 = z0 - kLHalo;
  const int per = L.global ? L.num_cams : 1;
  const int b = L.batch0;
  const int cam = L.global ? (int)blockIdx.y : b;
  const int slot = b * per + (L.global ? cam : 0);
  for (int k = threadIdx.x; k < kLRegion; k += 256) {
    acc[k] = 0.0;
    if (!EXACT) cts[k] = 0u;
  }
  __syncthreads();
  // one tap of weight w at grid cell (ix, iy, iz): the region when it is there, global otherwise
  auto emit = [&](int ix, int iy, int iz, float v, bool count) {
    const int lx = ix - rx0, ly = iy - ry0, lz = iz - rz0;
    if (((unsigned)lx < (unsigned)kLRX) & ((unsigned)ly < (unsigned)kLRY) & ((unsigned)lz < (unsigned)kLRZ)) {
      const int r = (lz * kLRY + ly) * kLRX + lx;
      __hip_atomic_fetch_add(a + r, (double)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (count) __hip_atomic_fetch_add(cn + r, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
      const size_t f = flat3(ix, iy, iz, din);
      atomic_addf(grad_b + f, v);
      if (count) atomicAdd(cnt + f, 1u);
    }
  };
  for (int k = threadIdx.x; k < kLTX * kLTY * kLTZ; k += 256) {
    const int x = x0 + k % kLTX, y = y0 + (k / kLTX) % kLTY, z = z0 + k / (kLTX * kLTY);
    if (x >= dout.x || y >= dout.y || z >= dout.z) continue;
    const size_t f = ((size_t)z * dout.y + y) * dout.x + x;
    float lod;
    const float3 p = lut_pos(lut, L, cam, f, x, y, z, lod);
    const float g = og[(size_t)slot * n + f];
    const float px = p.x - offset, py = p.y - offset, pz = p.z - offset;
    if (fm == F_LINEAR) {
      Vec<1> gx, gy, gz;
      read_grad<1, BM, EXACT>(in_b, din, px, py, pz, gx, gy, gz);
      float lx = g * gx.v[0], ly = g * gy.v[0], lz = g * gz.v[0];
      float4* dst = reinterpret_cast<float4*>(lut_grad + ((size_t)cam * n + f) * 4);
      if (EXACT) {
        if (L.normalized) {
          lx *= dout.x;
          ly *= dout.y;
          lz *= dout.z;
        }
        const float4 o = *dst;
        *dst = make_float4(o.x + lx, o.y + ly, o.z + lz, o.w);
      } else {
        *dst = make_float4(lx, ly, lz, 0.f);
      }
    }
    if (EXACT) {
      if (fm == F_LINEAR) {
        const Taps t = make_taps<BM>(px, py, pz, din);
        if (BM == B_BORDER && !t.inside) continue;
        const float wx[2] = {1.f - t.wx, t.wx}, wy[2] = {1.f - t.wy, t.wy}, wz[2] = {1.f - t.wz, t.wz};
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
              if (BM == B_BORDER && !(t.vx[i] & t.vy[j] & t.vz[kk])) continue;
              const float w = wx[i] * wy[j] * wz[kk];
              if (w == 0.f) continue;
              emit(t.ix[i], t.iy[j], t.iz[kk], g * w, false);
            }
      } else {
        size_t fi;
        if (nearest_index<BM>(px, py, pz, din, fi)) {
          const int ix = (int)(fi % (size_t)din.x), iy = (int)((fi / (size_t)din.x) % (size_t)din.y),
                    iz = (int)(fi / ((size_t)din.x * din.y));
          emit(ix, iy, iz, g, false);
        }
      }
    } else {
      // scatterGradInterpolated (as scatter_reference)
      if (BM == B_BORDER &&
          !((-1.f < px) & (px < (float)din.x) & (-1.f < py) & (py < (float)din.y) & (-1.f < pz) & (pz < (float)din.z)))
        continue;
      const float fxf = floorf(px), fyf = floorf(py), fzf = floorf(pz);
      const float cwx = px - fxf, cwy = py - fyf, cwz = pz - fzf;
      int fx = (int)fxf, fy = (int)fyf, fz = (int)fzf;
      int cx = (int)ceilf(px), cy = (int)ceilf(py), cz = (int)ceilf(pz);
      if (BM == B_CLAMP) {
        fx = clampi(fx, 0, din.x - 1); cx = clampi(cx, 0, din.x - 1);
        fy = clampi(fy, 0, din.y - 1); cy = clampi(cy, 0, din.y - 1);
        fz = clampi(fz, 0, din.z - 1); cz = clampi(cz, 0, din.z - 1);
      } else if (BM == B_WRAP) {
        fx = pmodi(fx, din.x); cx = pmodi(cx, din.x);
        fy = pmodi(fy, din.y); cy = pmodi(cy, din.y);
        fz = pmodi(fz, din.z); cz = pmodi(cz, din.z);
      }
      const int ix[2] = {fx, cx}, iy[2] = {fy, cy}, iz[2] = {fz, cz};
      const float wx[2] = {1.f - cwx, cwx}, wy[2] = {1.f - cwy, cwy}, wz[2] = {1.f - cwz, cwz};
#pragma unroll
      for (int kk = 0; kk < 2; ++kk)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            if (BM == B_BORDER &&
                !((ix[i] >= 0) & (ix[i] < din.x) & (iy[j] >= 0) & (iy[j] < din.y) & (iz[kk] >= 0) & (iz[kk] < din.z)))
              continue;
            emit(ix[i], iy[j], iz[kk], g * (wx[i] * wy[j] * wz[kk]), true);
          }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kLRegion; k += 256) {
    const double v = acc[k];
    const unsigned c = EXACT ? 0u : cts[k];
    if (v == 0.0 && c == 0u) continue;
    const int x = rx0 + k % kLRX, y = ry0 + (k / kLRX) % kLRY, z = rz0 + k / (kLRX * kLRY);
    // taps always land inside the grid (BORDER drops, CLAMP/WRAP fold them in)
    if (((unsigned)x < (unsigned)din.x) & ((unsigned)y < (unsigned)din.y) & ((unsigned)z < (unsigned)din.z)) {
      const size_t f = flat3(x, y, z, din);
      if (v != 0.0) atomic_addf(grad_b + f, (float)v);
      if (c) atomicAdd(cnt + f, c);
    }
  }
}

// ------------------------------------------------------------------- LoD
// kComputeLoD3D (resample_grid.cu.cc:1288-1300): position and LoD of every
// output cell (central differences, Jacobian rows). blockIdx.y = camera.
// One block per kLodTX x kLodTY x kLodTZ tile of output cells: the positions
// of the tile and a one-cell margin are mapped once into LDS (about 2 maps per
// cell instead of 7), then every cell takes its central differences from
// there. A neighbour's position is the same map of the same coordinates
// ((x - 1) + 0.5 = (x + 0.5) - 1 exactly), so the result is unchanged.
constexpr int kLodTX = 16, kLodTY = 8, kLodTZ = 4;
__global__ __launch_bounds__(kLodTX * kLodTY * kLodTZ) void k_lod_transform(Dims3 din, Dims3 dout, SampleLaunch L,
                                                                            float* __restrict__ out) {
  constexpr int HX = kLodTX + 2, HY = kLodTY + 2, HZ = kLodTZ + 2, NH = HX * HY * HZ;
  __shared__ float hp[3][NH];
  const SampleView& v = L.view[blockIdx.y];
  const int ntx = (dout.x + kLodTX - 1) / kLodTX, nty = (dout.y + kLodTY - 1) / kLodTY;
  const int t = (int)blockIdx.x;
  const int x0 = (t % ntx) * kLodTX, y0 = ((t / ntx) % nty) * kLodTY, z0 = (t / (ntx * nty)) * kLodTZ;
  const float3 oinv = make_float3(1.f / dout.x, 1.f / dout.y, 1.f / dout.z);
  const float3 idims = make_float3((float)din.x, (float)din.y, (float)din.z);
  for (int k = threadIdx.x; k < NH; k += blockDim.x) {
    const int hx = k % HX, hy = (k / HX) % HY, hz = k / (HX * HY);
    const float3 p = map_pos(v, L.cm, (float)(x0 + hx - 1) + 0.5f, (float)(y0 + hy - 1) + 0.5f,
                             (float)(z0 + hz - 1) + 0.5f, oinv, idims);
    hp[0][k] = p.x;
    hp[1][k] = p.y;
    hp[2][k] = p.z;
  }
  __syncthreads();
  const int lx = threadIdx.x % kLodTX, ly = (threadIdx.x / kLodTX) % kLodTY, lz = threadIdx.x / (kLodTX * kLodTY);
  const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
  if (x >= dout.x || y >= dout.y || z >= dout.z) return;
  auto at = [&](int dx, int dy, int dz) {
    const int k = ((lz + 1 + dz) * HY + (ly + 1 + dy)) * HX + (lx + 1 + dx);
    return make_float3(hp[0][k], hp[1][k], hp[2][k]);
  };
  const float3 p = at(0, 0, 0);
  const float3 xm = at(-1, 0, 0), xp = at(1, 0, 0), ym = at(0, -1, 0), yp = at(0, 1, 0), zm = at(0, 0, -1),
               zp = at(0, 0, 1);
  float3 dx = make_float3((xm.x - xp.x) * 0.5f, (xm.y - xp.y) * 0.5f, (xm.z - xp.z) * 0.5f);
  float3 dy = make_float3((ym.x - yp.x) * 0.5f, (ym.y - yp.y) * 0.5f, (ym.z - yp.z) * 0.5f);
  float3 dz = make_float3((zm.x - zp.x) * 0.5f, (zm.y - zp.y) * 0.5f, (zm.z - zp.z) * 0.5f);
  float fx = len3f(dx.x, dy.x, dz.x), fy = len3f(dx.y, dy.y, dz.y), fz = len3f(dx.z, dy.z, dz.z);
  float lod = fmaxf(0.f, log2f(fmaxf(fx, fmaxf(fy, fz))));
  const size_t n = (size_t)dout.x * dout.y * dout.z;
  const size_t f = ((size_t)z * dout.y + y) * dout.x + x;
  *reinterpret_cast<float4*>(out + ((size_t)v.out_slot * n + f) * 4) = make_float4(p.x, p.y, p.z, lod);
}

// ------------------------------------------------ shadow transmittance
// Fused spot-light shadow map (phitest/render/lighting.py:106-141): the
// density resampled into the light frustum (TRANSFORM_LINDEPTH), each tap
// scaled by the depth step as the step-premultiplied grid would hold it
// (renderer.py:365-385), its running sum along depth (reduce_grid_blend
// keep_dims on one channel is a plain prefix sum), the one-cell shift with
// the boundary's pad, the transmittance exp(-D) or 1 - clip(D, 0, 1) and
// the cone mask, in one pass that writes only the transmittance and,
// optionally, the prefix sums the backward needs.
//
// A block is 64 shadow-map pixels (one per lane, x fastest) by kShadowWaves
// waves; a round covers kShadowWaves * kShadowSeg depth cells, wave w
// sampling cells [w seg, (w + 1) seg) of it into LDS. Every wave then runs
// the round's serial sum over LDS in depth order, the order of the unfused
// running sum, so the prefix sums match it bit for bit, and keeps the values
// of its own cells.
constexpr int kShadowWaves = 8, kShadowSeg = 8, kShadowRound = kShadowWaves * kShadowSeg;
constexpr int kShadowPitch = kShadowRound + 4;  // floats per pixel row; +4 spreads lanes over banks
constexpr int kShadowUnroll = 1;

// A grid value times the step, rounded on its own (never fused into the
// lerp that follows), as the step-premultiplied grid stores it.
__device__ __forceinline__ float scaled_load(const float* __restrict__ p, float sc) {
#pragma clang fp contract(off)
  return *p * sc;
}

// One step-scaled tap of the v1 sampler (read_linear / read_nearest with
// every corner multiplied by `sc` before the lerps).
template <int BM, int FM>
__device__ __forceinline__ float fetch_scaled(const float* __restrict__ buf, Dims3 d, float px, float py, float pz,
                                              float sc) {
  if constexpr (FM == F_NEAREST) {
    size_t f;
    bool ok = nearest_index<BM>(px, py, pz, d, f);
    return ok ? scaled_load(buf + f, sc) : 0.f;
  } else {
    Taps t = make_taps<BM>(px, py, pz, d);
    if (BM == B_BORDER && !t.inside) return 0.f;
    float c[2][2][2];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          c[k][j][i] = scaled_load(buf + flat3(t.ix[i], t.iy[j], t.iz[k], d), sc);
          if (BM == B_BORDER && !(t.vx[i] & t.vy[j] & t.vz[k])) c[k][j][i] = 0.f;
        }
    float v0 = lerpf(lerpf(c[0][0][0], c[0][0][1], t.wx), lerpf(c[0][1][0], c[0][1][1], t.wx), t.wy);
    float v1 = lerpf(lerpf(c[1][0][0], c[1][0][1], t.wx), lerpf(c[1][1][0], c[1][1][1], t.wx), t.wy);
    return lerpf(v0, v1, t.wz);
  }
}

template <int BL> __device__ __forceinline__ float shadow_transmit(float d) {
  if constexpr (BL == BL_BEER_LAMBERT) return expf(-d);
  else return 1.f - fminf(fmaxf(d, 0.f), 1.f);
}

// sample_v1 with step-scaled taps: level 0 is the unscaled grid, read with
// `step`; the mip levels come from the pyramid of the scaled grid.
template <int BM, int FM>
__device__ __forceinline__ float sample_scaled(const float* __restrict__ base, const float* __restrict__ atlas,
                                               const MipTable& mt, const SampleCfg& s, float3 p, float lod,
                                               float step) {
  int lv[2];
  float w[2];
  int n = level_taps(s, lod, lv, w);
  float sc = 1.f / (float)(1 << lv[0]);
  float a = fetch_scaled<BM, FM>(level_ptr(base, atlas, mt, 0, lv[0]), mt.dims[lv[0]], p.x * sc - s.offset,
                                 p.y * sc - s.offset, p.z * sc - s.offset, lv[0] == 0 ? step : 1.f);
  if (n == 2) {
    float sc1 = 1.f / (float)(1 << lv[1]);
    float c = fetch_scaled<BM, FM>(level_ptr(base, atlas, mt, 0, lv[1]), mt.dims[lv[1]], p.x * sc1 - s.offset,
                                   p.y * sc1 - s.offset, p.z * sc1 - s.offset, lv[1] == 0 ? step : 1.f);
    a = lerpf(a, c, w[1]);
  }
  return a;
}

template <int BM, int FM, int BL, bool LOD>
__global__ __launch_bounds__(kShadowWaves * 64) void k_shadow_fwd(const float* __restrict__ dens,
                                                                  const float* __restrict__ atlas, MipTable mt,
                                                                  SampleCfg cfg, SampleView v, Dims3 dout, float step,
                                                                  const float* __restrict__ mask,
                                                                  float* __restrict__ saved,
                                                                  float* __restrict__ trans) {
  __shared__ float s_lds[64 * kShadowPitch];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const size_t plane = (size_t)dout.x * dout.y;
  const size_t p = (size_t)blockIdx.x * 64 + lane;
  const bool live = p < plane;
  const int x = live ? (int)(p % (size_t)dout.x) : 0, y = live ? (int)(p / (size_t)dout.x) : 0;
  const float3 oinv = make_float3(1.f / dout.x, 1.f / dout.y, 1.f / dout.z);
  const float3 idims = make_float3((float)mt.dims[0].x, (float)mt.dims[0].y, (float)mt.dims[0].z);
  const float mk = (mask && live) ? mask[p] : 1.f;
  float* row = s_lds + lane * kShadowPitch;
  float carry = 0.f;
  for (int r0 = 0; r0 < dout.z; r0 += kShadowRound) {
    // one sample per iteration keeps the kernel at the register budget of
    // the plain resample (8 waves per SIMD); the waves hide each other's latency
#pragma unroll kShadowUnroll
    for (int i = 0; i < kShadowSeg; ++i) {
      int z = r0 + w * kShadowSeg + i;
      float val = 0.f;
      if (live && z < dout.z) {
        float l;
        float3 q = sample_pos(v, CM_LINDEPTH, LOD, x + 0.5f, y + 0.5f, z + 0.5f, oinv, idims, l);
        val = sample_scaled<BM, FM>(dens, atlas, mt, cfg, q, l, step);
      }
      row[w * kShadowSeg + i] = val;
    }
    __syncthreads();
    float acc = carry;
    float mine[kShadowSeg];
    for (int sg = 0; sg < kShadowWaves; ++sg) {
#pragma unroll
      for (int i = 0; i < kShadowSeg; i += 4) {
        float4 q = *reinterpret_cast<const float4*>(row + sg * kShadowSeg + i);
        const float e[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          acc += e[u];
          mine[i + u] = sg == w ? acc : mine[i + u];
        }
      }
    }
    carry = acc;
    if (live) {
#pragma unroll
      for (int i = 0; i < kShadowSeg; ++i) {
        int z = r0 + w * kShadowSeg + i;
        if (z >= dout.z) break;
        float dz = mine[i];
        if (saved) saved[(size_t)z * plane + p] = dz;
        if (z + 1 < dout.z) trans[(size_t)(z + 1) * plane + p] = shadow_transmit<BL>(dz) * mk;
        if (z == 0) trans[p] = shadow_transmit<BL>(BM == B_CLAMP ? dz : 0.f) * mk;
      }
    }
    __syncthreads();
  }
}

// Reverse of k_shadow_fwd without the resample: u_z = dL/dD_z from the
// transmittance gradient of the cell it shades (and, under CLAMP, of the pad
// cell 0), and the running-sum adjoint g_s[z] = sum_{k >= z} u_k accumulated
// from the back, the order of the blend gradient (reduce_blend.cu.cc:34-79),
// so the result matches the op sequence's bit for bit given the same u.
// Same block shape as the forward, rounds walked from the last: waves store
// the u of their cells in LDS and every wave runs the round's serial sum.
template <int BL> __device__ __forceinline__ float shadow_transmit_grad(float gm, float d) {
  if constexpr (BL == BL_BEER_LAMBERT) return -(gm * expf(-d));
  else return (0.f <= d && d <= 1.f) ? -gm : 0.f;
}

template <int BL, bool CLAMP_PAD>
__global__ __launch_bounds__(kShadowWaves * 64) void k_shadow_bwd(const float* __restrict__ saved,
                                                                  const float* __restrict__ gt,
                                                                  const float* __restrict__ mask, Dims3 dout,
                                                                  float* __restrict__ gs) {
  __shared__ float s_lds[64 * kShadowPitch];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const size_t plane = (size_t)dout.x * dout.y;
  const size_t p = (size_t)blockIdx.x * 64 + lane;
  const bool live = p < plane;
  const float mk = (mask && live) ? mask[p] : 1.f;
  float* row = s_lds + lane * kShadowPitch;
  float carry = 0.f;
  const int rounds = (dout.z + kShadowRound - 1) / kShadowRound;
  for (int r = rounds - 1; r >= 0; --r) {
    const int r0 = r * kShadowRound;
    float uv[kShadowSeg];
#pragma unroll
    for (int i = 0; i < kShadowSeg; ++i) {
      int z = r0 + w * kShadowSeg + i;
      float u = 0.f;
      if (live && z < dout.z) {
        float dz = saved[(size_t)z * plane + p];
        if (z + 1 < dout.z) u = shadow_transmit_grad<BL>(gt[(size_t)(z + 1) * plane + p] * mk, dz);
        if (CLAMP_PAD && z == 0) u = u + shadow_transmit_grad<BL>(gt[p] * mk, dz);
      }
      uv[i] = u;
    }
#pragma unroll
    for (int i = 0; i < kShadowSeg; i += 4)
      *reinterpret_cast<float4*>(row + w * kShadowSeg + i) = make_float4(uv[i], uv[i + 1], uv[i + 2], uv[i + 3]);
    __syncthreads();
    float acc = carry;
    float mine[kShadowSeg];
    for (int sg = kShadowWaves - 1; sg >= 0; --sg) {
#pragma unroll
      for (int i = kShadowSeg - 4; i >= 0; i -= 4) {
        float4 q = *reinterpret_cast<const float4*>(row + sg * kShadowSeg + i);
        const float e[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int u = 3; u >= 0; --u) {
          acc += e[u];
          mine[i + u] = sg == w ? acc : mine[i + u];
        }
      }
    }
    carry = acc;
    if (live) {
#pragma unroll
      for (int i = 0; i < kShadowSeg; ++i) {
        int z = r0 + w * kShadowSeg + i;
        if (z < dout.z) gs[(size_t)z * plane + p] = mine[i];
      }
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------- launchers
static inline dim3 grid_for(size_t n, unsigned views) { return dim3((unsigned)((n + 255) / 256), views); }

#define GFT_DISPATCH_C(C, ...)                     \
  switch (C) {                                     \
    case 1: { constexpr int CC = 1; __VA_ARGS__; } break; \
    case 2: { constexpr int CC = 2; __VA_ARGS__; } break; \
    default: { constexpr int CC = 4; __VA_ARGS__; } break; \
  }
#define GFT_DISPATCH_BM(bm, ...)                                 \
  switch (bm) {                                                  \
    case B_BORDER: { constexpr int BB = B_BORDER; __VA_ARGS__; } break; \
    case B_CLAMP: { constexpr int BB = B_CLAMP; __VA_ARGS__; } break;   \
    default: { constexpr int BB = B_WRAP; __VA_ARGS__; } break;         \
  }

hipError_t mip_down_launch(int C, hipStream_t st, const float* src, size_t src_stride, Dims3 s, float* dst,
                           size_t dst_stride, Dims3 o, int batches, float scale) {
  size_t n = (size_t)o.x * o.y * o.z;
  GFT_DISPATCH_C(C, hipLaunchKernelGGL(k_mip_down<CC>, grid_for(n, batches), dim3(256), 0, st, src, src_stride, s,
                                       dst, dst_stride, o, scale));
  return hipGetLastError();
}

hipError_t mip_collapse_launch(int C, hipStream_t st, const float* gp, size_t gp_stride, Dims3 o, float* gc,
                               size_t gc_stride, Dims3 s, int batches) {
  size_t n = (size_t)s.x * s.y * s.z;
  GFT_DISPATCH_C(C, hipLaunchKernelGGL(k_mip_collapse<CC>, grid_for(n, batches), dim3(256), 0, st, gp, gp_stride, o,
                                       gc, gc_stride, s));
  return hipGetLastError();
}

hipError_t sample_transform_launch(int C, int bm, bool sum, hipStream_t st, const float* in, const float* atlas,
                                   const MipTable& mt, const SampleCfg& s, Dims3 dout, const SampleLaunch& L,
                                   float* out) {
  size_t n = (size_t)dout.x * dout.y * dout.z;
  GFT_DISPATCH_C(C, GFT_DISPATCH_BM(bm, {
    if (sum)
      hipLaunchKernelGGL((k_sample_transform<CC, BB, true>), grid_for(n, 1), dim3(256), 0, st, in, atlas, mt, s, dout,
                         L, out);
    else
      hipLaunchKernelGGL((k_sample_transform<CC, BB, false>), grid_for(n, L.num_views), dim3(256), 0, st, in, atlas,
                         mt, s, dout, L, out);
  }));
  return hipGetLastError();
}

// gather-form gradient over the voxel blocks of one batch (grad points at it)
hipError_t sample_transform_gather_launch(int C, bool ref, int fm, hipStream_t st, const float* og, Dims3 din,
                                          Dims3 dout, const SampleLaunch& L, const SampleLaunch& Linv, int accumulate,
                                          float* grad) {
  const unsigned nb = (unsigned)(((din.x + kGBX - 1) / kGBX) * ((din.y + kGBY - 1) / kGBY) *
                                 ((din.z + kGBZ - 1) / kGBZ));
  KernelTimer t(st, "transform_grad_gather");
  GFT_DISPATCH_C(C, {
    if (ref)
      hipLaunchKernelGGL((k_sample_transform_gather_adj<CC, true, F_LINEAR>), dim3(nb), dim3(256), 0, st, og, din,
                         dout, L, Linv, accumulate, grad);
    else if (fm == F_NEAREST)
      hipLaunchKernelGGL((k_sample_transform_gather_adj<CC, false, F_NEAREST>), dim3(nb), dim3(256), 0, st, og, din,
                         dout, L, Linv, accumulate, grad);
    else
      hipLaunchKernelGGL((k_sample_transform_gather_adj<CC, false, F_LINEAR>), dim3(nb), dim3(256), 0, st, og, din,
                         dout, L, Linv, accumulate, grad);
  });
  return hipGetLastError();
}

hipError_t sample_transform_scatter_ref_launch(int C, int bm, hipStream_t st, const float* og, Dims3 din, Dims3 dout,
                                               const SampleLaunch& L, float* grad, unsigned* cnt) {
  size_t n = (size_t)dout.x * dout.y * dout.z;
  GFT_DISPATCH_C(C, GFT_DISPATCH_BM(bm, hipLaunchKernelGGL((k_sample_transform_scatter_ref<CC, BB>),
                                                           grid_for(n, L.num_views), dim3(256), 0, st, og, din, dout,
                                                           L, grad, cnt)));
  return hipGetLastError();
}

hipError_t sample_transform_adjoint_launch(int C, int bm, hipStream_t st, const float* in, const float* atlas,
                                           const MipTable& mt, const SampleCfg& s, Dims3 dout, const SampleLaunch& L,
                                           const float* og, float* grad, float* gatlas) {
  size_t n = (size_t)dout.x * dout.y * dout.z;
  GFT_DISPATCH_C(C, GFT_DISPATCH_BM(bm, hipLaunchKernelGGL((k_sample_transform_adjoint<CC, BB>),
                                                           grid_for(n, L.num_views), dim3(256), 0, st, in, atlas, mt,
                                                           s, dout, L, og, grad, gatlas)));
  return hipGetLastError();
}

hipError_t normalize_count_launch(int C, hipStream_t st, float* grad, const unsigned* cnt, size_t nvox) {
  GFT_DISPATCH_C(C, hipLaunchKernelGGL(k_normalize_count<CC>, grid_for(nvox, 1), dim3(256), 0, st, grad, cnt, nvox));
  return hipGetLastError();
}

hipError_t sample_lut_launch(int C, int bm, hipStream_t st, const float* in, const float* atlas, const MipTable& mt,
                             const SampleCfg& s, const LutLaunch& L, int pairs, const float* lut, float* out) {
  size_t n = (size_t)L.dout.x * L.dout.y * L.dout.z;
  GFT_DISPATCH_C(C, GFT_DISPATCH_BM(bm, hipLaunchKernelGGL((k_sample_lut<CC, BB>), grid_for(n, pairs), dim3(256), 0,
                                                           st, in, atlas, mt, s, L, lut, out)));
  return hipGetLastError();
}

hipError_t sample_lut_bwd_launch(int C, int bm, bool exact, hipStream_t st, const float* in_b, int fm, float offset,
                                 const LutLaunch& L, int views, const float* lut, const float* og, Dims3 din,
                                 float* grad_b, unsigned* cnt, float* lut_grad) {
  size_t n = (size_t)L.dout.x * L.dout.y * L.dout.z;
  KernelTimer t(st, "lut_grad");
  if (C == 1 && L.relative && (fm == F_LINEAR || fm == F_NEAREST || !exact)) {
    const unsigned tiles = (unsigned)(((L.dout.x + kLTX - 1) / kLTX) * ((L.dout.y + kLTY - 1) / kLTY) *
                                      ((L.dout.z + kLTZ - 1) / kLTZ));
    const dim3 grid(tiles, (unsigned)views);
    GFT_DISPATCH_BM(bm, {
      if (exact)
        hipLaunchKernelGGL((k_sample_lut_bwd_tiled<BB, true>), grid, dim3(256), 0, st, in_b, fm, offset, L, lut, og,
                           din, grad_b, cnt, lut_grad);
      else
        hipLaunchKernelGGL((k_sample_lut_bwd_tiled<BB, false>), grid, dim3(256), 0, st, in_b, fm, offset, L, lut, og,
                           din, grad_b, cnt, lut_grad);
    });
    return hipGetLastError();
  }
  GFT_DISPATCH_C(C, GFT_DISPATCH_BM(bm, {
    if (exact)
      hipLaunchKernelGGL((k_sample_lut_bwd<CC, BB, true>), grid_for(n, views), dim3(256), 0, st, in_b, fm, offset, L,
                         lut, og, din, grad_b, cnt, lut_grad);
    else
      hipLaunchKernelGGL((k_sample_lut_bwd<CC, BB, false>), grid_for(n, views), dim3(256), 0, st, in_b, fm, offset, L,
                         lut, og, din, grad_b, cnt, lut_grad);
  }));
  return hipGetLastError();
}

hipError_t lod_transform_launch(hipStream_t st, Dims3 din, Dims3 dout, const SampleLaunch& L, float* out) {
  const unsigned tiles = (unsigned)(((dout.x + kLodTX - 1) / kLodTX) * ((dout.y + kLodTY - 1) / kLodTY) *
                                   ((dout.z + kLodTZ - 1) / kLodTZ));
  hipLaunchKernelGGL(k_lod_transform, dim3(tiles, (unsigned)L.num_views), dim3(kLodTX * kLodTY * kLodTZ), 0, st, din,
                     dout, L, out);
  return hipGetLastError();
}

hipError_t shadow_fwd_launch(int bm, int fm, int bl, hipStream_t st, const float* dens, const float* atlas,
                             const MipTable& mt, const SampleCfg& cfg, const SampleView& v, Dims3 dout, float step,
                             const float* mask, float* saved, float* trans) {
  KernelTimer t(st, "shadow_fwd");
  size_t plane = (size_t)dout.x * dout.y;
  dim3 grid((unsigned)((plane + 63) / 64)), block(kShadowWaves * 64);
  const bool lod = cfg.mm != MIP_NONE;
#define GFT_SHADOW_FWD(B, F, L)                                                                                    \
  do {                                                                                                             \
    if (lod)                                                                                                       \
      hipLaunchKernelGGL((k_shadow_fwd<B, F, L, true>), grid, block, 0, st, dens, atlas, mt, cfg, v, dout, step,   \
                         mask, saved, trans);                                                                      \
    else                                                                                                           \
      hipLaunchKernelGGL((k_shadow_fwd<B, F, L, false>), grid, block, 0, st, dens, atlas, mt, cfg, v, dout, step,  \
                         mask, saved, trans);                                                                      \
  } while (0)
#define GFT_SHADOW_FWD_BL(B, F)              \
  do {                                       \
    if (bl == BL_ALPHA)                      \
      GFT_SHADOW_FWD(B, F, BL_ALPHA);        \
    else                                     \
      GFT_SHADOW_FWD(B, F, BL_BEER_LAMBERT); \
  } while (0)
  if (bm == B_CLAMP) {
    if (fm == F_NEAREST) GFT_SHADOW_FWD_BL(B_CLAMP, F_NEAREST);
    else GFT_SHADOW_FWD_BL(B_CLAMP, F_LINEAR);
  } else {
    if (fm == F_NEAREST) GFT_SHADOW_FWD_BL(B_BORDER, F_NEAREST);
    else GFT_SHADOW_FWD_BL(B_BORDER, F_LINEAR);
  }
#undef GFT_SHADOW_FWD_BL
#undef GFT_SHADOW_FWD
  return hipGetLastError();
}

hipError_t shadow_bwd_launch(int bm, int bl, hipStream_t st, const float* saved, const float* gt, const float* mask,
                             Dims3 dout, float* gs) {
  KernelTimer t(st, "shadow_bwd");
  size_t plane = (size_t)dout.x * dout.y;
  dim3 grid((unsigned)((plane + 63) / 64)), block(kShadowWaves * 64);
  if (bl == BL_ALPHA) {
    if (bm == B_CLAMP) hipLaunchKernelGGL((k_shadow_bwd<BL_ALPHA, true>), grid, block, 0, st, saved, gt, mask, dout, gs);
    else hipLaunchKernelGGL((k_shadow_bwd<BL_ALPHA, false>), grid, block, 0, st, saved, gt, mask, dout, gs);
  } else {
    if (bm == B_CLAMP)
      hipLaunchKernelGGL((k_shadow_bwd<BL_BEER_LAMBERT, true>), grid, block, 0, st, saved, gt, mask, dout, gs);
    else hipLaunchKernelGGL((k_shadow_bwd<BL_BEER_LAMBERT, false>), grid, block, 0, st, saved, gt, mask, dout, gs);
  }
  return hipGetLastError();
}

}  // namespace gft
