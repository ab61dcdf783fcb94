// Batched mipmapped density render: the density-reconstruction step of many
// independent reconstructions of uneven grid sizes and view counts (1-8
// views each) in a handful of launches over one device table of
// reconstruction records.
//
// Per reconstruction the step is the reference's non-fused render of a
// density under an ambient light (Renderer.render_density with fused=False,
// renderer.py:555-568,603-651): the step-scaled light/density grid resampled
// into every camera frustum with SampleGridTransform TRANSFORM_LINDEPTH and
// LINEAR mipmapping, the LoD of every sample taken from its frustum footprint
// (samplePos3D<TransformLinDepth, true>, sampling.hpp:114-123; levels 0..L,
// L <= 3), the frustum grid blended along depth with _blend_grid
// BEER_LAMBERT (renderer.py:488-499: D_i = sum_{j<=i} sigma_j, rgb = sum_i
// L_i exp(-D_i)), the mean squared rgb error against the targets, and the
// exact gradient w.r.t. the density through the blend, the resample and the
// mip pyramid.
//
// Launches (each one over every reconstruction, grid-stride over its items):
//   k_batch_level(m)   m = 0: copy the density into level 0 of the pyramid
//                      and clear the density gradient; m >= 1: build mip
//                      level m from level m-1 (kGenerateMip3D's pairwise
//                      means) and clear its gradient level;
//   k_batch_render     forward and backward of every ray in one pass: items
//                      are (reconstruction, view, 16x16-pixel tile); a ray,
//                      clipped to the grid box, marches its samples keeping
//                      each sample's c_k, writes its pixel and loss partial,
//                      forms its loss gradient and scatters the sample
//                      gradients alpha c_k + beta through the level taps;
//                      level-0 and level-1 taps go to FP64 LDS bricks per tile
//                      (or 8x8 sub-tile) and segment of at most K samples,
//                      flushed with float atomics; coarser levels (rare) take
//                      float atomics directly;
//   k_batch_finish     the mip adjoint (every level-0 voxel adds 1/8^m of its
//                      level-m ancestor's gradient) and, in the first item of
//                      each reconstruction, its loss: the per-tile partials
//                      summed in a fixed order.
// The brick of the backward (tile size x longest segment) is chosen per
// reconstruction: 16x16-pixel tiles with 8-sample segments (about 8^3 cells
// at two pixels per cell), 8x8 tiles with 4-sample segments (about 4^3), or
// 8x8 tiles with 8 (about 4x4x8). A segment whose bricks would not fit the
// LDS is halved until they do.
#include <algorithm>
#include <cstring>
#include <vector>

#include "gft_host.hpp"
#include "sampler_v1.hpp"

namespace gft {

constexpr int kBatchMaxViews = 8;
constexpr int kBatchMaxLevels = 3;
constexpr int kBatchCap = 6144;  // FP64 brick cells per workgroup (48 KiB)
constexpr int kRenderGrid = 768;  // render workgroups: as many as are resident (three per CU, by LDS)

struct BatchRec {
  MipTable mt;            // dims of levels 0..L; offset[m]: float offset of level m in pyr (offset[0] = 0)
  const float* dens;
  float* pyr;             // density pyramid, levels 0..L (level 0 a copy of dens)
  float* grad;            // level-0 gradient, overwritten
  float* gatlas;          // gradients of levels 1..L, level m at offset[m] - offset[1]
  const float* target;    // [V,H,W,3]
  float* images;          // [V,H,W,4]
  float* partials;        // per render item
  float* loss;            // [1]
  float light[3];
  float step;
  float loss_scale;
  int num_views, out_w, out_h, out_d;
  int tiles_x, tiles_y;   // 16x16-pixel tiles per view
  int tile;               // brick tile: 16 (one brick per workgroup) or 8 (one per wave)
  int seg;                // samples per backward segment
  int item0;              // first render item
  int lvl_item0[kBatchMaxLevels + 1];  // first item of the level passes
  SampleView view[kBatchMaxViews];
};

// index of the reconstruction owning global item g (records sorted by first item)
__device__ __forceinline__ int rec_of(const BatchRec* __restrict__ recs, int nrec, int g, int level) {
  int lo = 0, hi = nrec - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int first = level < 0 ? recs[mid].item0 : recs[mid].lvl_item0[level];
    if (first <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ float* level_grad(const BatchRec& R, int m) {
  return m == 0 ? R.grad : R.gatlas + (R.mt.offset[m] - R.mt.offset[1]);
}

// ------------------------------------------------------------ level passes
// m = 0: copy the density into the pyramid and zero the level-0 gradient;
// m >= 1: level m of the pyramid (pairwise means in z, then y, then x, as
// kGenerateMip3D) and zero its gradient.
__global__ __launch_bounds__(256) void k_batch_level(const BatchRec* __restrict__ recs, int nrec, int m,
                                                     int total_items) {
#pragma clang fp contract(off)
  for (int item = blockIdx.x; item < total_items; item += gridDim.x) {
    const BatchRec& R = recs[rec_of(recs, nrec, item, m)];
    if (m > R.mt.levels) continue;
    const Dims3 o = R.mt.dims[m];
    const size_t n = (size_t)o.x * o.y * o.z;
    const size_t f = (size_t)(item - R.lvl_item0[m]) * 256 + threadIdx.x;
    if (f >= n) continue;
    if (m == 0) {
      R.pyr[f] = R.dens[f];
      R.grad[f] = 0.f;
      continue;
    }
    const Dims3 s = R.mt.dims[m - 1];
    const float* in = m == 1 ? R.dens : R.pyr + R.mt.offset[m - 1];
    const int x = (int)(f % (size_t)o.x), y = (int)((f / (size_t)o.x) % (size_t)o.y), z = (int)(f / ((size_t)o.x * o.y));
    float ya[2];
#pragma unroll
    for (int xi = 0; xi < 2; ++xi) {
      float za[2];
#pragma unroll
      for (int yi = 0; yi < 2; ++yi) {
        const float a = in[flat3(2 * x + xi, 2 * y + yi, 2 * z, s)], b = in[flat3(2 * x + xi, 2 * y + yi, 2 * z + 1, s)];
        za[yi] = (a + b) * 0.5f;
      }
      ya[xi] = (za[0] + za[1]) * 0.5f;
    }
    R.pyr[R.mt.offset[m] + f] = (ya[0] + ya[1]) * 0.5f;
    level_grad(R, m)[f] = 0.f;
  }
}

// ---------------------------------------------------------------- render
// One pixel column of the TRANSFORM_LINDEPTH resample. LINDEPTH spaces the
// samples uniformly in view depth and a pixel's unprojection is affine in
// view depth, so the positions are affine in the depth index: p(i) = p0 +
// i dp, from the two end samples of the column. The LoD footprint's
// differences to the +x and +y pixels are affine in i as well (e0 + i e1)
// and the difference to sample i + 1 is dp.
struct Column {
  float3 p0, dp, ex0, ex1, ey0, ey1;
};

__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 scale3(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 fma3(float s, float3 a, float3 b) {
  return make_float3(fmaf(s, a.x, b.x), fmaf(s, a.y, b.y), fmaf(s, a.z, b.z));
}

__device__ __forceinline__ Column lindepth_column(const SampleView& v, float qx, float qy, int S, float3 oinv,
                                                  bool lod) {
  const float3 z3 = make_float3(0.f, 0.f, 0.f);
  const float qa = 0.5f, qb = (float)S - 0.5f, inv = S > 1 ? 1.f / (float)(S - 1) : 0.f;
  Column c;
  const float3 pa = map_pos(v, CM_LINDEPTH, qx, qy, qa, oinv, z3), pb = map_pos(v, CM_LINDEPTH, qx, qy, qb, oinv, z3);
  c.p0 = pa;
  c.dp = scale3(sub3(pb, pa), inv);
  c.ex0 = c.ex1 = c.ey0 = c.ey1 = z3;
  if (lod) {
    const float3 xa = sub3(pa, map_pos(v, CM_LINDEPTH, qx + 1.f, qy, qa, oinv, z3));
    const float3 xb = sub3(pb, map_pos(v, CM_LINDEPTH, qx + 1.f, qy, qb, oinv, z3));
    const float3 ya = sub3(pa, map_pos(v, CM_LINDEPTH, qx, qy + 1.f, qa, oinv, z3));
    const float3 yb = sub3(pb, map_pos(v, CM_LINDEPTH, qx, qy + 1.f, qb, oinv, z3));
    c.ex0 = xa;
    c.ex1 = scale3(sub3(xb, xa), inv);
    c.ey0 = ya;
    c.ey1 = scale3(sub3(yb, ya), inv);
  }
  return c;
}

// samplePos3D's LoD (sampling.hpp:114-123): log2 of the largest per-axis
// footprint, the norms of the Jacobian rows of the forward differences
__device__ __forceinline__ float column_lod(const Column& c, float fi) {
  const float3 ex = fma3(fi, c.ex1, c.ex0), ey = fma3(fi, c.ey1, c.ey0);
  const float fx = fmaf(ex.x, ex.x, fmaf(ey.x, ey.x, c.dp.x * c.dp.x));
  const float fy = fmaf(ex.y, ex.y, fmaf(ey.y, ey.y, c.dp.y * c.dp.y));
  const float fz = fmaf(ex.z, ex.z, fmaf(ey.z, ey.z, c.dp.z * c.dp.z));
  return fmaxf(0.f, 0.5f * log2f(fmaxf(fx, fmaxf(fy, fz))));
}

// weight of level m among a sample's mip taps (0 if the sample does not read it)
__device__ __forceinline__ float level_weight(int m, int nl, const int lv[2], const float lw[2]) {
  return lv[0] == m ? lw[0] : (((nl == 2) & (lv[1] == m)) ? lw[1] : 0.f);
}

// BORDER trilinear value at level base `buf` (make_taps' corners, zeroing and
// lerp order, as read_linear): flat indices of one level fit 32 bits
__device__ __forceinline__ float level_read(const float* __restrict__ buf, Dims3 d, float px, float py, float pz) {
  const Taps t = make_taps<B_BORDER>(px, py, pz, d);
  if (!t.inside) return 0.f;
  float c[2][2][2];
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const unsigned row = ((unsigned)t.iz[k] * (unsigned)d.y + (unsigned)t.iy[j]) * (unsigned)d.x;
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const float a = buf[row + (unsigned)t.ix[i]];
        c[k][j][i] = (t.vx[i] & t.vy[j] & t.vz[k]) ? a : 0.f;
      }
    }
  const float v0 = lerpf(lerpf(c[0][0][0], c[0][0][1], t.wx), lerpf(c[0][1][0], c[0][1][1], t.wx), t.wy);
  const float v1 = lerpf(lerpf(c[1][0][0], c[1][0][1], t.wx), lerpf(c[1][1][0], c[1][1][1], t.wx), t.wy);
  return lerpf(v0, v1, t.wz);
}

// the mipmapped sample: level taps of the LoD, each level read at a
// wave-uniform base (the level loop is uniform, the taps per lane)
__device__ __forceinline__ float batch_sample(const BatchRec& R, const SampleCfg& cfg, float3 p, float lod) {
  int lv[2];
  float lw[2];
  const int nl = level_taps(cfg, lod, lv, lw);
  float vf = 0.f, vc = 0.f;
#pragma unroll
  for (int m = 0; m <= kBatchMaxLevels; ++m) {
    if (m > R.mt.levels) break;
    const bool f = lv[0] == m, c = (nl == 2) & (lv[1] == m);
    if (f | c) {
      const float sc = 1.f / (float)(1 << m);
      const float a = level_read(R.pyr + R.mt.offset[m], R.mt.dims[m], p.x * sc - 0.5f, p.y * sc - 0.5f,
                                 p.z * sc - 0.5f);
      if (f) vf = a;
      else vc = a;
    }
  }
  return nl == 2 ? lerpf(vf, vc, lw[1]) : vf;
}

// Bounding bricks of levels 0 and 1 (corner indices) of the samples ia..ib
// of a tile's pixel columns: positions are affine along a column and, at one
// depth, across pixels, so the four corner columns at the two segment ends
// bound every sample of the tile. A tap outside its brick takes a float
// atomic, so the bound only decides where a tap is summed, never whether.
struct BatchBrick {
  int lo0[3], n0[3], lo1[3], n1[3];
  int size0, size;
};

__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

// corner columns of a tile in LDS: [c][0..2] = p0, [c][3..5] = dp; a 16-pixel
// brick uses corners 0..3 of a 2x2 table, the 8-pixel brick s of a 4x4 table
// (pixel offsets 0, 7, 8, 15) the corners of its quadrant
__device__ __forceinline__ BatchBrick batch_brick(const float (*corner)[6], int tile, int sub, int ia, int ib,
                                                  const MipTable& mt) {
  float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int cx = c & 1, cy = (c >> 1) & 1;
    const int k = tile == 16 ? cy * 2 + cx : ((sub >> 1) * 2 + cy) * 4 + (sub & 1) * 2 + cx;
    const float fi = (float)(c < 4 ? ia : ib);
    const float q[3] = {fmaf(fi, corner[k][3], corner[k][0]), fmaf(fi, corner[k][4], corner[k][1]),
                        fmaf(fi, corner[k][5], corner[k][2])};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      mn[a] = fminf(mn[a], q[a]);
      mx[a] = fmaxf(mx[a], q[a]);
    }
  }
  const int d0[3] = {mt.dims[0].x, mt.dims[0].y, mt.dims[0].z};
  const int d1[3] = {mt.dims[1].x, mt.dims[1].y, mt.dims[1].z};
  BatchBrick b;
  int s0 = 1, s1 = 1;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int lo = clampi((int)floorf(mn[a] - 0.5f - 1e-3f), 0, d0[a] - 1);
    int hi = clampi((int)floorf(mx[a] - 0.5f + 1e-3f) + 1, 0, d0[a] - 1);
    b.lo0[a] = uniform_i(lo);
    b.n0[a] = uniform_i(hi - lo + 1);
    s0 *= b.n0[a];
    if (mt.levels > 0) {
      lo = clampi((int)floorf(mn[a] * 0.5f - 0.5f - 1e-3f), 0, d1[a] - 1);
      hi = clampi((int)floorf(mx[a] * 0.5f - 0.5f + 1e-3f) + 1, 0, d1[a] - 1);
    } else {
      lo = 0;
      hi = -1;
    }
    b.lo1[a] = uniform_i(lo);
    b.n1[a] = uniform_i(hi - lo + 1);
    s1 *= b.n1[a];
  }
  b.size0 = s0;
  b.size = s0 + s1;
  return b;
}

// gradient of one level tap (BORDER corners, weight products in the order
// ((g w_x) w_y) w_z): into the level's LDS brick when the whole corner cube
// lies in it, float atomics otherwise
__device__ __forceinline__ void level_scatter(double* __restrict__ lds, bool staged, const int blo[3],
                                              const int bn[3], float* __restrict__ gbuf, Dims3 d, float px, float py,
                                              float pz, float gw) {
  const Taps t = make_taps<B_BORDER>(px, py, pz, d);
  if (!t.inside) return;
  const float wx[2] = {1.f - t.wx, t.wx}, wy[2] = {1.f - t.wy, t.wy}, wz[2] = {1.f - t.wz, t.wz};
  float w[2][2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const float gx = gw * wx[i];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const float gxy = gx * wy[j];
#pragma unroll
      for (int k = 0; k < 2; ++k) w[k][j][i] = (t.vx[i] & t.vy[j] & t.vz[k]) ? gxy * wz[k] : 0.f;
    }
  }
  const int lx = t.ix[0] - blo[0], ly = t.iy[0] - blo[1], lz = t.iz[0] - blo[2];
  const int hx = t.ix[1] - blo[0], hy = t.iy[1] - blo[1], hz = t.iz[1] - blo[2];
  if (staged & (lx >= 0) & (ly >= 0) & (lz >= 0) & (hx < bn[0]) & (hy < bn[1]) & (hz < bn[2])) {
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int row = ((k ? hz : lz) * bn[1] + (j ? hy : ly)) * bn[0];
#pragma unroll
        for (int i = 0; i < 2; ++i) atomicAdd(&lds[row + (i ? hx : lx)], (double)w[k][j][i]);
      }
  } else {
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const unsigned row = ((unsigned)t.iz[k] * (unsigned)d.y + (unsigned)t.iy[j]) * (unsigned)d.x;
#pragma unroll
        for (int i = 0; i < 2; ++i)
          if (w[k][j][i] != 0.f) atomic_addf(gbuf + row + (unsigned)t.ix[i], w[k][j][i]);
      }
  }
}

// The forward keeps c_k = exp(-D_k) + E_{k-1} of every sample of its column
// (D_k the optical depth through sample k, E_{k-1} the emission before it):
// with E the pixel's total, dE/dsigma_k = exp(-D_k) - sum_{i>=k} exp(-D_i)
// sigma_i = c_k - E, so the backward needs no second march. The values live
// in a per-workgroup scratch of smax samples x 256 columns (sample-major, so
// a wave's 64 columns are one coalesced run), reused by every item of the
// workgroup and read back soon after they are written.
__global__ __launch_bounds__(256) void k_batch_render(const BatchRec* __restrict__ recs, int nrec, int total_items,
                                                      float* __restrict__ csave, int smax) {
  __shared__ double lds[kBatchCap];
  __shared__ float red[256];
  __shared__ int span[2];
  __shared__ float corner[16][6];
  const int wave = uniform_i((int)threadIdx.x >> 6), lane = threadIdx.x & 63;
  float* __restrict__ cs = csave + (size_t)blockIdx.x * smax * 256 + threadIdx.x;
  for (int i = threadIdx.x; i < kBatchCap; i += 256) lds[i] = 0.0;
  for (int item = blockIdx.x; item < total_items; item += gridDim.x) {
    const BatchRec& R = recs[rec_of(recs, nrec, item, -1)];
    const int local = item - R.item0;
    const int per_view = R.tiles_x * R.tiles_y;
    const int vi = local / per_view, t = local % per_view;
    const int tx0 = (t % R.tiles_x) * 16, ty0 = (t / R.tiles_x) * 16;
    const SampleView& V = R.view[vi];
    // this thread's pixel: 16x16 bricks take the tile in rows, 8x8 bricks give each wave a quadrant
    int x, y, cap, cap0;
    if (R.tile == 16) {
      x = tx0 + (threadIdx.x & 15);
      y = ty0 + (threadIdx.x >> 4);
      cap = kBatchCap;
      cap0 = 0;
    } else {
      x = tx0 + (wave & 1) * 8 + (lane & 7);
      y = ty0 + (wave >> 1) * 8 + (lane >> 3);
      cap = kBatchCap / 4;
      cap0 = wave * cap;
    }
    const bool valid = x < R.out_w && y < R.out_h;
    const int S = R.out_d;
    const float3 oinv = make_float3(1.f / (float)R.out_w, 1.f / (float)R.out_h, 1.f / (float)S);
    const bool mips = R.mt.levels > 0;
    const SampleCfg cfg{F_LINEAR, mips ? MIP_LINEAR : MIP_NONE, R.mt.levels, 0.f, 0.5f};
    const Column col = lindepth_column(V, x + 0.5f, y + 0.5f, S, oinv, mips);
    // samples outside the box [0, dims) read zero at every mip level (a level-m
    // sample is inside only where 0 <= p < dims_m 2^m <= dims), so they add
    // nothing to the blend or the gradient: the column is clipped to the box
    int lo = 1, hi = 0;
    if (valid) clip_ray(col.p0, col.dp, R.mt.dims[0], S, 0.f, lo, hi);
    // ---- forward: BEER_LAMBERT blend of the sampled frustum column
    float D = 0.f, E = 0.f;
    for (int i = lo; i <= hi; ++i) {
      const float fi = (float)i;
      const float lod = mips ? column_lod(col, fi) : 0.f;
      const float sigma = R.step * batch_sample(R, cfg, fma3(fi, col.dp, col.p0), lod);
      D += sigma;
      const float u = __expf(-D);
      cs[(size_t)i * 256] = u + E;
      E = fmaf(u, sigma, E);
    }
    float G = 0.f, sq = 0.f;
    if (valid) {
      const size_t pix = ((size_t)vi * R.out_h + y) * R.out_w + x;
      reinterpret_cast<float4*>(R.images)[pix] =
          make_float4(R.light[0] * E, R.light[1] * E, R.light[2] * E, D);
      const float* tg = R.target + pix * 3;
      const float e0 = R.light[0] * E - tg[0], e1 = R.light[1] * E - tg[1], e2 = R.light[2] * E - tg[2];
      sq = e0 * e0 + e1 * e1 + e2 * e2;
      G = R.loss_scale * 2.f * (R.light[0] * e0 + R.light[1] * e1 + R.light[2] * e2);
    }
    const bool walk = valid && G != 0.f && lo <= hi;
    const float alpha = R.step * G, beta = -(R.step * G) * E;
    // per-item loss partial: a fixed-order tree over the tile's 256 pixel
    // slots, the same sum whatever the brick's thread-to-pixel mapping; the
    // union of the walking columns' sample ranges; the bricks' corner columns
    red[(y - ty0) * 16 + (x - tx0)] = sq;
    if (threadIdx.x == 0) {
      span[0] = S;
      span[1] = -1;
    }
    if (threadIdx.x < (R.tile == 16 ? 4 : 16)) {
      const int ncx = R.tile == 16 ? 2 : 4;
      const int ox = (int)threadIdx.x % ncx, oy = (int)threadIdx.x / ncx;
      const int offs[4] = {0, 7, 8, 15};
      const int dx = R.tile == 16 ? 15 * ox : offs[ox], dy = R.tile == 16 ? 15 * oy : offs[oy];
      const Column cc = lindepth_column(V, min(tx0 + dx, R.out_w - 1) + 0.5f, min(ty0 + dy, R.out_h - 1) + 0.5f, S,
                                        oinv, false);
      float* cr = corner[threadIdx.x];
      cr[0] = cc.p0.x;
      cr[1] = cc.p0.y;
      cr[2] = cc.p0.z;
      cr[3] = cc.dp.x;
      cr[4] = cc.dp.y;
      cr[5] = cc.dp.z;
    }
    __syncthreads();
    if (walk) {
      atomicMin(&span[0], lo);
      atomicMax(&span[1], hi);
    }
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) R.partials[local] = red[0];
    const int s_lo = span[0], s_hi = span[1];
    __syncthreads();  // span and red are rewritten by the next item
    // ---- backward: segments of at most R.seg samples, each as long as its
    // bricks fit the LDS (halved until they do, down to one sample; in the
    // 8-pixel mode until every wave's brick fits, so all waves take the same
    // segments), the sample gradients c_k -> alpha c_k + beta scattered
    // through the level taps
    const int sub = R.tile == 16 ? 0 : wave;
    for (int ia = s_lo; ia <= s_hi;) {
      int k = R.seg, ib;
      BatchBrick br;
      for (;;) {
        ib = min(ia + k - 1, s_hi);
        br = batch_brick(corner, R.tile, sub, ia, ib, R.mt);
        int need = br.size;
        if (R.tile == 8)
          for (int s2 = 0; s2 < 4; ++s2)
            if (s2 != sub) need = max(need, batch_brick(corner, 8, s2, ia, ib, R.mt).size);
        if (need <= cap || k == 1) break;
        k >>= 1;
      }
      const bool use_lds = br.size <= cap;
      if (walk) {
        for (int i = max(ia, lo); i <= min(ib, hi); ++i) {
          const float fi = (float)i;
          const float3 p = fma3(fi, col.dp, col.p0);
          const float lod = mips ? column_lod(col, fi) : 0.f;
          const float g = fmaf(alpha, cs[(size_t)i * 256], beta);
          int lv[2];
          float lw[2];
          const int nl = level_taps(cfg, lod, lv, lw);
#pragma unroll
          for (int m = 0; m <= kBatchMaxLevels; ++m) {
            if (m > R.mt.levels) break;
            const float wm = level_weight(m, nl, lv, lw);
            if (wm != 0.f) {
              const float sc = 1.f / (float)(1 << m);
              level_scatter(lds + cap0 + (m == 1 ? br.size0 : 0), use_lds & (m <= 1), m == 0 ? br.lo0 : br.lo1,
                            m == 0 ? br.n0 : br.n1, level_grad(R, m), R.mt.dims[m], p.x * sc - 0.5f,
                            p.y * sc - 0.5f, p.z * sc - 0.5f, g * wm);
            }
          }
        }
      }
      // (8-pixel bricks: each wave owns its brick, so use_lds may differ between waves and the
      // barriers stay outside the branch)
      __syncthreads();
      if (use_lds) {
        // flush the bricks (all 256 threads for a 16-pixel tile, the wave's 64 lanes for an 8-pixel one)
        const int nthr = R.tile == 16 ? 256 : 64, tid = R.tile == 16 ? (int)threadIdx.x : lane;
        for (int idx = tid; idx < br.size; idx += nthr) {
          const double val = lds[cap0 + idx];
          if (val != 0.0) {
            const bool l0 = idx < br.size0;
            const int j = l0 ? idx : idx - br.size0;
            const int nx = l0 ? br.n0[0] : br.n1[0], ny = l0 ? br.n0[1] : br.n1[1];
            const int lx = j % nx, ly = (j / nx) % ny, lz = j / (nx * ny);
            if (l0)
              atomic_addf(R.grad + flat3(br.lo0[0] + lx, br.lo0[1] + ly, br.lo0[2] + lz, R.mt.dims[0]), (float)val);
            else
              atomic_addf(level_grad(R, 1) + flat3(br.lo1[0] + lx, br.lo1[1] + ly, br.lo1[2] + lz, R.mt.dims[1]),
                          (float)val);
            lds[cap0 + idx] = 0.0;
          }
        }
      }
      __syncthreads();
      ia = ib + 1;
    }
  }
}

// ------------------------------------------------------------- finishing
// Mip adjoint into the level-0 gradient (g0 += sum_m 8^-m g_m(ancestor)) and,
// in each reconstruction's first item, its loss from the render partials.
__global__ __launch_bounds__(256) void k_batch_finish(const BatchRec* __restrict__ recs, int nrec, int total_items) {
  __shared__ float red[256];
  for (int item = blockIdx.x; item < total_items; item += gridDim.x) {
    const BatchRec& R = recs[rec_of(recs, nrec, item, 0)];
    const int local = item - R.lvl_item0[0];
    if (local == 0) {
      const int n = R.num_views * R.tiles_x * R.tiles_y;
      float t = 0.f;
      for (int i = threadIdx.x; i < n; i += 256) t += R.partials[i];
      red[threadIdx.x] = t;
      __syncthreads();
      for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
      }
      if (threadIdx.x == 0) R.loss[0] = red[0] * R.loss_scale;
      __syncthreads();
    }
    if (R.mt.levels == 0) continue;
    const Dims3 d = R.mt.dims[0];
    const size_t f = (size_t)local * 256 + threadIdx.x;
    if (f >= (size_t)d.x * d.y * d.z) continue;
    const int x = (int)(f % (size_t)d.x), y = (int)((f / (size_t)d.x) % (size_t)d.y), z = (int)(f / ((size_t)d.x * d.y));
    float g = 0.f, s = 1.f;
    for (int m = 1; m <= R.mt.levels; ++m) {
      s *= 0.125f;
      const Dims3 o = R.mt.dims[m];
      const int px = x >> m, py = y >> m, pz = z >> m;
      if (px < o.x && py < o.y && pz < o.z) g = fmaf(s, level_grad(R, m)[flat3(px, py, pz, o)], g);
    }
    R.grad[f] += g;
  }
}

// ------------------------------------------------------------------ host
static size_t align256b(size_t b) { return (b + 255) & ~size_t(255); }

static MipTable batch_mip_table(Dims3 d, int levels) {
  MipTable mt{};
  levels = std::min(levels, kBatchMaxLevels);
  for (int m = 1; m <= levels; ++m)
    if ((d.x >> m) <= 0 || (d.y >> m) <= 0 || (d.z >> m) <= 0) {
      levels = m - 1;
      break;
    }
  mt.levels = levels;
  mt.C = 1;
  mt.dims[0] = d;
  mt.base_stride = (size_t)d.x * d.y * d.z;
  // the pyramid: level 0 at offset 0, each level on a 256-byte boundary
  size_t off = (mt.base_stride + 63) & ~size_t(63);
  for (int m = 1; m <= levels; ++m) {
    mt.dims[m] = Dims3{d.x >> m, d.y >> m, d.z >> m};
    mt.offset[m] = off;
    off += ((size_t)mt.dims[m].x * mt.dims[m].y * mt.dims[m].z + 63) & ~size_t(63);
  }
  mt.atlas_stride = off;
  return mt;
}

static int brick_tile(int brick) { return brick == 0 ? 16 : 8; }
static int brick_seg(int brick) { return brick == 1 ? 4 : 8; }

// workspace layout: the record table, then per reconstruction its density
// pyramid (levels 0..L), the gradients of levels 1..L and the render
// partials, then the render's c_k scratch
struct BatchLayout {
  size_t table, total;
  std::vector<size_t> atlas, gatlas, partials;
  size_t csave;  // the render's per-workgroup c_k scratch (kRenderGrid x 256 x smax floats)
  int smax;      // the most depth samples of any reconstruction
};

static int batch_layout(int32_t n, const gft_batch_recon* rs, int32_t num_mipmaps, BatchLayout& L) {
  if (n < 1) return fail(GFT_ERR_INVALID_ARGUMENT, "the batch needs at least one reconstruction");
  if (!rs) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  if (num_mipmaps < 0 || num_mipmaps > kBatchMaxLevels)
    return fail(GFT_ERR_INVALID_ARGUMENT, "num_mipmaps must be in [0, %d]", kBatchMaxLevels);
  size_t off = align256b(sizeof(BatchRec) * (size_t)n);
  L.table = off;
  L.atlas.resize(n);
  L.gatlas.resize(n);
  L.partials.resize(n);
  for (int r = 0; r < n; ++r) {
    const gft_batch_recon& c = rs[r];
    for (int a = 0; a < 3; ++a)
      if (c.grid_dhw[a] < 1 || c.output_shape[a] < 1)
        return fail(GFT_ERR_INVALID_ARGUMENT, "reconstruction %d: shapes must be positive", r);
    if (c.num_views < 1 || c.num_views > kBatchMaxViews)
      return fail(GFT_ERR_INVALID_ARGUMENT, "reconstruction %d: 1 to %d views, got %d", r, kBatchMaxViews,
                  c.num_views);
    if (c.brick < 0 || c.brick > 2) return fail(GFT_ERR_INVALID_ARGUMENT, "reconstruction %d: brick must be 0, 1 or 2", r);
    const Dims3 d{c.grid_dhw[2], c.grid_dhw[1], c.grid_dhw[0]};
    if ((int64_t)d.x * d.y * d.z >= ((int64_t)1 << 31)) return fail(GFT_ERR_INVALID_ARGUMENT, "grid too large");
    const MipTable mt = batch_mip_table(d, num_mipmaps);
    L.atlas[r] = off;
    off += align256b(mt.atlas_stride * sizeof(float));
    L.gatlas[r] = off;
    off += align256b((mt.levels > 0 ? mt.atlas_stride - mt.offset[1] : 0) * sizeof(float));
    const size_t items = (size_t)c.num_views * ((c.output_shape[2] + 15) / 16) * ((c.output_shape[1] + 15) / 16);
    L.partials[r] = off;
    off += align256b(items * sizeof(float));
  }
  L.smax = 1;
  for (int r = 0; r < n; ++r) L.smax = std::max(L.smax, (int)rs[r].output_shape[0]);
  L.csave = off;
  off += align256b((size_t)kRenderGrid * 256 * (size_t)L.smax * sizeof(float));
  L.total = off;
  return GFT_OK;
}

// item counts of the launches: render items, then per level
static void batch_items(int32_t n, const gft_batch_recon* rs, int32_t num_mipmaps, int& render, int level[4]) {
  render = 0;
  for (int m = 0; m <= kBatchMaxLevels; ++m) level[m] = 0;
  for (int r = 0; r < n; ++r) {
    const gft_batch_recon& c = rs[r];
    render += c.num_views * ((c.output_shape[2] + 15) / 16) * ((c.output_shape[1] + 15) / 16);
    const MipTable mt = batch_mip_table(Dims3{c.grid_dhw[2], c.grid_dhw[1], c.grid_dhw[0]}, num_mipmaps);
    for (int m = 0; m <= mt.levels; ++m)
      level[m] += (int)(((size_t)mt.dims[m].x * mt.dims[m].y * mt.dims[m].z + 255) / 256);
  }
}

static unsigned persistent_grid(int items) { return (unsigned)std::max(1, std::min(items, 256 * 8)); }

}  // namespace gft

using namespace gft;

extern "C" {

size_t gft_batch_render_workspace(int32_t n, const gft_batch_recon* recons, int32_t num_mipmaps) {
  BatchLayout L;
  if (batch_layout(n, recons, num_mipmaps, L)) return 0;
  return L.total;
}

int gft_batch_render_prepare(void* stream, int32_t n, const gft_batch_recon* recons, int32_t num_mipmaps,
                             void* workspace, size_t workspace_bytes) {
  BatchLayout L;
  int rc;
  if ((rc = batch_layout(n, recons, num_mipmaps, L))) return rc;
  if (!workspace || workspace_bytes < L.total)
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                L.total);
  std::vector<BatchRec> recs(n);
  char* ws = (char*)workspace;
  int item = 0, lvl_item[kBatchMaxLevels + 1] = {0, 0, 0, 0};
  for (int r = 0; r < n; ++r) {
    const gft_batch_recon& c = recons[r];
    if (!c.density || !c.targets || !c.images || !c.density_grad || !c.loss || !c.M || !c.V || !c.P || !c.frustum)
      return fail(GFT_ERR_INVALID_ARGUMENT, "reconstruction %d: null pointer argument", r);
    BatchRec& R = recs[r];
    std::memset(&R, 0, sizeof(R));
    const Dims3 d{c.grid_dhw[2], c.grid_dhw[1], c.grid_dhw[0]};
    R.mt = batch_mip_table(d, num_mipmaps);
    R.dens = c.density;
    R.pyr = (float*)(ws + L.atlas[r]);
    R.grad = c.density_grad;
    R.gatlas = (float*)(ws + L.gatlas[r]);
    R.target = c.targets;
    R.images = c.images;
    R.partials = (float*)(ws + L.partials[r]);
    R.loss = c.loss;
    for (int k = 0; k < 3; ++k) R.light[k] = c.light[k];
    R.step = c.depth_step;
    R.loss_scale = c.loss_scale;
    R.num_views = c.num_views;
    R.out_d = c.output_shape[0];
    R.out_h = c.output_shape[1];
    R.out_w = c.output_shape[2];
    R.tiles_x = (R.out_w + 15) / 16;
    R.tiles_y = (R.out_h + 15) / 16;
    R.tile = brick_tile(c.brick);
    R.seg = brick_seg(c.brick);
    R.item0 = item;
    item += R.num_views * R.tiles_x * R.tiles_y;
    for (int m = 0; m <= kBatchMaxLevels; ++m) {
      R.lvl_item0[m] = lvl_item[m];
      if (m <= R.mt.levels)
        lvl_item[m] += (int)(((size_t)R.mt.dims[m].x * R.mt.dims[m].y * R.mt.dims[m].z + 255) / 256);
    }
    const Mat4d m = mat_from_colmajor(c.M);
    for (int v = 0; v < c.num_views; ++v) {
      const Mat4d mv = mat_mul(mat_from_colmajor(c.V + 16 * v), m), p = mat_from_colmajor(c.P + 16 * v);
      Mat4d mvinv, pinv;
      if (!mat_inverse(mv, mvinv)) return fail(GFT_ERR_INVALID_ARGUMENT, "model-view matrix is singular");
      if (!mat_inverse(p, pinv)) return fail(GFT_ERR_INVALID_ARGUMENT, "projection matrix is singular");
      SampleView& sv = R.view[v];
      mat_to_float(pinv, sv.A);
      mat_to_float(mvinv, sv.B);
      sv.p22 = c.P[16 * v + 4 * 2 + 2];
      sv.p32 = c.P[16 * v + 4 * 3 + 2];
      sv.nearp = c.frustum[6 * v + 0];
      sv.farp = c.frustum[6 * v + 1];
      sv.out_slot = v;
    }
  }
  // a set-up call: the table is copied before it returns
  if ((rc = check_hip(hipMemcpyAsync(ws, recs.data(), sizeof(BatchRec) * (size_t)n, hipMemcpyHostToDevice,
                                     (hipStream_t)stream),
                      "copy the reconstruction table")))
    return rc;
  return check_hip(hipStreamSynchronize((hipStream_t)stream), "copy the reconstruction table");
}

int gft_batch_render_run(void* stream, int32_t n, const gft_batch_recon* recons, int32_t num_mipmaps,
                         void* workspace, size_t workspace_bytes) {
  BatchLayout L;
  int rc;
  if ((rc = batch_layout(n, recons, num_mipmaps, L))) return rc;
  if (!workspace || workspace_bytes < L.total)
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                L.total);
  hipStream_t s = (hipStream_t)stream;
  const BatchRec* recs = (const BatchRec*)workspace;
  int render = 0, level[kBatchMaxLevels + 1];
  batch_items(n, recons, num_mipmaps, render, level);
  {
    KernelTimer t(s, "batch_levels");
    for (int m = 0; m <= kBatchMaxLevels; ++m)
      if (level[m])
        hipLaunchKernelGGL(k_batch_level, dim3(persistent_grid(level[m])), dim3(256), 0, s, recs, n, m, level[m]);
  }
  if ((rc = check_hip(hipGetLastError(), "batched mip levels"))) return rc;
  {
    KernelTimer t(s, "batch_render");
    hipLaunchKernelGGL(k_batch_render, dim3((unsigned)std::max(1, std::min(render, kRenderGrid))), dim3(256), 0, s,
                       recs, n, render, (float*)((char*)workspace + L.csave), L.smax);
  }
  if ((rc = check_hip(hipGetLastError(), "batched render"))) return rc;
  {
    KernelTimer t(s, "batch_finish");
    hipLaunchKernelGGL(k_batch_finish, dim3(persistent_grid(level[0])), dim3(256), 0, s, recs, n, level[0]);
  }
  return check_hip(hipGetLastError(), "batched finish");
}

}  // extern "C"
