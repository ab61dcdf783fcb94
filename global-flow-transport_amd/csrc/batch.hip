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
// Launches (each one over the items of every reconstruction; the forward is
// persistent over a device work queue, the others run a workgroup per item):
//   k_batch_prep(m)    level m+1 of the zero-bordered pyramid
//                      (kGenerateMip3D's pairwise means of level m, level 0
//                      being the density) and the corner-quad copy of level
//                      m, in one launch per m;
//   k_batch_fwd        items (reconstruction, view, 16x16-pixel tile),
//                      taken from the queue by resident workgroups: a
//                      ray, clipped to the grid box, marches its samples
//                      keeping each sample's c_k, writes its pixel, its loss
//                      term, its gradient record and the view's largest LoD;
//   k_batch_gather(m)  one wave per block of level-m voxels: for every view
//                      whose samples read level m, the candidate pixels of the
//                      block's support box, each column cut by the box, every
//                      sample's gradient alpha c_k + beta times its level-m
//                      weight and its trilinear weights summed in FP64 LDS,
//                      then one store per voxel (no atomics in global memory:
//                      the gradient is the same bit for bit on every run).
//                      Levels 3 and 2 first; then levels 0 and 1 in one walk,
//                      whose store is the density gradient itself: every
//                      voxel's level-0 gradient plus 1/8^m of its level-m
//                      ancestor's (the mip adjoint; level 1 from the same
//                      wave's region); the first block of each reconstruction
//                      also sums its loss from the per-tile partials in a
//                      fixed order.
// The gather's block shape (16x16x2, 8x8x8 or 16x8x4 voxels) is chosen per
// reconstruction.
#include <algorithm>
#include <array>
#include <cstring>
#include <mutex>
#include <vector>

#include "gft_host.hpp"
#include "quad.hpp"
#include "sampler_v1.hpp"

namespace gft {

constexpr int kBatchMaxViews = 8;
constexpr int kBatchMaxLevels = 3;
constexpr int kBatchLevels = kBatchMaxLevels + 1;
// item tables: level passes (level 0 has none: it is the density),
// corner-quad passes, gather passes (one per level each; level 1 has none:
// it is gathered with level 0)
constexpr int kQuadPass = kBatchLevels;
constexpr int kGatherPass = 2 * kBatchLevels;
constexpr int kBatchPasses = 3 * kBatchLevels;
constexpr int kBlockVox = 512;  // voxels of a gather block (every shape)
constexpr int kFwdBatch = 4;    // samples per forward round trip
// the forward's item: a kTileW x kTileH-pixel tile, one thread per pixel.
// Smaller tiles wait less at an item's barriers for its slowest ray but pay
// the queue, record lookup and column set-up more often (c4 forward: 16x16
// 2.55 ms, 16x8 2.53, 8x8 2.59, 16x4 2.81)
constexpr int kTileW = 16, kTileH = 16, kTilePix = kTileW * kTileH;

// four floats in global memory (loads through it are global, not flat,
// 16-byte instructions)
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const f32x4 gcf32x4;
typedef __attribute__((address_space(1))) const float gcfloat;

// One view of a reconstruction. TRANSFORM_LINDEPTH samples pixel (x, y) at
// depth index q in object space at H.xyz / H.w, H = C (nx, ny, nz(q), 1) with
// C the inverse model-view times the inverse projection (the model-view is
// affine, so its product with the un-divided point commutes with the
// perspective divide) and nx = (x + 0.5) 2 / W - 1: at the first and last
// depth index (indices 0 and 1 for a single sample) H is n0 + x gx + y gy
// and f0 + x gx + y gy. pmv is the
// object -> clip projection used to find a voxel block's pixels.
struct ColumnView {
  float n0[4], f0[4], gx[4], gy[4];
  float pmv[16];
};

// Every pyramid level above 0 is stored with a one-cell zero border,
// (W+2)(H+2)(D+2) cells, from which the next level is averaged and the
// level's corner-quad copy (the forward's taps: two 16-byte loads per level)
// is taken; level 0 is read from the density. The gradients of levels 2 and
// up use the same bordered cells (their gathers write every interior cell,
// the border is never read); those of levels 0 and 1 stay in the level-0/1
// gather's LDS regions.
struct BatchRec {
  MipTable mt;            // dims: levels 0..L without the border
  const float* dens;
  float* pyr;             // bordered density pyramid, levels 1..L (level m at offset[m] - offset[1])
  float* gpyr;            // bordered gradients of levels 2..L (level m at offset[m] - offset[2]; the
                          // gather writes the interior)
  float4* quad[kBatchLevels];   // corner-quad copy of each level
  QuadDims qd[kBatchLevels];
  float* grad;            // level-0 gradient output
  const float* target;    // [V,H,W,3]
  float* images;          // [V,H,W,4]
  float* partials;        // per render item
  float* records;         // per render item, kTilePix float4 pixel records (lo, hi, alpha, beta)
  float* csave;           // per render item, out_d x kTilePix sample values c_k
  int* lodmax;            // per view, the largest LoD of its samples (float bits)
  float* loss;            // [1]
  float light[3];
  float step;
  float loss_scale;
  int num_views, out_w, out_h, out_d;
  int tiles_x, tiles_y;   // kTileW x kTileH-pixel tiles per view
  int bw, bh, bd;         // gather block (voxels)
  int item0;              // first render item
  int pass_item0[kBatchPasses];  // first item of the level, quad, gather and finishing passes
  ColumnView view[kBatchMaxViews];
};

// index of the reconstruction owning global item g of a pass (records sorted
// by first item; one without items in the pass shares its successor's first
// item and is never chosen)
__device__ __forceinline__ int rec_of(const BatchRec* __restrict__ recs, int nrec, int g, int pass) {
  int lo = 0, hi = nrec - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int first = pass < 0 ? recs[mid].item0 : recs[mid].pass_item0[pass];
    if (first <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// rec_of in a grid-stride loop: items only grow, so the owner of the next
// item is found by stepping on from the previous one (r < 0: none yet)
__device__ __forceinline__ int rec_from(const BatchRec* __restrict__ recs, int nrec, int g, int pass, int r) {
  if (r < 0) return rec_of(recs, nrec, g, pass);
  while (r + 1 < nrec && (pass < 0 ? recs[r + 1].item0 : recs[r + 1].pass_item0[pass]) <= g) ++r;
  return r;
}

// The device work queue of a persistent launch (the forward's (reconstruction,
// view, tile) items): each workgroup takes its next item from the pass's
// counter (one atomic per item), so items of uneven cost (reconstructions of
// different sizes and depths) spread over the workgroups as they finish, and
// each workgroup sees its items in increasing order. The first prep launch of
// a run zeroes the counters; stream order makes that visible to the later
// launches of the run. The other passes are launched one workgroup per item:
// their items are small or even, and the hardware's dispatch measured faster
// than a queue's atomic and two barriers per item.
enum BatchQueue { Q_FWD, kBatchQueues };

__device__ __forceinline__ int next_item(int* __restrict__ ctr, int* slot) {
  if (threadIdx.x == 0) *slot = atomicAdd(ctr, 1);
  __syncthreads();
  const int item = *slot;
  __syncthreads();  // the slot is rewritten by the next call
  return item;
}

// runs of 256 cells per item of the level and quad passes: a thread's runs
// are independent, so their loads are in flight together
constexpr int kPrepRuns = 2;

// flat index of bordered cell (x, y, z) (border coordinates: 0 and dims + 1
// are the border) of a level of dims d; the host keeps (H+2)(D+2) and W+2
// below 2^24
__device__ __forceinline__ unsigned pad_index(int x, int y, int z, Dims3 d) {
  return __umul24(__umul24((unsigned)z, (unsigned)(d.y + 2)) + (unsigned)y, (unsigned)(d.x + 2)) + (unsigned)x;
}

// bordered cells of pyramid level m >= 1 (pyr starts at level 1)
__device__ __forceinline__ float* level_cells(const BatchRec& R, int m) {
  return R.pyr + (R.mt.offset[m] - R.mt.offset[1]);
}

// bordered gradient cells of level m >= 2 (gpyr starts at level 2)
__device__ __forceinline__ float* grad_cells(const BatchRec& R, int m) {
  return R.gpyr + (R.mt.offset[m] - R.mt.offset[2]);
}

// ------------------------------------------------------------- prep passes
// Level m >= 1 of the zero-bordered pyramid (bordered cells, border cells get
// zero): kGenerateMip3D's pairwise means of level m-1 in z, then y, then x.
// Level 0 is the density itself: the means of level 1 read cells 2x+1 <=
// 2 (W>>1) - 1 < W only, so they never reach the border, and the level-0
// corner quads take their zeros from bounds checks.
__device__ __forceinline__ void batch_level_item(const BatchRec& R, int m, int local) {
#pragma clang fp contract(off)
  const Dims3 o = R.mt.dims[m], s = R.mt.dims[m - 1];
  // 32-bit cell indices: the host keeps every level's bordered cells below 2^31
  const unsigned px = (unsigned)o.x + 2u, py = (unsigned)o.y + 2u;
  const unsigned n = px * py * ((unsigned)o.z + 2u);
  const float* __restrict__ in = m == 1 ? R.dens : level_cells(R, m - 1);
  float* __restrict__ out = level_cells(R, m);
  auto at = [&](int cx, int cy, int cz) -> float {
    return m == 1 ? in[flat3(cx, cy, cz, s)] : in[pad_index(cx + 1, cy + 1, cz + 1, s)];
  };
  float v[kPrepRuns];
#pragma unroll
  for (int k = 0; k < kPrepRuns; ++k) {
    const unsigned f = ((unsigned)local * kPrepRuns + k) * 256u + threadIdx.x, fy = f / px;
    const int x = (int)(f - fy * px) - 1, y = (int)(fy % py) - 1, z = (int)(fy / py) - 1;
    v[k] = 0.f;
    if (f < n && x >= 0 && x < o.x && y >= 0 && y < o.y && z >= 0 && z < o.z) {
      float ya[2];
#pragma unroll
      for (int xi = 0; xi < 2; ++xi) {
        float za[2];
#pragma unroll
        for (int yi = 0; yi < 2; ++yi) {
          const float a = at(2 * x + xi, 2 * y + yi, 2 * z);
          const float b = at(2 * x + xi, 2 * y + yi, 2 * z + 1);
          za[yi] = (a + b) * 0.5f;
        }
        ya[xi] = (za[0] + za[1]) * 0.5f;
      }
      v[k] = (ya[0] + ya[1]) * 0.5f;
    }
  }
#pragma unroll
  for (int k = 0; k < kPrepRuns; ++k) {
    const unsigned f = ((unsigned)local * kPrepRuns + k) * 256u + threadIdx.x;
    if (f < n) out[f] = v[k];
  }
}

// Corner-quad copy of level m (the layout of the fused render's copy): entry
// (x, y, z) holds v(x..x+1, y..y+1, z) in bordered coordinates, zero outside
// the grid. The first item of level 0 also clears the views' LoD maxima.
__device__ __forceinline__ void batch_quad_item(const BatchRec& R, int m, int local) {
  using B = QuadBrick<float4>;
  if (m == 0 && local == 0 && threadIdx.x < (unsigned)R.num_views) R.lodmax[threadIdx.x] = 0;
  const QuadDims q = R.qd[m];
  const Dims3 d = R.mt.dims[m];
  const unsigned n = (unsigned)quad_entries(q);  // below 2^31 (host check)
  const float* __restrict__ lv = m == 0 ? R.dens : level_cells(R, m);
  float4* __restrict__ out = R.quad[m];
  float4 e[kPrepRuns];
#pragma unroll
  for (int k = 0; k < kPrepRuns; ++k) {
    const unsigned j = ((unsigned)local * kPrepRuns + k) * 256u + threadIdx.x;
    const unsigned b = (unsigned)(j / B::N), o = (unsigned)(j % B::N);
    const unsigned ix = (b % q.nbx) * B::X + (o & (B::X - 1));
    const unsigned iy = ((b / q.nbx) % q.nby) * B::Y + ((o >> B::Xl) & (B::Y - 1));
    const unsigned iz = (b / (q.nbx * q.nby)) * B::Z + (o >> (B::Xl + B::Yl));
    e[k] = make_float4(0.f, 0.f, 0.f, 0.f);  // padding entries of partial bricks: never read
    if (j < n && ix < q.qx && iy < q.qy && iz < q.qz) {
      if (m == 0) {
        const int x = (int)ix - 1, y = (int)iy - 1, z = (int)iz - 1;
        auto v = [&](int xx, int yy) -> float {
          return (((unsigned)xx < (unsigned)d.x) & ((unsigned)yy < (unsigned)d.y) & ((unsigned)z < (unsigned)d.z))
                     ? lv[flat3(xx, yy, z, d)]
                     : 0.f;
        };
        e[k] = make_float4(v(x, y), v(x + 1, y), v(x, y + 1), v(x + 1, y + 1));
      } else {
        const float* c = lv + pad_index((int)ix, (int)iy, (int)iz, d);
        const unsigned sy = (unsigned)(d.x + 2);
        e[k] = make_float4(c[0], c[1], c[sy], c[sy + 1]);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kPrepRuns; ++k) {
    const unsigned j = ((unsigned)local * kPrepRuns + k) * 256u + threadIdx.x;
    if (j < n) out[j] = e[k];
  }
}

// One launch per m: the items of level m+1 (they read level m), then those of
// the corner quads of level m.
__global__ __launch_bounds__(256) void k_batch_prep(const BatchRec* __restrict__ recs, int nrec, int m,
                                                    int level_items, int total_items, int* __restrict__ queues) {
  if (m == 0 && blockIdx.x == 0 && threadIdx.x < kBatchQueues) queues[threadIdx.x] = 0;
  int rl = -1, rq = -1;
  for (int item = blockIdx.x; item < total_items; item += gridDim.x) {
    if (item < level_items) {
      rl = rec_from(recs, nrec, item, m + 1, rl);
      const BatchRec& R = recs[rl];
      if (m + 1 <= R.mt.levels) batch_level_item(R, m + 1, item - R.pass_item0[m + 1]);
    } else {
      const int g = item - level_items;
      rq = rec_from(recs, nrec, g, kQuadPass + m, rq);
      const BatchRec& R = recs[rq];
      if (m <= R.mt.levels) batch_quad_item(R, m, g - R.pass_item0[kQuadPass + m]);
    }
  }
}

// ----------------------------------------------------------------- columns
// A pixel's column: positions p(i) = p0 + i dp from its first and last
// sample, and the LoD footprint's differences to the +x and +y pixels,
// e0 + i e1 (the +z difference is dp). The forward and the gather evaluate a
// column with this one function, so they see the same positions and LoDs.
struct Column {
  float3 p0, dp, ex0, ex1, ey0, ey1;
};

__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 scale3(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float3 fma3(float s, float3 a, float3 b) {
  return make_float3(fmaf(s, a.x, b.x), fmaf(s, a.y, b.y), fmaf(s, a.z, b.z));
}

__device__ __forceinline__ float3 column_point(const float* b, const ColumnView& v, float x, float y) {
  const float h0 = fmaf(y, v.gy[0], fmaf(x, v.gx[0], b[0])), h1 = fmaf(y, v.gy[1], fmaf(x, v.gx[1], b[1]));
  const float h2 = fmaf(y, v.gy[2], fmaf(x, v.gx[2], b[2])), h3 = fmaf(y, v.gy[3], fmaf(x, v.gx[3], b[3]));
  const float iw = __builtin_amdgcn_rcpf(h3);
  return make_float3(h0 * iw, h1 * iw, h2 * iw);
}

// 1 / (index of the column's second defining sample): S - 1, or 1 for a
// single sample (whose column is defined by indices 0 and 1, so that its step
// is the difference to the next depth index, which the LoD uses)
__device__ __forceinline__ float column_rcp(int S) { return S > 1 ? 1.f / (float)(S - 1) : 1.f; }

__device__ __forceinline__ Column column_of(const ColumnView& v, int x, int y, float inv_s1, bool lod) {
  const float fx = (float)x, fy = (float)y;
  Column c;
  const float3 pa = column_point(v.n0, v, fx, fy), pb = column_point(v.f0, v, fx, fy);
  c.p0 = pa;
  c.dp = scale3(sub3(pb, pa), inv_s1);
  c.ex0 = c.ex1 = c.ey0 = c.ey1 = make_float3(0.f, 0.f, 0.f);
  if (lod) {
    const float3 xa = sub3(pa, column_point(v.n0, v, fx + 1.f, fy)), xb = sub3(pb, column_point(v.f0, v, fx + 1.f, fy));
    const float3 ya = sub3(pa, column_point(v.n0, v, fx, fy + 1.f)), yb = sub3(pb, column_point(v.f0, v, fx, fy + 1.f));
    c.ex0 = xa;
    c.ex1 = scale3(sub3(xb, xa), inv_s1);
    c.ey0 = ya;
    c.ey1 = scale3(sub3(yb, ya), inv_s1);
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

// U samples' BORDER trilinear values at one level from its corner-quad copy,
// (p / 2^m - 0.5) in cell coordinates, every load in flight before any value
// is used: make_taps' inside test, corners and lerp order, as read_linear (an
// inside sample's corners are entries x, y in [-1, dims-1] and slices z, z+1
// in [-1, dims]); samples without use[u] read a clamped entry and give 0
template <int U>
__device__ __forceinline__ void quad_read_batch(const gcf32x4* __restrict__ quad, const QuadDims& qd, Dims3 d,
                                                const float3* p, float sc, const bool* use, float* val) {
  unsigned a0[U], a1[U];
  float w[U][3];
  bool in[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const float px = p[u].x * sc - 0.5f, py = p[u].y * sc - 0.5f, pz = p[u].z * sc - 0.5f;
    const float qx = px + 0.5f, qy = py + 0.5f, qz = pz + 0.5f;
    in[u] = use[u] & (0.f <= qx) & (qx < (float)d.x) & (0.f <= qy) & (qy < (float)d.y) & (0.f <= qz) &
            (qz < (float)d.z);
    const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
    w[u][0] = px - fx;
    w[u][1] = py - fy;
    w[u][2] = pz - fz;
    const unsigned ix = (unsigned)clampi((int)fx + 1, 0, (int)qd.qx - 1);
    const unsigned iy = (unsigned)clampi((int)fy + 1, 0, (int)qd.qy - 1);
    const unsigned iz = (unsigned)clampi((int)fz + 1, 0, (int)qd.qz - 2);
    a0[u] = quad_index(qd, ix, iy, iz);
    a1[u] = quad_index(qd, ix, iy, iz + 1);
  }
  f32x4 c0[U], c1[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    c0[u] = quad[a0[u]];
    c1[u] = quad[a1[u]];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const float v0 = lerpf(lerpf(c0[u].x, c0[u].y, w[u][0]), lerpf(c0[u].z, c0[u].w, w[u][0]), w[u][1]);
    const float v1 = lerpf(lerpf(c1[u].x, c1[u].y, w[u][0]), lerpf(c1[u].z, c1[u].w, w[u][0]), w[u][1]);
    val[u] = in[u] ? lerpf(v0, v1, w[u][2]) : 0.f;
  }
}

// ------------------------------------------------------------------ forward
// One thread per pixel of a kTileW x kTileH tile: the clipped column's BEER_LAMBERT
// march, the image, the pixel's loss term and gradient, and for the backward
// c_k = exp(-D_k) + E_{k-1} of every sample it marched (D_k the optical depth
// through sample k, E_{k-1} the emission before it): with E the pixel's
// total, dE/dsigma_k = exp(-D_k) - sum_{i>=k} exp(-D_i) sigma_i = c_k - E,
// so the gradient of sample k is alpha c_k + beta with alpha = step G and
// beta = -step G E, kept in the pixel's record (lo, hi, alpha, beta). c_k is
// stored item-major, [item][k][kTilePix pixels] (a wave's 64 pixels in one run).
// The view's largest LoD (the LoD is convex along a column, so a column's
// largest is at one of its ends) lets the gather skip the levels no sample
// of a view reads.
__global__ __launch_bounds__(kTilePix) void k_batch_fwd(const BatchRec* __restrict__ recs, int nrec,
                                                        int total_items, int* __restrict__ queues) {
  __shared__ float red[kTilePix];
  __shared__ int slot;
  int r = -1;
  for (int item = next_item(queues + Q_FWD, &slot); item < total_items; item = next_item(queues + Q_FWD, &slot)) {
    r = rec_from(recs, nrec, item, -1, r);
    const BatchRec& R = recs[r];
    const int local = item - R.item0;
    const int per_view = R.tiles_x * R.tiles_y;
    const int vi = local / per_view, t = local % per_view;
    const int x = (t % R.tiles_x) * kTileW + (int)(threadIdx.x % kTileW);
    const int y = (t / R.tiles_x) * kTileH + (int)(threadIdx.x / kTileW);
    const bool valid = x < R.out_w && y < R.out_h;
    const int S = R.out_d;
    const bool mips = R.mt.levels > 0;
    const SampleCfg cfg{F_LINEAR, mips ? MIP_LINEAR : MIP_NONE, R.mt.levels, 0.f, 0.5f};
    float D = 0.f, E = 0.f, sq = 0.f, lmax = 0.f;
    float4 rec = make_float4(__int_as_float(1), __int_as_float(0), 0.f, 0.f);
    if (valid) {
      const Column col = column_of(R.view[vi], x, y, column_rcp(S), mips);
      // samples outside the box [0, dims) read zero at every mip level (a
      // level-m sample is inside only where 0 <= p < dims_m 2^m <= dims), so
      // they add nothing to the blend or the gradient: the column is clipped
      int lo, hi;
      clip_ray(col.p0, col.dp, R.mt.dims[0], S, 0.f, lo, hi);
      float* __restrict__ cs = R.csave + (size_t)local * S * kTilePix + threadIdx.x;
      for (int ib = lo; ib <= hi; ib += kFwdBatch) {
        // kFwdBatch samples at a time: per mip level any of them reads, all
        // their loads in flight
        float3 p[kFwdBatch];
        int lv[kFwdBatch][2], nl[kFwdBatch];
        float lw[kFwdBatch][2], vf[kFwdBatch], vc[kFwdBatch];
#pragma unroll
        for (int u = 0; u < kFwdBatch; ++u) {
          const float fi = (float)(ib + u);
          p[u] = fma3(fi, col.dp, col.p0);
          nl[u] = level_taps(cfg, mips ? column_lod(col, fi) : 0.f, lv[u], lw[u]);
          vf[u] = vc[u] = 0.f;
        }
#pragma unroll
        for (int m = 0; m <= kBatchMaxLevels; ++m) {
          if (m > R.mt.levels) break;
          bool use[kFwdBatch], any = false;
#pragma unroll
          for (int u = 0; u < kFwdBatch; ++u) {
            use[u] = (ib + u <= hi) & ((lv[u][0] == m) | ((nl[u] == 2) & (lv[u][1] == m)));
            any |= use[u];
          }
          if (!any) continue;
          float val[kFwdBatch];
          quad_read_batch<kFwdBatch>((const gcf32x4*)R.quad[m], R.qd[m], R.mt.dims[m], p, 1.f / (float)(1 << m),
                                     use, val);
#pragma unroll
          for (int u = 0; u < kFwdBatch; ++u)
            if (use[u]) {
              if (lv[u][0] == m) vf[u] = val[u];
              else vc[u] = val[u];
            }
        }
#pragma unroll
        for (int u = 0; u < kFwdBatch; ++u) {
          if (ib + u > hi) break;
          const float sigma = R.step * (nl[u] == 2 ? lerpf(vf[u], vc[u], lw[u][1]) : vf[u]);
          D += sigma;
          const float e = __expf(-D);
          cs[(size_t)(ib + u) * kTilePix] = e + E;
          E = fmaf(e, sigma, E);
        }
      }
      if (mips && lo <= hi) lmax = fmaxf(column_lod(col, (float)lo), column_lod(col, (float)hi));
      const size_t pix = ((size_t)vi * R.out_h + y) * R.out_w + x;
      reinterpret_cast<float4*>(R.images)[pix] = make_float4(R.light[0] * E, R.light[1] * E, R.light[2] * E, D);
      const float* tg = R.target + pix * 3;
      const float e0 = R.light[0] * E - tg[0], e1 = R.light[1] * E - tg[1], e2 = R.light[2] * E - tg[2];
      sq = e0 * e0 + e1 * e1 + e2 * e2;
      const float G = R.loss_scale * 2.f * (R.light[0] * e0 + R.light[1] * e1 + R.light[2] * e2);
      if (G != 0.f && lo <= hi)
        rec = make_float4(__int_as_float(lo), __int_as_float(hi), R.step * G, -(R.step * G) * E);
    }
    reinterpret_cast<float4*>(R.records)[(size_t)local * kTilePix + threadIdx.x] = rec;
    for (int off = 32; off > 0; off >>= 1) lmax = fmaxf(lmax, __shfl_xor(lmax, off, 64));
    if ((threadIdx.x & 63) == 0 && lmax > 0.f) atomicMax(R.lodmax + vi, __float_as_int(lmax));
    // per-item loss partial: a fixed-order tree over the tile's pixels
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int w = kTilePix / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) R.partials[local] = red[0];
    __syncthreads();  // red is rewritten by the next item
  }
}

// ------------------------------------------------------------------- gather
// The transpose of the resample's trilinear scatter: one wave per voxel block,
// accumulated in wave-private FP64 LDS regions. PAIR: levels 0 and 1 at once,
// a block of bw x bh x bd level-0 voxels and the (bw/2) x (bh/2) x (bd/2)
// level-1 cells over it (the level-1 cells' support box contains the level-0
// voxels', so one walk over the candidate samples serves both); otherwise
// level m alone (m >= 2, for views whose samples reach it). For each view the
// wave projects the support box (level-m cell coordinates (x0-1, x0+bw) ...,
// i.e. object positions ((x0 - 0.5) 2^m, (x0 + bw + 0.5) 2^m) ...), takes the
// pixels whose columns can cross it (lanes: consecutive pixels in row order,
// the next pixel's record loaded while one is walked), cuts each column by
// the box, and adds every sample's gradient (alpha c_k + beta) times its
// level weight to the block's corners (weight prod_a max(0, 1 - |q_a -
// c_a|)); samples outside a level (BORDER) add nothing to it. The sums run in
// a fixed order and one store per voxel follows.
typedef __attribute__((address_space(3))) double lds_f64;  // an FP64 LDS region
constexpr int kGatherBatch = 4;  // c_k loaded per round trip
// voxel blocks (one per wave) per gather item: the waves of an item meet at
// its barriers, so with several the fast blocks wait for the slow one; one
// wave per workgroup measured best (c4: 4.84 ms against 4.95 with 2 and 5.15
// with 4)
constexpr int kGatherWaves = 1;
// copies of the level-1 region, lane l adding into copy l % kL1Copies: the
// level-1 cells are 8 times the volume of level 0's, so neighbouring pixels'
// samples often hit the same cell, and same-address LDS atomics of one
// instruction serialize; the copies are summed in a fixed order at the store
constexpr int kL1Copies = 4;
constexpr int kL1Stride = kBlockVox / 8 + 1;  // odd: the copies start on different banks
// the same for the block's own level: one copy (two, twice the LDS and the
// clearing, measured 5.05 ms against 4.74 for the c4 gathers)
constexpr int kL0Copies = 1;
constexpr int kL0Stride = kBlockVox + 1;

// add a sample's level tap to the block region acc (block W x H x Z with
// WXY = W H, origin o, (qx, qy, qz) in the level's cell coordinates, g the
// weighted gradient); corners outside the block are skipped
__device__ __forceinline__ void gather_tap(__attribute__((address_space(3))) double* acc, int WX, int WY, int WZ,
                                           int WXY, int ox, int oy, int oz, float qx, float qy, float qz, float g) {
  const float rx = qx - (float)ox, ry = qy - (float)oy, rz = qz - (float)oz;
  const float flx = floorf(rx), fly = floorf(ry), flz = floorf(rz);
  const float tx = rx - flx, ty = ry - fly, tz = rz - flz;
  const int lx = (int)flx, ly = (int)fly, lz = (int)flz;
  const float wx[2] = {g * (1.f - tx), g * tx}, wy[2] = {1.f - ty, ty}, wz[2] = {1.f - tz, tz};
  const bool vx[2] = {(unsigned)lx < (unsigned)WX, (unsigned)(lx + 1) < (unsigned)WX};
  const bool vy[2] = {(unsigned)ly < (unsigned)WY, (unsigned)(ly + 1) < (unsigned)WY};
  const bool vz[2] = {(unsigned)lz < (unsigned)WZ, (unsigned)(lz + 1) < (unsigned)WZ};
  const int base = __mul24(__mul24(lz, WY) + ly, WX) + lx;
#pragma unroll
  for (int kz = 0; kz < 2; ++kz) {
    if (!vz[kz]) continue;
#pragma unroll
    for (int ky = 0; ky < 2; ++ky) {
      if (!vy[ky]) continue;
      const float wyz = wy[ky] * wz[kz];
      const int row = base + (kz ? WXY : 0) + (ky ? WX : 0);
#pragma unroll
      for (int kx = 0; kx < 2; ++kx) {
        if (!vx[kx]) continue;
        __hip_atomic_fetch_add(acc + row + kx, (double)(wx[kx] * wyz), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
  }
}

// LINEAR-mip level weights (level_taps for lod >= 0 and bias 0): with
// l = min(lod, L), level floor(l) has weight 1 - frac(l) and level
// floor(l) + 1 has frac(l)
__device__ __forceinline__ float mip_weight(int m, int fl, float fr) {
  return m == fl ? 1.f - fr : (m == fl + 1 ? fr : 0.f);
}

__device__ __forceinline__ bool inside_level(float qx, float qy, float qz, Dims3 d) {
  return (0.f <= qx + 0.5f) & (qx + 0.5f < (float)d.x) & (0.f <= qy + 0.5f) & (qy + 0.5f < (float)d.y) &
         (0.f <= qz + 0.5f) & (qz + 0.5f < (float)d.z);
}

// One item of the gather: the block (blk = item's block) of reconstruction
// R at level m (PAIR: levels 0 and 1), with the block shape WX x WY x WZ as
// compile-time constants (index arithmetic and bounds on immediates).
template <bool PAIR, int WX, int WY, int WZ>
__device__ __forceinline__ void gather_item(const BatchRec& R, int m, int item, int wave, int lane, lds_f64* accs,
                                            lds_f64* acc, lds_f64* acc1s, lds_f64* acc1) {
  constexpr int kVox = kL0Copies * kL0Stride + (PAIR ? kL1Copies * kL1Stride : 0);
  const Dims3 dm = R.mt.dims[m];
  const bool two = PAIR && R.mt.levels > 0;
  const Dims3 d1 = R.mt.dims[two ? 1 : 0];
  const int nbx = (dm.x + WX - 1) / WX, nby = (dm.y + WY - 1) / WY, nbz = (dm.z + WZ - 1) / WZ;
  const int blk = (item - R.pass_item0[kGatherPass + m]) * kGatherWaves + wave;
  const bool live = m <= R.mt.levels && blk < nbx * nby * nbz;  // other waves only join the barriers
  const int x0 = (blk % nbx) * WX, y0 = ((blk / nbx) % nby) * WY, z0 = (blk / (nbx * nby)) * WZ;
  for (int k = lane; k < kVox; k += 64) accs[k] = 0.0;
  // support box (object space) of the block's cells; PAIR: of its level-1 cells
  const int mb = two ? 1 : m;
  const float s2 = (float)(1 << mb), sc = 1.f / (float)(1 << m);
  const int bx = two ? x0 / 2 : x0, by = two ? y0 / 2 : y0, bz = two ? z0 / 2 : z0;
  const int bwx = two ? WX / 2 : WX, bwy = two ? WY / 2 : WY, bwz = two ? WZ / 2 : WZ;
  const float b0[3] = {((float)bx - 0.5f) * s2, ((float)by - 0.5f) * s2, ((float)bz - 0.5f) * s2};
  const float b1[3] = {((float)(bx + bwx) + 0.5f) * s2, ((float)(by + bwy) + 0.5f) * s2,
                       ((float)(bz + bwz) + 0.5f) * s2};
  const int S = R.out_d;
  const float inv_s1 = column_rcp(S);
  const bool mips = R.mt.levels > 0;
  const float lmaxf = (float)R.mt.levels;
  const int per_view = R.tiles_x * R.tiles_y;
  const float4* __restrict__ recv = reinterpret_cast<const float4*>(R.records);
  for (int vi = 0; vi < R.num_views && live; ++vi) {
    if (!PAIR && __int_as_float(R.lodmax[vi]) <= (float)(m - 1)) continue;  // no sample reads level m
    const ColumnView& cv = R.view[vi];
    const float* A = cv.pmv;
    float mnx = 3.4e38f, mxx = -3.4e38f, mny = 3.4e38f, mxy = -3.4e38f;
    bool behind = false;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const float ox = (k & 1) ? b1[0] : b0[0], oy = (k & 2) ? b1[1] : b0[1], oz = (k & 4) ? b1[2] : b0[2];
      const float4 c = mat_vec(A, ox, oy, oz, 1.f);
      behind |= !(c.w > 0.f);
      const float iw = __builtin_amdgcn_rcpf(c.w);
      const float px = fmaf(c.x * iw, 0.5f, 0.5f) * (float)R.out_w, py = fmaf(c.y * iw, 0.5f, 0.5f) * (float)R.out_h;
      mnx = fminf(mnx, px);
      mxx = fmaxf(mxx, px);
      mny = fminf(mny, py);
      mxy = fmaxf(mxy, py);
    }
    int xs = 0, xe = R.out_w - 1, ys = 0, ye = R.out_h - 1;
    if (!behind) {
      xs = max(xs, (int)ceilf(mnx - 0.5f - 1e-2f));
      xe = min(xe, (int)floorf(mxx - 0.5f + 1e-2f));
      ys = max(ys, (int)ceilf(mny - 0.5f - 1e-2f));
      ye = min(ye, (int)floorf(mxy - 0.5f + 1e-2f));
    }
    if (xs > xe || ys > ye) continue;
    const int nxr = xe - xs + 1, cnt = nxr * (ye - ys + 1);
    // lane k of the box's row-major pixel list; the next record (k + 64) is
    // loaded while the current one is walked
    const int dy64 = 64 / nxr, dx64 = 64 - dy64 * nxr;
    int px = xs + lane % nxr, py = ys + lane / nxr;
    auto rec_at = [&](int qx, int qy, int& it, int& slot) {
      it = vi * per_view + (qy / kTileH) * R.tiles_x + qx / kTileW;
      slot = (qy % kTileH) * kTileW + qx % kTileW;
    };
    int it_n = 0, slot_n = 0;
    float4 rec_n = make_float4(__int_as_float(1), __int_as_float(0), 0.f, 0.f);
    if (lane < cnt) {
      rec_at(px, py, it_n, slot_n);
      rec_n = recv[(size_t)it_n * kTilePix + slot_n];
    }
    for (int k = lane; k < cnt; k += 64) {
      const float4 rec = rec_n;
      const int it = it_n, slot = slot_n, cpx = px, cpy = py;
      px += dx64;
      py += dy64;
      if (px > xe) {
        px -= nxr;
        ++py;
      }
      if (k + 64 < cnt) {
        rec_at(px, py, it_n, slot_n);
        rec_n = recv[(size_t)it_n * kTilePix + slot_n];
      }
      const int lo = __float_as_int(rec.x), hi = __float_as_int(rec.y);
      if (lo > hi) continue;
      const Column col = column_of(cv, cpx, cpy, inv_s1, mips);
      // the column's samples in the support box (one hundredth of a sample of
      // slack); a zero step component (reciprocal inf) gives (-inf, inf) for
      // a column inside that slab and an empty range outside it
      const float p0[3] = {col.p0.x, col.p0.y, col.p0.z}, dp[3] = {col.dp.x, col.dp.y, col.dp.z};
      float tlo = (float)lo, thi = (float)hi;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float ir = __builtin_amdgcn_rcpf(dp[a]);
        const float t0 = (b0[a] - p0[a]) * ir, t1 = (b1[a] - p0[a]) * ir;
        tlo = fmaxf(tlo, fminf(t0, t1));
        thi = fminf(thi, fmaxf(t0, t1));
      }
      if (!(tlo <= thi + 1.f)) continue;  // (also a NaN slab: 0 * inf)
      // tlo, thi lie within [lo, hi] here, so the conversions stay in range
      const int i0 = max(lo, (int)ceilf(tlo - 1e-2f)), i1 = min(hi, (int)floorf(thi + 1e-2f));
      const gcfloat* __restrict__ cs = (const gcfloat*)R.csave + ((size_t)it * S) * kTilePix + slot;
      for (int ib = i0; ib <= i1; ib += kGatherBatch) {
        float cv4[kGatherBatch];
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) cv4[u] = cs[(unsigned)min(ib + u, i1) * (unsigned)kTilePix];
#pragma unroll
        for (int u = 0; u < kGatherBatch; ++u) {
          const int i = ib + u;
          if (i > i1) break;
          const float fi = (float)i;
          const float l = mips ? fminf(column_lod(col, fi), lmaxf) : 0.f;
          const int fl = (int)l;
          const float fr = l - (float)fl;
          const float g = fmaf(rec.z, cv4[u], rec.w);
          const float3 p = fma3(fi, col.dp, col.p0);
          const float wm = mip_weight(m, fl, fr);
          if (wm != 0.f) {
            const float qx = p.x * sc - 0.5f, qy = p.y * sc - 0.5f, qz = p.z * sc - 0.5f;
            if (inside_level(qx, qy, qz, dm)) gather_tap(acc, WX, WY, WZ, WX * WY, x0, y0, z0, qx, qy, qz, g * wm);
          }
          if (two) {
            const float w1 = mip_weight(1, fl, fr);
            if (w1 != 0.f) {
              const float qx = p.x * 0.5f - 0.5f, qy = p.y * 0.5f - 0.5f, qz = p.z * 0.5f - 0.5f;
              if (inside_level(qx, qy, qz, d1))
                gather_tap(acc1, WX / 2, WY / 2, WZ / 2, (WX / 2) * (WY / 2), x0 / 2, y0 / 2, z0 / 2, qx, qy, qz,
                           g * w1);
            }
          }
        }
      }
    }
  }
  __syncthreads();
  if (live && PAIR) {
    // the density gradient: g_0 + sum_l 8^-l g_l(ancestor), level 1 from
    // this wave's region (the level-1 cells over the block), levels 2 and 3
    // from their gathers, which ran before; each level's value rounded to
    // float and accumulated by fma from level 1 up
    const int W1 = WX / 2, H1 = WY / 2;
    for (int k = lane; k < kBlockVox; k += 64) {
      const int lx = k % WX, ly = (k / WX) % WY, lz = k / (WX * WY);
      const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
      if (x >= dm.x || y >= dm.y || z >= dm.z) continue;
      float g = 0.f, s = 1.f;
      for (int l = 1; l <= R.mt.levels; ++l) {
        s *= 0.125f;
        const Dims3 o = R.mt.dims[l];
        const int px = x >> l, py = y >> l, pz = z >> l;
        if (px < o.x && py < o.y && pz < o.z) {
          float gl;
          if (l == 1) {
            const int c1 = (lx >> 1) + W1 * ((ly >> 1) + H1 * (lz >> 1));
            double t1 = acc1s[c1];
#pragma unroll
            for (int c = 1; c < kL1Copies; ++c) t1 += acc1s[c * kL1Stride + c1];
            gl = (float)t1;
          } else {
            gl = grad_cells(R, l)[pad_index(px + 1, py + 1, pz + 1, o)];
          }
          g = fmaf(s, gl, g);
        }
      }
      double t0 = accs[k];
#pragma unroll
      for (int c = 1; c < kL0Copies; ++c) t0 += accs[c * kL0Stride + k];
      R.grad[flat3(x, y, z, dm)] = (float)t0 + g;
    }
    if (blk == 0) {
      // the reconstruction's loss: lane-strided sums of its per-tile
      // partials, then a butterfly over the wave (a fixed order)
      const int np = R.num_views * R.tiles_x * R.tiles_y;
      float t = 0.f;
      for (int i = lane; i < np; i += 64) t += R.partials[i];
      for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
      if (lane == 0) R.loss[0] = t * R.loss_scale;
    }
  } else if (live) {
    float* __restrict__ gout = grad_cells(R, m);
    for (int k = lane; k < kBlockVox; k += 64) {
      const int x = x0 + k % WX, y = y0 + (k / WX) % WY, z = z0 + k / (WX * WY);
      if (x >= dm.x || y >= dm.y || z >= dm.z) continue;
      double t0 = accs[k];
#pragma unroll
      for (int c = 1; c < kL0Copies; ++c) t0 += accs[c * kL0Stride + k];
      gout[pad_index(x + 1, y + 1, z + 1, dm)] = (float)t0;
    }
  }
  __syncthreads();  // the regions are cleared by the next item
}

template <bool PAIR>
__global__ __launch_bounds__(256) void k_batch_gather(const BatchRec* __restrict__ recs, int nrec, int m,
                                                      int total_items) {
  constexpr int kVox = kL0Copies * kL0Stride + (PAIR ? kL1Copies * kL1Stride : 0);
  __shared__ double region[kGatherWaves][kVox];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  lds_f64* accs = (lds_f64*)region[wave];                        // the block's cells' copies
  lds_f64* acc = accs + (threadIdx.x % kL0Copies) * kL0Stride;    // this lane's copy
  lds_f64* acc1s = accs + kL0Copies * kL0Stride;                  // PAIR: the level-1 cells' copies
  lds_f64* acc1 = acc1s + (threadIdx.x % kL1Copies) * kL1Stride;  // this lane's copy
  int r = -1;
  for (int item = blockIdx.x; item < total_items; item += gridDim.x) {
    r = rec_from(recs, nrec, item, kGatherPass + m, r);
    const BatchRec& R = recs[r];
    // the reconstruction's block shape (bricks 0, 1, 2: 16x16x2, 8x8x8, 16x8x4)
    if (R.bd == 2) gather_item<PAIR, 16, 16, 2>(R, m, item, wave, lane, accs, acc, acc1s, acc1);
    else if (R.bd == 8) gather_item<PAIR, 8, 8, 8>(R, m, item, wave, lane, accs, acc, acc1s, acc1);
    else gather_item<PAIR, 16, 8, 4>(R, m, item, wave, lane, accs, acc, acc1s, acc1);
  }
}

// ------------------------------------------------------------------ host
static size_t align256b(size_t b) { return (b + 255) & ~size_t(255); }

static size_t padded_cells(Dims3 d) { return (size_t)(d.x + 2) * (d.y + 2) * (d.z + 2); }

// levels 0..L (L <= num_mipmaps, while every axis halves to at least one
// cell), each bordered and starting on a 256-byte boundary
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
  mt.base_stride = (size_t)d.x * d.y * d.z;
  size_t off = 0;
  for (int m = 0; m <= levels; ++m) {
    mt.dims[m] = Dims3{d.x >> m, d.y >> m, d.z >> m};
    mt.offset[m] = off;
    off += (padded_cells(mt.dims[m]) + 63) & ~size_t(63);
  }
  mt.atlas_stride = off;
  return mt;
}

// gather block (voxels) of each brick choice: 16x16x2, 8x8x8, 16x8x4
static void brick_block(int brick, int& bw, int& bh, int& bd) {
  bw = brick == 1 ? 8 : 16;
  bh = brick == 0 ? 16 : 8;
  bd = brick == 0 ? 2 : (brick == 1 ? 8 : 4);
}

// workspace layout: the record table, then per reconstruction its bordered
// density pyramid above level 0, the bordered gradients of levels 2 and up, the
// corner-quad copy of each level, the render partials, the views' LoD maxima,
// the pixel records and the sample values c_k (items x out_d x kTilePix floats:
// the largest part, 4 bytes per frustum sample)
struct BatchLayout {
  size_t table, queues, total;
  std::vector<size_t> pyr, gpyr, partials, lodmax, records, csave;
  std::vector<std::array<size_t, kBatchLevels>> quad;
};

// items of pass p (prep passes: 256 threads each) for a reconstruction with pyramid mt and
// gather block bw x bh x bd
static int pass_items(const MipTable& mt, int bw, int bh, int bd, int p) {
  constexpr size_t kPrep = 256 * kPrepRuns;
  if (p < kQuadPass) return p >= 1 && p <= mt.levels ? (int)((padded_cells(mt.dims[p]) + kPrep - 1) / kPrep) : 0;
  if (p < kGatherPass) {
    const int m = p - kQuadPass;
    return m <= mt.levels ? (int)((quad_entries(make_quad_dims(mt.dims[m])) + kPrep - 1) / kPrep) : 0;
  }
  const int m = p - kGatherPass;
  if (m > mt.levels || m == 1) return 0;  // level 1 is gathered with level 0
  const Dims3 d = mt.dims[m];
  const int64_t blocks = (int64_t)((d.x + bw - 1) / bw) * ((d.y + bh - 1) / bh) * ((d.z + bd - 1) / bd);
  return (int)((blocks + kGatherWaves - 1) / kGatherWaves);  // a block per wave
}

static int batch_layout(int32_t n, const gft_batch_recon* rs, int32_t num_mipmaps, BatchLayout& L) {
  if (n < 1) return fail(GFT_ERR_INVALID_ARGUMENT, "the batch needs at least one reconstruction");
  if (!rs) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  if (num_mipmaps < 0 || num_mipmaps > kBatchMaxLevels)
    return fail(GFT_ERR_INVALID_ARGUMENT, "num_mipmaps must be in [0, %d]", kBatchMaxLevels);
  size_t off = align256b(sizeof(BatchRec) * (size_t)n);
  L.queues = off;
  off += align256b(kBatchQueues * sizeof(int));
  L.table = off;
  L.pyr.resize(n);
  L.gpyr.resize(n);
  L.records.resize(n);
  L.csave.resize(n);
  L.lodmax.resize(n);
  L.partials.resize(n);
  L.quad.resize(n);
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
    // bordered cell indices are formed with 24-bit multiplies and stay 32-bit
    if ((int64_t)(d.y + 2) * (d.z + 2) >= ((int64_t)1 << 24) || d.x + 2 >= (1 << 24) ||
        (int64_t)(d.x + 4) * (d.y + 2) * (d.z + 3) >= ((int64_t)1 << 31))
      return fail(GFT_ERR_INVALID_ARGUMENT,
                  "reconstruction %d: grid too large for the batched step ((H+2)(D+2) < 2^24, cells < 2^31)", r);
    const MipTable mt = batch_mip_table(d, num_mipmaps);
    L.pyr[r] = off;
    if (mt.levels >= 1) off += align256b((mt.atlas_stride - mt.offset[1]) * sizeof(float));
    L.gpyr[r] = off;
    if (mt.levels >= 2) off += align256b((mt.atlas_stride - mt.offset[2]) * sizeof(float));
    for (int m = 0; m <= mt.levels; ++m) {
      L.quad[r][m] = off;
      off += align256b(quad_entries(make_quad_dims(mt.dims[m])) * sizeof(float4));
    }
    const size_t items = (size_t)c.num_views * ((c.output_shape[2] + kTileW - 1) / kTileW) *
                         ((c.output_shape[1] + kTileH - 1) / kTileH);
    L.partials[r] = off;
    off += align256b(items * sizeof(float));
    L.lodmax[r] = off;
    off += align256b(kBatchMaxViews * sizeof(int));
    L.records[r] = off;
    off += align256b(items * kTilePix * 4 * sizeof(float));
    L.csave[r] = off;
    off += align256b(items * (size_t)c.output_shape[0] * kTilePix * sizeof(float));
  }
  L.total = off;
  return GFT_OK;
}

// item counts of the launches: render items; the level passes (bordered
// cells of each level), the quad passes and the finishing pass (level-0
// voxels)
static void batch_items(int32_t n, const gft_batch_recon* rs, int32_t num_mipmaps, int& render, int pass[kBatchPasses]) {
  render = 0;
  for (int p = 0; p < kBatchPasses; ++p) pass[p] = 0;
  for (int r = 0; r < n; ++r) {
    const gft_batch_recon& c = rs[r];
    render += c.num_views * ((c.output_shape[2] + kTileW - 1) / kTileW) * ((c.output_shape[1] + kTileH - 1) / kTileH);
    const MipTable mt = batch_mip_table(Dims3{c.grid_dhw[2], c.grid_dhw[1], c.grid_dhw[0]}, num_mipmaps);
    int bw, bh, bd;
    brick_block(c.brick, bw, bh, bd);
    for (int p = 0; p < kBatchPasses; ++p) pass[p] += pass_items(mt, bw, bh, bd, p);
  }
}

// Workgroups of the persistent forward: as many as the device holds at once
// (the kernel's occupancy per CU times the CUs); a workgroup that starts late
// finds fewer items left, so the count only needs to fill the chip. Queried
// once per device (by prepare, outside any graph capture).
struct ResidentCap {
  int dev = -1, cap = 0;
};
static std::mutex g_caps_mu;
static std::vector<ResidentCap> g_caps;

static int forward_cap(int& cap) {
  int dev = 0, rc;
  if ((rc = check_hip(hipGetDevice(&dev), "current device"))) return rc;
  std::lock_guard<std::mutex> lock(g_caps_mu);
  for (const ResidentCap& c : g_caps)
    if (c.dev == dev) {
      cap = c.cap;
      return GFT_OK;
    }
  int cus = 0, per_cu = 0;
  if ((rc = check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count")))
    return rc;
  if ((rc = check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                          &per_cu, reinterpret_cast<const void*>(&k_batch_fwd), kTilePix, 0),
                      "occupancy query")))
    return rc;
  cap = std::max(1, per_cu) * std::max(1, cus);
  g_caps.push_back(ResidentCap{dev, cap});
  return GFT_OK;
}

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
  int item = 0, pass_item[kBatchPasses] = {};
  for (int r = 0; r < n; ++r) {
    const gft_batch_recon& c = recons[r];
    if (!c.density || !c.targets || !c.images || !c.density_grad || !c.loss || !c.M || !c.V || !c.P || !c.frustum)
      return fail(GFT_ERR_INVALID_ARGUMENT, "reconstruction %d: null pointer argument", r);
    BatchRec& R = recs[r];
    std::memset(&R, 0, sizeof(R));
    const Dims3 d{c.grid_dhw[2], c.grid_dhw[1], c.grid_dhw[0]};
    R.mt = batch_mip_table(d, num_mipmaps);
    R.dens = c.density;
    R.pyr = (float*)(ws + L.pyr[r]);
    R.gpyr = (float*)(ws + L.gpyr[r]);
    R.grad = c.density_grad;
    R.target = c.targets;
    R.images = c.images;
    R.partials = (float*)(ws + L.partials[r]);
    R.lodmax = (int*)(ws + L.lodmax[r]);
    R.records = (float*)(ws + L.records[r]);
    R.csave = (float*)(ws + L.csave[r]);
    R.loss = c.loss;
    for (int k = 0; k < 3; ++k) R.light[k] = c.light[k];
    R.step = c.depth_step;
    R.loss_scale = c.loss_scale;
    R.num_views = c.num_views;
    R.out_d = c.output_shape[0];
    R.out_h = c.output_shape[1];
    R.out_w = c.output_shape[2];
    R.tiles_x = (R.out_w + kTileW - 1) / kTileW;
    R.tiles_y = (R.out_h + kTileH - 1) / kTileH;
    brick_block(c.brick, R.bw, R.bh, R.bd);
    R.item0 = item;
    item += R.num_views * R.tiles_x * R.tiles_y;
    for (int p = 0; p < kBatchPasses; ++p) {
      R.pass_item0[p] = pass_item[p];
      pass_item[p] += pass_items(R.mt, R.bw, R.bh, R.bd, p);
    }
    for (int m = 0; m <= R.mt.levels; ++m) {
      R.quad[m] = (float4*)(ws + L.quad[r][m]);
      R.qd[m] = make_quad_dims(R.mt.dims[m]);
    }
    const Mat4d m = mat_from_colmajor(c.M);
    const double W = R.out_w, Hh = R.out_h, S = R.out_d;
    for (int v = 0; v < c.num_views; ++v) {
      const Mat4d mv = mat_mul(mat_from_colmajor(c.V + 16 * v), m), p = mat_from_colmajor(c.P + 16 * v);
      Mat4d mvinv, pinv;
      if (!mat_inverse(mv, mvinv)) return fail(GFT_ERR_INVALID_ARGUMENT, "model-view matrix is singular");
      if (!mat_inverse(p, pinv)) return fail(GFT_ERR_INVALID_ARGUMENT, "projection matrix is singular");
      // object position of pixel (x, y) at depth index q: (C (nx, ny, nz(q), 1)).xyz / .w with C = mv^-1 p^-1,
      // nx = (x + 0.5) 2 / W - 1 and the TRANSFORM_LINDEPTH depth nz(q) = -p22 - p32 / z, z = -(near + (far -
      // near) q / S) (transformations.hpp:86-105)
      const Mat4d C = mat_mul(mvinv, pinv);
      const double p22 = c.P[16 * v + 4 * 2 + 2], p32 = c.P[16 * v + 4 * 3 + 2];
      const double nearp = c.frustum[6 * v + 0], farp = c.frustum[6 * v + 1];
      auto nz_of = [&](double q) { return -p22 - p32 / (-(nearp + (farp - nearp) * (q / S))); };
      ColumnView& cv = R.view[v];
      const double ax = 1.0 / W - 1.0, ay = 1.0 / Hh - 1.0;
      for (int r4 = 0; r4 < 4; ++r4) {
        const double c0 = C.m[0 * 4 + r4], c1 = C.m[1 * 4 + r4], c2 = C.m[2 * 4 + r4], c3 = C.m[3 * 4 + r4];
        cv.n0[r4] = (float)(c0 * ax + c1 * ay + c2 * nz_of(0.5) + c3);
        cv.f0[r4] = (float)(c0 * ax + c1 * ay + c2 * nz_of(S > 1 ? S - 0.5 : 1.5) + c3);
        cv.gx[r4] = (float)(c0 * 2.0 / W);
        cv.gy[r4] = (float)(c1 * 2.0 / Hh);
      }
      mat_to_float(mat_mul(p, mv), cv.pmv);
    }
  }
  int cap;  // the forward's launch size, queried here (never under capture)
  if ((rc = forward_cap(cap))) return rc;
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
  int* queues = (int*)((char*)workspace + L.queues);
  int cap;
  if ((rc = forward_cap(cap))) return rc;
  int render = 0, pass[kBatchPasses];
  batch_items(n, recons, num_mipmaps, render, pass);
  {
    KernelTimer t(s, "batch_levels");
    for (int m = 0; m <= kBatchMaxLevels; ++m) {
      const int nl = m < kBatchMaxLevels ? pass[m + 1] : 0, total = nl + pass[kQuadPass + m];
      if (total)
        hipLaunchKernelGGL(k_batch_prep, dim3(total), dim3(256), 0, s, recs, n, m, nl, total, queues);
    }
  }
  if ((rc = check_hip(hipGetLastError(), "batched mip levels"))) return rc;
  {
    KernelTimer t(s, "batch_fwd");
    hipLaunchKernelGGL(k_batch_fwd, dim3(std::max(1, std::min(render, cap))), dim3(kTilePix), 0, s, recs, n, render,
                       queues);
  }
  if ((rc = check_hip(hipGetLastError(), "batched render forward"))) return rc;
  {
    KernelTimer t(s, "batch_gather");
    // the coarse levels first: the level-0/1 gather adds them into the density gradient
    for (int m = kBatchMaxLevels; m >= 2; --m)
      if (pass[kGatherPass + m])
        hipLaunchKernelGGL(k_batch_gather<false>, dim3(pass[kGatherPass + m]), dim3(64 * kGatherWaves), 0, s, recs, n, m,
                           pass[kGatherPass + m]);
    hipLaunchKernelGGL(k_batch_gather<true>, dim3(pass[kGatherPass]), dim3(64 * kGatherWaves), 0, s, recs, n, 0,
                       pass[kGatherPass]);
  }
  return check_hip(hipGetLastError(), "batched render gather");
}

}  // extern "C"
