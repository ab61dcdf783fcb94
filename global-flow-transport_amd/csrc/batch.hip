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
//   k_batch_level(m)   level m of the zero-bordered pyramid (m = 0: the
//                      density; m >= 1: kGenerateMip3D's pairwise means of
//                      level m-1) and a cleared gradient level;
//   k_batch_fwd        items (reconstruction, view, 16x16-pixel tile): a
//                      ray, clipped to the grid box, marches its samples
//                      keeping each sample's c_k, writes its pixel, its loss
//                      term and its gradient record;
//   k_batch_bwd        the same items: every sample's gradient alpha c_k +
//                      beta scattered through its level taps; level-0 and
//                      level-1 taps go to FP64 LDS bricks per tile (or 8x8
//                      sub-tile) and segment of at most K samples, flushed
//                      with float atomics; coarser levels (rare) take float
//                      atomics directly;
//   k_batch_finish     the density gradient: every voxel's level-0 gradient
//                      plus 1/8^m of its level-m ancestor's (the mip
//                      adjoint); in the first item of each reconstruction,
//                      its loss: the per-tile partials summed in a fixed
//                      order.
// The brick of the backward (tile size x longest segment) is chosen per
// reconstruction: 16x16-pixel tiles with 8-sample segments (about 8^3 cells
// at two pixels per cell), 8x8 tiles with 4-sample segments (about 4^3), or
// 8x8 tiles with 8 (about 4x4x8). A segment whose bricks would not fit the
// LDS is halved until they do.
#include <algorithm>
#include <array>
#include <cstring>
#include <vector>

#include "gft_host.hpp"
#include "quad.hpp"
#include "sampler_v1.hpp"

namespace gft {

constexpr int kBatchMaxViews = 8;
constexpr int kBatchMaxLevels = 3;
constexpr int kBatchLevels = kBatchMaxLevels + 1;
// item tables: level passes 0..3, corner-quad passes 0..3, the finishing pass
constexpr int kBatchPasses = 2 * kBatchLevels + 1;
constexpr int kFinishPass = kBatchPasses - 1;
constexpr int kBatchCap = 6144;  // FP64 brick cells per workgroup (48 KiB)

// four floats in global memory (loads through it are global, not flat,
// 16-byte instructions)
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const f32x4 gcf32x4;

// Every pyramid level is stored with a one-cell zero border, (W+2)(H+2)(D+2)
// cells, from which the next level is averaged and the level's corner-quad
// copy (the forward's taps: two 16-byte loads per level) is taken. The
// gradient levels have the same border: the backward's taps of corners
// outside the grid land there and nothing reads them.
struct BatchRec {
  MipTable mt;            // dims: levels 0..L without the border; offset[m]: level m in pyr and gpyr
  const float* dens;
  float* pyr;             // bordered density pyramid, levels 0..L
  float* gpyr;            // bordered gradients of the same cells
  float4* quad[kBatchLevels];   // corner-quad copy of each level
  QuadDims qd[kBatchLevels];
  float* grad;            // level-0 gradient output
  const float* target;    // [V,H,W,3]
  float* images;          // [V,H,W,4]
  float* partials;        // per render item
  float* records;         // per render item, 256 float4 pixel records (lo, hi, alpha, beta)
  float* csave;           // per render item, out_d x 256 sample values c_k
  float* loss;            // [1]
  float light[3];
  float step;
  float loss_scale;
  int num_views, out_w, out_h, out_d;
  int tiles_x, tiles_y;   // 16x16-pixel tiles per view
  int tile;               // brick tile: 16 (one brick per workgroup) or 8 (one per wave)
  int seg;                // longest backward segment in samples
  int item0;              // first render item
  int pass_item0[kBatchPasses];  // first item of the level, quad and finishing passes
  SampleView view[kBatchMaxViews];
};

// index of the reconstruction owning global item g of a pass (records sorted by first item)
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

// flat index of bordered cell (x, y, z) (border coordinates: 0 and dims + 1
// are the border) of a level of dims d; the host keeps (H+2)(D+2) and W+2
// below 2^24
__device__ __forceinline__ unsigned pad_index(int x, int y, int z, Dims3 d) {
  return __umul24(__umul24((unsigned)z, (unsigned)(d.y + 2)) + (unsigned)y, (unsigned)(d.x + 2)) + (unsigned)x;
}

// ------------------------------------------------------------ level passes
// Over the bordered cells of level m (border cells get zero): m = 0 copies
// the density; m >= 1 takes pairwise means of level m-1 in z, then y, then x
// (kGenerateMip3D's order). Both clear the level's gradient.
__global__ __launch_bounds__(256) void k_batch_level(const BatchRec* __restrict__ recs, int nrec, int m,
                                                     int total_items) {
#pragma clang fp contract(off)
  for (int item = blockIdx.x; item < total_items; item += gridDim.x) {
    const BatchRec& R = recs[rec_of(recs, nrec, item, m)];
    if (m > R.mt.levels) continue;
    const Dims3 o = R.mt.dims[m];
    const int px = o.x + 2, py = o.y + 2;
    const size_t n = (size_t)px * py * (o.z + 2);
    const size_t f = (size_t)(item - R.pass_item0[m]) * 256 + threadIdx.x;
    if (f >= n) continue;
    const int x = (int)(f % (size_t)px) - 1, y = (int)((f / (size_t)px) % (size_t)py) - 1,
              z = (int)(f / ((size_t)px * py)) - 1;
    float v = 0.f;
    if (x >= 0 && x < o.x && y >= 0 && y < o.y && z >= 0 && z < o.z) {
      if (m == 0) {
        v = R.dens[flat3(x, y, z, o)];
      } else {
        const Dims3 s = R.mt.dims[m - 1];
        const float* in = R.pyr + R.mt.offset[m - 1];
        float ya[2];
#pragma unroll
        for (int xi = 0; xi < 2; ++xi) {
          float za[2];
#pragma unroll
          for (int yi = 0; yi < 2; ++yi) {
            const float a = in[pad_index(2 * x + xi + 1, 2 * y + yi + 1, 2 * z + 1, s)];
            const float b = in[pad_index(2 * x + xi + 1, 2 * y + yi + 1, 2 * z + 2, s)];
            za[yi] = (a + b) * 0.5f;
          }
          ya[xi] = (za[0] + za[1]) * 0.5f;
        }
        v = (ya[0] + ya[1]) * 0.5f;
      }
    }
    R.pyr[R.mt.offset[m] + f] = v;
    R.gpyr[R.mt.offset[m] + f] = 0.f;
  }
}

// Corner-quad copy of level m (the layout of the fused render's copy), taken
// from the bordered level: entry (x, y, z) holds v(x..x+1, y..y+1, z), the
// border supplying the zeros of corners outside the grid.
__global__ __launch_bounds__(256) void k_batch_quad(const BatchRec* __restrict__ recs, int nrec, int m,
                                                    int total_items) {
  using B = QuadBrick<float4>;
  for (int item = blockIdx.x; item < total_items; item += gridDim.x) {
    const BatchRec& R = recs[rec_of(recs, nrec, item, kBatchLevels + m)];
    if (m > R.mt.levels) continue;
    const QuadDims q = R.qd[m];
    const size_t j = (size_t)(item - R.pass_item0[kBatchLevels + m]) * 256 + threadIdx.x;
    if (j >= quad_entries(q)) continue;
    const unsigned b = (unsigned)(j / B::N), o = (unsigned)(j % B::N);
    const unsigned ix = (b % q.nbx) * B::X + (o & (B::X - 1));
    const unsigned iy = ((b / q.nbx) % q.nby) * B::Y + ((o >> B::Xl) & (B::Y - 1));
    const unsigned iz = (b / (q.nbx * q.nby)) * B::Z + (o >> (B::Xl + B::Yl));
    float4 e = make_float4(0.f, 0.f, 0.f, 0.f);  // padding entries of partial bricks: never read
    if (ix < q.qx && iy < q.qy && iz < q.qz) {
      const Dims3 d = R.mt.dims[m];
      const float* c = R.pyr + R.mt.offset[m] + pad_index((int)ix, (int)iy, (int)iz, d);
      const unsigned sy = (unsigned)(d.x + 2);
      e = make_float4(c[0], c[1], c[sy], c[sy + 1]);
    }
    R.quad[m][j] = e;
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

// BORDER trilinear value from a level's corner-quad copy, (px, py, pz) in
// cell coordinates (p / 2^m - 0.5): make_taps' inside test, corners and lerp
// order, as read_linear (an inside sample's corners are entries x, y in
// [-1, dims-1] and slices z, z+1 in [-1, dims])
__device__ __forceinline__ float quad_read(const gcf32x4* __restrict__ quad, const QuadDims& qd, Dims3 d, float px,
                                           float py, float pz) {
  const float qx = px + 0.5f, qy = py + 0.5f, qz = pz + 0.5f;
  if (!((0.f <= qx) & (qx < (float)d.x) & (0.f <= qy) & (qy < (float)d.y) & (0.f <= qz) & (qz < (float)d.z)))
    return 0.f;
  const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
  const float tx = px - fx, ty = py - fy, tz = pz - fz;
  const unsigned ix = (unsigned)((int)fx + 1), iy = (unsigned)((int)fy + 1), iz = (unsigned)((int)fz + 1);
  const f32x4 c0 = quad[quad_index(qd, ix, iy, iz)], c1 = quad[quad_index(qd, ix, iy, iz + 1)];
  const float v0 = lerpf(lerpf(c0.x, c0.y, tx), lerpf(c0.z, c0.w, tx), ty);
  const float v1 = lerpf(lerpf(c1.x, c1.y, tx), lerpf(c1.z, c1.w, tx), ty);
  return lerpf(v0, v1, tz);
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
      const float a = quad_read((const gcf32x4*)R.quad[m], R.qd[m], R.mt.dims[m], p.x * sc - 0.5f,
                                p.y * sc - 0.5f, p.z * sc - 0.5f);
      if (f) vf = a;
      else vc = a;
    }
  }
  return nl == 2 ? lerpf(vf, vc, lw[1]) : vf;
}

// Bounding bricks of levels 0 and 1 (bordered cell coordinates) of the
// samples ia..ib of a tile's pixel columns: positions are affine along a
// column and, at one depth, across pixels, so the four corner columns at the
// two segment ends bound every sample of the tile. A tap outside its brick
// takes a float atomic, so the bound only decides where a tap is summed,
// never whether.
struct BatchBrick {
  int lo0[3], n0[3], lo1[3], n1[3];
  int size0, size;
};

__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

// j / n for 0 <= j, n < 2^20 with rn = 1 / n in float: the float quotient
// is within 2^20 * 2^-23 = 0.125 of j / n, so truncation is the quotient or
// one off, and one correction step makes it exact
__device__ __forceinline__ int div_small(int j, int n, float rn) {
  const int q = (int)((float)j * rn);
  const int r = j - q * n;
  return q + (r >= n) - (r < 0);
}

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
    int lo = clampi((int)floorf(mn[a] - 0.5f - 1e-3f) + 1, 0, d0[a] + 1);
    int hi = clampi((int)floorf(mx[a] - 0.5f + 1e-3f) + 2, 0, d0[a] + 1);
    b.lo0[a] = uniform_i(lo);
    b.n0[a] = uniform_i(hi - lo + 1);
    s0 *= b.n0[a];
    if (mt.levels > 0) {
      lo = clampi((int)floorf(mn[a] * 0.5f - 0.5f - 1e-3f) + 1, 0, d1[a] + 1);
      hi = clampi((int)floorf(mx[a] * 0.5f - 0.5f + 1e-3f) + 2, 0, d1[a] + 1);
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

// gradient of one level tap (an inside sample's eight corners, weight
// products in the order ((g w_x) w_y) w_z, corners outside the grid landing
// in the border): into the level's LDS brick when the whole corner cube lies
// in it, float atomics into the bordered gradient level otherwise
__device__ __forceinline__ void level_scatter(double* __restrict__ lds, bool staged, const int blo[3],
                                              const int bn[3], float* __restrict__ gbuf, Dims3 d, float px, float py,
                                              float pz, float gw) {
  const float qx = px + 0.5f, qy = py + 0.5f, qz = pz + 0.5f;
  if (!((0.f <= qx) & (qx < (float)d.x) & (0.f <= qy) & (qy < (float)d.y) & (0.f <= qz) & (qz < (float)d.z)))
    return;
  const float fx = floorf(px), fy = floorf(py), fz = floorf(pz);
  const float tx = px - fx, ty = py - fy, tz = pz - fz;
  const float wx[2] = {1.f - tx, tx}, wy[2] = {1.f - ty, ty}, wz[2] = {1.f - tz, tz};
  float w[2][2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const float gx = gw * wx[i];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const float gxy = gx * wy[j];
#pragma unroll
      for (int k = 0; k < 2; ++k) w[k][j][i] = gxy * wz[k];
    }
  }
  const int ix = (int)fx + 1, iy = (int)fy + 1, iz = (int)fz + 1;
  const int lx = ix - blo[0], ly = iy - blo[1], lz = iz - blo[2];
  if (staged & (lx >= 0) & (ly >= 0) & (lz >= 0) & (lx + 1 < bn[0]) & (ly + 1 < bn[1]) & (lz + 1 < bn[2])) {
    const int sy = bn[0], sz = bn[0] * bn[1];
    double* c = lds + (lz * bn[1] + ly) * bn[0] + lx;
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 2; ++i) atomicAdd(c + k * sz + j * sy + i, (double)w[k][j][i]);
  } else {
    const unsigned sy = (unsigned)(d.x + 2), sz = __umul24(sy, (unsigned)(d.y + 2));
    float* c = gbuf + pad_index(ix, iy, iz, d);
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 2; ++i) atomic_addf(c + k * sz + j * sy + i, w[k][j][i]);
  }
}

// ---- forward
// One thread per pixel of a 16x16 tile: the clipped column's BEER_LAMBERT
// march, the image, the pixel's loss term and gradient, and for the backward
// c_k = exp(-D_k) + E_{k-1} of every sample it marched (D_k the optical depth
// through sample k, E_{k-1} the emission before it): with E the pixel's
// total, dE/dsigma_k = exp(-D_k) - sum_{i>=k} exp(-D_i) sigma_i = c_k - E,
// so the gradient of sample k is alpha c_k + beta with alpha = step G and
// beta = -step G E, kept in the pixel's record (lo, hi, alpha, beta). c_k is
// stored item-major, [item][k][256 pixels] (a wave's 64 pixels in one run).
__global__ __launch_bounds__(256) void k_batch_fwd(const BatchRec* __restrict__ recs, int nrec, int total_items) {
  __shared__ float red[256];
  for (int item = blockIdx.x; item < total_items; item += gridDim.x) {
    const BatchRec& R = recs[rec_of(recs, nrec, item, -1)];
    const int local = item - R.item0;
    const int per_view = R.tiles_x * R.tiles_y;
    const int vi = local / per_view, t = local % per_view;
    const int x = (t % R.tiles_x) * 16 + (threadIdx.x & 15), y = (t / R.tiles_x) * 16 + (threadIdx.x >> 4);
    const SampleView& V = R.view[vi];
    const bool valid = x < R.out_w && y < R.out_h;
    const int S = R.out_d;
    const float3 oinv = make_float3(1.f / (float)R.out_w, 1.f / (float)R.out_h, 1.f / (float)S);
    const bool mips = R.mt.levels > 0;
    const SampleCfg cfg{F_LINEAR, mips ? MIP_LINEAR : MIP_NONE, R.mt.levels, 0.f, 0.5f};
    float D = 0.f, E = 0.f, sq = 0.f;
    float4 rec = make_float4(__int_as_float(1), __int_as_float(0), 0.f, 0.f);
    if (valid) {
      const Column col = lindepth_column(V, x + 0.5f, y + 0.5f, S, oinv, mips);
      // samples outside the box [0, dims) read zero at every mip level (a
      // level-m sample is inside only where 0 <= p < dims_m 2^m <= dims), so
      // they add nothing to the blend or the gradient: the column is clipped
      int lo, hi;
      clip_ray(col.p0, col.dp, R.mt.dims[0], S, 0.f, lo, hi);
      float* __restrict__ cs = R.csave + (size_t)local * S * 256 + threadIdx.x;
      for (int i = lo; i <= hi; ++i) {
        const float fi = (float)i;
        const float lod = mips ? column_lod(col, fi) : 0.f;
        const float sigma = R.step * batch_sample(R, cfg, fma3(fi, col.dp, col.p0), lod);
        D += sigma;
        const float u = __expf(-D);
        cs[(size_t)i * 256] = u + E;
        E = fmaf(u, sigma, E);
      }
      const size_t pix = ((size_t)vi * R.out_h + y) * R.out_w + x;
      reinterpret_cast<float4*>(R.images)[pix] = make_float4(R.light[0] * E, R.light[1] * E, R.light[2] * E, D);
      const float* tg = R.target + pix * 3;
      const float e0 = R.light[0] * E - tg[0], e1 = R.light[1] * E - tg[1], e2 = R.light[2] * E - tg[2];
      sq = e0 * e0 + e1 * e1 + e2 * e2;
      const float G = R.loss_scale * 2.f * (R.light[0] * e0 + R.light[1] * e1 + R.light[2] * e2);
      if (G != 0.f && lo <= hi)
        rec = make_float4(__int_as_float(lo), __int_as_float(hi), R.step * G, -(R.step * G) * E);
    }
    reinterpret_cast<float4*>(R.records)[(size_t)local * 256 + threadIdx.x] = rec;
    // per-item loss partial: a fixed-order tree over the tile's 256 pixels
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) R.partials[local] = red[0];
    __syncthreads();  // red is rewritten by the next item
  }
}

// ---- backward
// Per item the tile's pixels (16x16 bricks: the tile in rows; 8x8 bricks:
// each wave a quadrant), in segments of at most R.seg samples, each as long
// as its bricks fit the LDS (halved until they do, down to one sample; in the
// 8-pixel mode until every wave's brick fits, so all waves take the same
// segments). Each sample's gradient alpha c_k + beta is scattered through
// its level taps.

__global__ __launch_bounds__(256) void k_batch_bwd(const BatchRec* __restrict__ recs, int nrec, int total_items) {
  __shared__ double lds[kBatchCap];
  __shared__ int span[2];
  __shared__ float corner[16][6];
  const int wave = uniform_i((int)threadIdx.x >> 6), lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < kBatchCap; i += 256) lds[i] = 0.0;
  for (int item = blockIdx.x; item < total_items; item += gridDim.x) {
    const BatchRec& R = recs[rec_of(recs, nrec, item, -1)];
    const int local = item - R.item0;
    const int per_view = R.tiles_x * R.tiles_y;
    const int vi = local / per_view, t = local % per_view;
    const int tx0 = (t % R.tiles_x) * 16, ty0 = (t / R.tiles_x) * 16;
    const SampleView& V = R.view[vi];
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
    const int slot = (y - ty0) * 16 + (x - tx0);  // the forward's thread of this pixel
    const float4 rec = reinterpret_cast<const float4*>(R.records)[(size_t)local * 256 + slot];
    const int lo = __float_as_int(rec.x), hi = __float_as_int(rec.y);
    const float alpha = rec.z, beta = rec.w;
    const int S = R.out_d;
    const float3 oinv = make_float3(1.f / (float)R.out_w, 1.f / (float)R.out_h, 1.f / (float)S);
    const bool mips = R.mt.levels > 0;
    const SampleCfg cfg{F_LINEAR, mips ? MIP_LINEAR : MIP_NONE, R.mt.levels, 0.f, 0.5f};
    Column col{};
    if (lo <= hi) col = lindepth_column(V, x + 0.5f, y + 0.5f, S, oinv, mips);
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
    if (lo <= hi) {
      atomicMin(&span[0], lo);
      atomicMax(&span[1], hi);
    }
    __syncthreads();
    const int s_lo = span[0], s_hi = span[1];
    const float* __restrict__ cs = R.csave + (size_t)local * S * 256 + slot;
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
      const int i0 = max(ia, lo), i1 = min(ib, hi);
      if (i0 <= i1) {
        // c_k of the next sample is loaded while this one is scattered
        float cnext = cs[(size_t)i0 * 256];
        for (int i = i0; i <= i1; ++i) {
          const float c = cnext;
          if (i < i1) cnext = cs[(size_t)(i + 1) * 256];
          const float fi = (float)i;
          const float3 p = fma3(fi, col.dp, col.p0);
          const float lod = mips ? column_lod(col, fi) : 0.f;
          const float g = fmaf(alpha, c, beta);
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
                            m == 0 ? br.n0 : br.n1, R.gpyr + R.mt.offset[m], R.mt.dims[m], p.x * sc - 0.5f,
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
        const float rx0 = 1.f / (float)br.n0[0], rxy0 = 1.f / (float)(br.n0[0] * br.n0[1]);
        const float rx1 = 1.f / (float)br.n1[0], rxy1 = 1.f / (float)(br.n1[0] * br.n1[1]);
        for (int idx = tid; idx < br.size; idx += nthr) {
          const double val = lds[cap0 + idx];
          if (val != 0.0) {
            const bool l0 = idx < br.size0;
            const int j = l0 ? idx : idx - br.size0;
            const int nx = l0 ? br.n0[0] : br.n1[0], nxy = nx * (l0 ? br.n0[1] : br.n1[1]);
            const int lz = div_small(j, nxy, l0 ? rxy0 : rxy1), rz = j - lz * nxy;
            const int ly = div_small(rz, nx, l0 ? rx0 : rx1), lx = rz - ly * nx;
            if (l0)
              atomic_addf(R.gpyr + pad_index(br.lo0[0] + lx, br.lo0[1] + ly, br.lo0[2] + lz, R.mt.dims[0]),
                          (float)val);
            else
              atomic_addf(R.gpyr + R.mt.offset[1] +
                              pad_index(br.lo1[0] + lx, br.lo1[1] + ly, br.lo1[2] + lz, R.mt.dims[1]),
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
// The level-0 gradient with the mip adjoint (g0 + sum_m 8^-m g_m(ancestor))
// and, in each reconstruction's first item, its loss from the render
// partials.
__global__ __launch_bounds__(256) void k_batch_finish(const BatchRec* __restrict__ recs, int nrec, int total_items) {
  __shared__ float red[256];
  for (int item = blockIdx.x; item < total_items; item += gridDim.x) {
    const BatchRec& R = recs[rec_of(recs, nrec, item, kFinishPass)];
    const int local = item - R.pass_item0[kFinishPass];
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
    const Dims3 d = R.mt.dims[0];
    const size_t f = (size_t)local * 256 + threadIdx.x;
    if (f >= (size_t)d.x * d.y * d.z) continue;
    const int x = (int)(f % (size_t)d.x), y = (int)((f / (size_t)d.x) % (size_t)d.y), z = (int)(f / ((size_t)d.x * d.y));
    float g = 0.f, s = 1.f;
    for (int m = 1; m <= R.mt.levels; ++m) {
      s *= 0.125f;
      const Dims3 o = R.mt.dims[m];
      const int px = x >> m, py = y >> m, pz = z >> m;
      if (px < o.x && py < o.y && pz < o.z)
        g = fmaf(s, R.gpyr[R.mt.offset[m] + pad_index(px + 1, py + 1, pz + 1, o)], g);
    }
    R.grad[f] = R.gpyr[pad_index(x + 1, y + 1, z + 1, d)] + g;
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

static int brick_tile(int brick) { return brick == 0 ? 16 : 8; }
static int brick_seg(int brick) { return brick == 1 ? 4 : 8; }

// workspace layout: the record table, then per reconstruction its bordered
// density pyramid, the gradients of the same cells, the corner-quad copy of
// each level, the render partials, the
// pixel records and the sample values c_k (items x out_d x 256 floats: the
// largest part, 4 bytes per frustum sample)
struct BatchLayout {
  size_t table, total;
  std::vector<size_t> pyr, gpyr, partials, records, csave;
  std::vector<std::array<size_t, kBatchLevels>> quad;
};

// items (of 256 threads) of pass p for a reconstruction with pyramid mt
static int pass_items(const MipTable& mt, int p) {
  if (p < kBatchLevels) return p <= mt.levels ? (int)((padded_cells(mt.dims[p]) + 255) / 256) : 0;
  if (p < 2 * kBatchLevels) {
    const int m = p - kBatchLevels;
    return m <= mt.levels ? (int)((quad_entries(make_quad_dims(mt.dims[m])) + 255) / 256) : 0;
  }
  return (int)((mt.base_stride + 255) / 256);
}

static int batch_layout(int32_t n, const gft_batch_recon* rs, int32_t num_mipmaps, BatchLayout& L) {
  if (n < 1) return fail(GFT_ERR_INVALID_ARGUMENT, "the batch needs at least one reconstruction");
  if (!rs) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  if (num_mipmaps < 0 || num_mipmaps > kBatchMaxLevels)
    return fail(GFT_ERR_INVALID_ARGUMENT, "num_mipmaps must be in [0, %d]", kBatchMaxLevels);
  size_t off = align256b(sizeof(BatchRec) * (size_t)n);
  L.table = off;
  L.pyr.resize(n);
  L.gpyr.resize(n);
  L.records.resize(n);
  L.csave.resize(n);
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
    off += align256b(mt.atlas_stride * sizeof(float));
    L.gpyr[r] = off;
    off += align256b(mt.atlas_stride * sizeof(float));
    for (int m = 0; m <= mt.levels; ++m) {
      L.quad[r][m] = off;
      off += align256b(quad_entries(make_quad_dims(mt.dims[m])) * sizeof(float4));
    }
    const size_t items = (size_t)c.num_views * ((c.output_shape[2] + 15) / 16) * ((c.output_shape[1] + 15) / 16);
    L.partials[r] = off;
    off += align256b(items * sizeof(float));
    L.records[r] = off;
    off += align256b(items * 256 * 4 * sizeof(float));
    L.csave[r] = off;
    off += align256b(items * (size_t)c.output_shape[0] * 256 * sizeof(float));
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
    render += c.num_views * ((c.output_shape[2] + 15) / 16) * ((c.output_shape[1] + 15) / 16);
    const MipTable mt = batch_mip_table(Dims3{c.grid_dhw[2], c.grid_dhw[1], c.grid_dhw[0]}, num_mipmaps);
    for (int p = 0; p < kBatchPasses; ++p) pass[p] += pass_items(mt, p);
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
    R.tiles_x = (R.out_w + 15) / 16;
    R.tiles_y = (R.out_h + 15) / 16;
    R.tile = brick_tile(c.brick);
    R.seg = brick_seg(c.brick);
    R.item0 = item;
    item += R.num_views * R.tiles_x * R.tiles_y;
    for (int p = 0; p < kBatchPasses; ++p) {
      R.pass_item0[p] = pass_item[p];
      pass_item[p] += pass_items(R.mt, p);
    }
    for (int m = 0; m <= R.mt.levels; ++m) {
      R.quad[m] = (float4*)(ws + L.quad[r][m]);
      R.qd[m] = make_quad_dims(R.mt.dims[m]);
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
  int render = 0, pass[kBatchPasses];
  batch_items(n, recons, num_mipmaps, render, pass);
  {
    KernelTimer t(s, "batch_levels");
    for (int m = 0; m <= kBatchMaxLevels; ++m)
      if (pass[m])
        hipLaunchKernelGGL(k_batch_level, dim3(persistent_grid(pass[m])), dim3(256), 0, s, recs, n, m, pass[m]);
    for (int m = 0; m <= kBatchMaxLevels; ++m)
      if (pass[kBatchLevels + m])
        hipLaunchKernelGGL(k_batch_quad, dim3(persistent_grid(pass[kBatchLevels + m])), dim3(256), 0, s, recs, n, m,
                           pass[kBatchLevels + m]);
  }
  if ((rc = check_hip(hipGetLastError(), "batched mip levels"))) return rc;
  {
    KernelTimer t(s, "batch_fwd");
    hipLaunchKernelGGL(k_batch_fwd, dim3(persistent_grid(render)), dim3(256), 0, s, recs, n, render);
  }
  if ((rc = check_hip(hipGetLastError(), "batched render forward"))) return rc;
  {
    KernelTimer t(s, "batch_bwd");
    hipLaunchKernelGGL(k_batch_bwd, dim3(persistent_grid(render)), dim3(256), 0, s, recs, n, render);
  }
  if ((rc = check_hip(hipGetLastError(), "batched render backward"))) return rc;
  {
    KernelTimer t(s, "batch_finish");
    hipLaunchKernelGGL(k_batch_finish, dim3(persistent_grid(pass[kFinishPass])), dim3(256), 0, s, recs, n,
                       pass[kFinishPass]);
  }
  return check_hip(hipGetLastError(), "batched finish");
}

}  // extern "C"
