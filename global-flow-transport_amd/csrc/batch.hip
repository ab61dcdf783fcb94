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
//   k_batch_level(m)   m = 0: clear the density gradients; m >= 1: build mip
//                      level m from level m-1 (kGenerateMip3D's pairwise
//                      means) and clear its gradient level;
//   k_batch_render     forward and backward of every ray in one pass: items
//                      are (reconstruction, view, 16x16-pixel tile); a ray
//                      marches its samples, writes its pixel and loss partial,
//                      forms its loss gradient and walks back from far to
//                      near recomputing the transmittance, scattering the
//                      sample gradients through the level taps; level-0 taps
//                      go to an FP64 LDS brick per tile (or 8x8 sub-tile) and
//                      K-sample segment, flushed with float atomics; coarser
//                      levels take float atomics directly;
//   k_batch_finish     the mip adjoint (every level-0 voxel adds 1/8^m of its
//                      level-m ancestor's gradient) and, in the first item of
//                      each reconstruction, its loss: the per-tile partials
//                      summed in a fixed order.
// The brick of the backward (tile size x segment length) is chosen per
// reconstruction: 16x16-pixel tiles with 8-sample segments (about 8^3 cells
// at two pixels per cell), 8x8 tiles with 4-sample segments (about 4^3), or
// 8x8 tiles with 8 (about 4x4x8).
#include <algorithm>
#include <cstring>
#include <vector>

#include "gft_host.hpp"
#include "sampler_v1.hpp"

namespace gft {

constexpr int kBatchMaxViews = 8;
constexpr int kBatchMaxLevels = 3;
constexpr int kBatchCap = 6144;  // FP64 brick cells per workgroup (48 KiB)

struct BatchRec {
  MipTable mt;            // density pyramid: level 0 = dens, levels 1..L in atlas
  const float* dens;
  float* atlas;
  float* grad;            // level-0 gradient, overwritten
  float* gatlas;          // gradients of levels 1..L (same layout as atlas)
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

// ------------------------------------------------------------ level passes
// m = 0: zero the level-0 gradient; m >= 1: level m of the pyramid (pairwise
// means in z, then y, then x, as kGenerateMip3D) and zero its gradient.
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
      R.grad[f] = 0.f;
      continue;
    }
    const Dims3 s = R.mt.dims[m - 1];
    const float* in = m == 1 ? R.dens : R.atlas + R.mt.offset[m - 1];
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
    R.atlas[R.mt.offset[m] + f] = (ya[0] + ya[1]) * 0.5f;
    R.gatlas[R.mt.offset[m] + f] = 0.f;
  }
}

// ---------------------------------------------------------------- render
__device__ __forceinline__ float batch_sample(const BatchRec& R, const SampleCfg& cfg, float3 p, float lod) {
  int lv[2];
  float w[2];
  const int n = level_taps(cfg, lod, lv, w);
  float v = 0.f;
  for (int k = 0; k < n; ++k) {
    const float sc = 1.f / (float)(1 << lv[k]);
    const float* buf = lv[k] == 0 ? R.dens : R.atlas + R.mt.offset[lv[k]];
    const float a = read_linear<1, B_BORDER>(buf, R.mt.dims[lv[k]], p.x * sc - 0.5f, p.y * sc - 0.5f,
                                             p.z * sc - 0.5f).v[0];
    v = k == 0 ? a : lerpf(v, a, w[1]);
  }
  return v;
}

struct BatchBrick {
  int lo[3], n[3];
  int size;
};

// bounding brick (level-0 corner indices) of the samples ia..ib of the rays
// of pixels [x0, x1] x [y0, y1]: the index -> object map is affine along a
// ray (uniform view depth) and across pixels at one depth, so the four corner
// rays at the two segment ends bound every sample of the tile
__device__ __forceinline__ BatchBrick batch_brick(const SampleView& v, float3 oinv, int x0, int x1, int y0, int y1,
                                                  int ia, int ib, Dims3 d) {
  float mn[3] = {3.4e38f, 3.4e38f, 3.4e38f}, mx[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
  const int px[2] = {x0, x1}, py[2] = {y0, y1}, pz[2] = {ia, ib};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float3 p = map_pos(v, CM_LINDEPTH, px[c & 1] + 0.5f, py[(c >> 1) & 1] + 0.5f, pz[c >> 2] + 0.5f, oinv,
                             make_float3(0.f, 0.f, 0.f));
    const float q[3] = {p.x - 0.5f, p.y - 0.5f, p.z - 0.5f};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      mn[a] = fminf(mn[a], q[a]);
      mx[a] = fmaxf(mx[a], q[a]);
    }
  }
  const int dim[3] = {d.x, d.y, d.z};
  BatchBrick b;
  b.size = 1;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int lo = clampi((int)floorf(mn[a] - 1e-3f), 0, dim[a] - 1);
    const int hi = clampi((int)floorf(mx[a] + 1e-3f) + 1, 0, dim[a] - 1);
    b.lo[a] = lo;
    b.n[a] = hi - lo + 1;
    b.size *= b.n[a];
  }
  return b;
}

__global__ __launch_bounds__(256) void k_batch_render(const BatchRec* __restrict__ recs, int nrec, int total_items) {
  __shared__ double lds[kBatchCap];
  __shared__ float red[256];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < kBatchCap; i += 256) lds[i] = 0.0;
  for (int item = blockIdx.x; item < total_items; item += gridDim.x) {
    const BatchRec& R = recs[rec_of(recs, nrec, item, -1)];
    const int local = item - R.item0;
    const int per_view = R.tiles_x * R.tiles_y;
    const int vi = local / per_view, t = local % per_view;
    const int tx0 = (t % R.tiles_x) * 16, ty0 = (t / R.tiles_x) * 16;
    const SampleView& V = R.view[vi];
    // this thread's pixel and the pixels of its brick
    int x, y, bx0, by0, bw, cap, cap0;
    if (R.tile == 16) {
      x = tx0 + (threadIdx.x & 15);
      y = ty0 + (threadIdx.x >> 4);
      bx0 = tx0;
      by0 = ty0;
      bw = 16;
      cap = kBatchCap;
      cap0 = 0;
    } else {
      bx0 = tx0 + (wave & 1) * 8;
      by0 = ty0 + (wave >> 1) * 8;
      x = bx0 + (lane & 7);
      y = by0 + (lane >> 3);
      bw = 8;
      cap = kBatchCap / 4;
      cap0 = wave * cap;
    }
    const bool valid = x < R.out_w && y < R.out_h;
    const int S = R.out_d;
    const float3 oinv = make_float3(1.f / (float)R.out_w, 1.f / (float)R.out_h, 1.f / (float)S);
    const SampleCfg cfg{F_LINEAR, R.mt.levels > 0 ? MIP_LINEAR : MIP_NONE, R.mt.levels, 0.f, 0.5f};
    // ---- forward: BEER_LAMBERT blend of the sampled frustum column
    float D = 0.f, E = 0.f;
    if (valid) {
      for (int i = 0; i < S; ++i) {
        float lod;
        const float3 p = sample_pos(V, CM_LINDEPTH, R.mt.levels > 0, x + 0.5f, y + 0.5f, i + 0.5f, oinv,
                                    make_float3(0.f, 0.f, 0.f), lod);
        const float sigma = R.step * batch_sample(R, cfg, p, lod);
        D += sigma;
        E = fmaf(__expf(-D), sigma, E);
      }
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
    // per-item loss partial: a fixed-order tree over the tile's 256 pixel
    // slots, the same sum whatever the brick's thread-to-pixel mapping
    red[(y - ty0) * 16 + (x - tx0)] = sq;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) R.partials[local] = red[0];
    // ---- backward: walk far to near, recomputing sigma and the transmittance
    const bool walk = valid && G != 0.f;
    float Racc = 0.f;
    const int bx1 = min(bx0 + bw - 1, R.out_w - 1), by1 = min(by0 + bw - 1, R.out_h - 1);
    const int K = R.seg;
    for (int sgi = (S + K - 1) / K - 1; sgi >= 0; --sgi) {
      const int ia = sgi * K, ib = min(ia + K, S) - 1;
      const BatchBrick br = batch_brick(V, oinv, bx0, bx1, by0, by1, ia, ib, R.mt.dims[0]);
      const bool use_lds = br.size <= cap;
      if (walk) {
        for (int i = ib; i >= ia; --i) {
          float lod;
          const float3 p = sample_pos(V, CM_LINDEPTH, R.mt.levels > 0, x + 0.5f, y + 0.5f, i + 0.5f, oinv,
                                      make_float3(0.f, 0.f, 0.f), lod);
          const float sigma = R.step * batch_sample(R, cfg, p, lod);
          const float u = __expf(-D);
          Racc = fmaf(u, sigma, Racc);
          const float g = R.step * G * (u - Racc);
          D -= sigma;
          int lv[2];
          float lw[2];
          const int nl = level_taps(cfg, lod, lv, lw);
          for (int k = 0; k < nl; ++k) {
            const int m = lv[k];
            const float sc = 1.f / (float)(1 << m);
            const Dims3 dm = R.mt.dims[m];
            const Taps tp = make_taps<B_BORDER>(p.x * sc - 0.5f, p.y * sc - 0.5f, p.z * sc - 0.5f, dm);
            if (!tp.inside) continue;
            const float gw = g * lw[k];
            const float wx[2] = {1.f - tp.wx, tp.wx}, wy[2] = {1.f - tp.wy, tp.wy}, wz[2] = {1.f - tp.wz, tp.wz};
            for (int kz = 0; kz < 2; ++kz)
              for (int ky = 0; ky < 2; ++ky)
                for (int kx = 0; kx < 2; ++kx) {
                  if (!(tp.vx[kx] & tp.vy[ky] & tp.vz[kz])) continue;
                  const float w = gw * wx[kx] * wy[ky] * wz[kz];
                  if (w == 0.f) continue;
                  const int cx = tp.ix[kx], cy = tp.iy[ky], cz = tp.iz[kz];
                  if (m == 0) {
                    const int lx = cx - br.lo[0], ly = cy - br.lo[1], lz = cz - br.lo[2];
                    if (use_lds & ((unsigned)lx < (unsigned)br.n[0]) & ((unsigned)ly < (unsigned)br.n[1]) &
                        ((unsigned)lz < (unsigned)br.n[2]))
                      atomicAdd(&lds[cap0 + (lz * br.n[1] + ly) * br.n[0] + lx], (double)w);
                    else
                      atomic_addf(R.grad + flat3(cx, cy, cz, dm), w);
                  } else {
                    atomic_addf(R.gatlas + R.mt.offset[m] + flat3(cx, cy, cz, dm), w);
                  }
                }
          }
        }
      }
      // (8-pixel tiles: each wave owns its brick, so use_lds may differ between waves and the
      // barriers stay outside the branch)
      __syncthreads();
      if (use_lds) {
        // flush the brick (all 256 threads for a 16-pixel tile, the wave's 64 lanes for an 8-pixel one)
        const int nthr = R.tile == 16 ? 256 : 64, tid = R.tile == 16 ? (int)threadIdx.x : lane;
        for (int idx = tid; idx < br.size; idx += nthr) {
          const double val = lds[cap0 + idx];
          if (val != 0.0) {
            const int lx = idx % br.n[0], ly = (idx / br.n[0]) % br.n[1], lz = idx / (br.n[0] * br.n[1]);
            atomic_addf(R.grad + flat3(br.lo[0] + lx, br.lo[1] + ly, br.lo[2] + lz, R.mt.dims[0]), (float)val);
            lds[cap0 + idx] = 0.0;
          }
        }
      }
      __syncthreads();
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
      if (px < o.x && py < o.y && pz < o.z) g = fmaf(s, R.gatlas[R.mt.offset[m] + flat3(px, py, pz, o)], g);
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
  size_t off = 0;
  for (int m = 1; m <= levels; ++m) {
    mt.dims[m] = Dims3{d.x >> m, d.y >> m, d.z >> m};
    mt.offset[m] = off;
    off += (size_t)mt.dims[m].x * mt.dims[m].y * mt.dims[m].z;
  }
  mt.atlas_stride = (off + 63) & ~size_t(63);
  return mt;
}

static int brick_tile(int brick) { return brick == 0 ? 16 : 8; }
static int brick_seg(int brick) { return brick == 1 ? 4 : 8; }

// workspace layout: the record table, then per reconstruction its atlas,
// gradient atlas and render partials
struct BatchLayout {
  size_t table, total;
  std::vector<size_t> atlas, gatlas, partials;
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
    off += align256b(mt.atlas_stride * sizeof(float));
    const size_t items = (size_t)c.num_views * ((c.output_shape[2] + 15) / 16) * ((c.output_shape[1] + 15) / 16);
    L.partials[r] = off;
    off += align256b(items * sizeof(float));
  }
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
    R.atlas = (float*)(ws + L.atlas[r]);
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
    hipLaunchKernelGGL(k_batch_render, dim3(persistent_grid(render)), dim3(256), 0, s, recs, n, render);
  }
  if ((rc = check_hip(hipGetLastError(), "batched render"))) return rc;
  {
    KernelTimer t(s, "batch_finish");
    hipLaunchKernelGGL(k_batch_finish, dim3(persistent_grid(level[0])), dim3(256), 0, s, recs, n, level[0]);
  }
  return check_hip(hipGetLastError(), "batched finish");
}

}  // extern "C"
