// C ABI of the resampling ops: SampleGridTransform (+ gradient and the
// view-summed gather gradient of the renderer), SampleGridLut (+ gradient),
// LodTransform and the mip pyramid. Validation follows the op kernels of
// phitest/render/cuda/src/resample_grid.cc:268-1105.
#include <algorithm>

#include "gft_host.hpp"

namespace gft {


hipError_t fill_launch(hipStream_t s, float* p, size_t n, float v);
hipError_t mip_down_launch(int C, hipStream_t st, const float* src, size_t src_stride, Dims3 s, float* dst,
                           size_t dst_stride, Dims3 o, int batches, float scale);
hipError_t mip_collapse_launch(int C, hipStream_t st, const float* gp, size_t gp_stride, Dims3 o, float* gc,
                               size_t gc_stride, Dims3 s, int batches);
hipError_t sample_transform_launch(int C, int bm, bool sum, hipStream_t st, const float* in, const float* atlas,
                                   const MipTable& mt, const SampleCfg& s, Dims3 dout, const SampleLaunch& L,
                                   float* out);
hipError_t sample_transform_scatter_ref_launch(int C, int bm, hipStream_t st, const float* og, Dims3 din, Dims3 dout,
                                               const SampleLaunch& L, float* grad, unsigned* cnt);
hipError_t sample_transform_adjoint_launch(int C, int bm, hipStream_t st, const float* in, const float* atlas,
                                           const MipTable& mt, const SampleCfg& s, Dims3 dout, const SampleLaunch& L,
                                           const float* og, float* grad, float* gatlas);
hipError_t shadow_fwd_launch(int bm, int fm, int bl, hipStream_t st, const float* dens, const float* atlas,
                             const MipTable& mt, const SampleCfg& cfg, const SampleView& v, Dims3 dout, float step,
                             const float* mask, float* saved, float* trans);
hipError_t shadow_bwd_launch(int bm, int bl, hipStream_t st, const float* saved, const float* gt, const float* mask,
                             Dims3 dout, float* gs);
hipError_t sample_transform_gather_launch(int C, bool ref, int fm, hipStream_t st, const float* og, Dims3 din,
                                          Dims3 dout, const SampleLaunch& L, const SampleLaunch& Linv, int accumulate,
                                          float* grad);
hipError_t normalize_count_launch(int C, hipStream_t st, float* grad, const unsigned* cnt, size_t nvox);
hipError_t sample_lut_launch(int C, int bm, hipStream_t st, const float* in, const float* atlas, const MipTable& mt,
                             const SampleCfg& s, const LutLaunch& L, int pairs, const float* lut, float* out);
hipError_t sample_lut_bwd_launch(int C, int bm, bool exact, hipStream_t st, const float* in_b, int fm, float offset,
                                 const LutLaunch& L, int views, const float* lut, const float* og, Dims3 din,
                                 float* grad_b, unsigned* cnt, float* lut_grad);
hipError_t lod_transform_launch(hipStream_t st, Dims3 din, Dims3 dout, const SampleLaunch& L, float* out);

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

static Dims3 dims_of(const int64_t* shape5) { return Dims3{(int)shape5[3], (int)shape5[2], (int)shape5[1]}; }
static Dims3 dims_dhw(const int32_t* dhw) { return Dims3{dhw[2], dhw[1], dhw[0]}; }
static size_t nvox(Dims3 d) { return (size_t)d.x * d.y * d.z; }

// setupMipAtlas (resample_grid.cc:201-220): levels whose dims would reach 0
// are dropped.
static int usable_levels(Dims3 d, int levels) {
  levels = std::min(levels, kMaxMips);
  for (int m = 1; m <= levels; ++m)
    if ((d.x >> m) <= 0 || (d.y >> m) <= 0 || (d.z >> m) <= 0) return m - 1;
  return levels;
}

static MipTable make_mip_table(Dims3 d, int C, int levels, size_t base_stride) {
  MipTable mt{};
  mt.levels = levels;
  mt.C = C;
  mt.dims[0] = d;
  mt.base_stride = base_stride;
  size_t off = 0;
  for (int m = 1; m <= levels; ++m) {
    mt.dims[m] = Dims3{d.x >> m, d.y >> m, d.z >> m};
    mt.offset[m] = off;
    off += nvox(mt.dims[m]) * C;
  }
  mt.atlas_stride = (off + 63) & ~size_t(63);  // 256-byte aligned blocks
  return mt;
}

static size_t atlas_bytes(Dims3 d, int C, int levels, int batches) {
  if (levels <= 0) return 0;
  MipTable mt = make_mip_table(d, C, levels, 0);
  return align256(mt.atlas_stride * sizeof(float) * (size_t)batches);
}

// Mip levels 1.. of `base` into the atlas; base_scale multiplies the base
// grid's values as they are read (the pyramid of base * base_scale).
static hipError_t build_atlas(hipStream_t st, const float* base, const MipTable& mt, int batches, float* atlas,
                              float base_scale = 1.f) {
  for (int m = 1; m <= mt.levels; ++m) {
    const float* src = m == 1 ? base : atlas + mt.offset[m - 1];
    size_t src_stride = m == 1 ? mt.base_stride : mt.atlas_stride;
    hipError_t e = mip_down_launch(mt.C, st, src, src_stride, mt.dims[m - 1], atlas + mt.offset[m], mt.atlas_stride,
                                   mt.dims[m], batches, m == 1 ? base_scale : 1.f);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

static int check_grid(const int64_t* shape, const char* what) {
  for (int i = 0; i < 5; ++i)
    if (shape[i] < 1)
      return fail(GFT_ERR_INVALID_ARGUMENT, "Empty %s (NDHWC): dimension %d is %lld", what, i, (long long)shape[i]);
  if (shape[4] > 4) return fail(GFT_ERR_INVALID_ARGUMENT, "Invalid %s shape (NDHWC): channel %lld > 4", what,
                                (long long)shape[4]);
  if (shape[4] == 3) return fail(GFT_ERR_UNIMPLEMENTED, "Only 1,2 and 4 Channel supported.");
  return GFT_OK;
}

static int check_out_shape(const int32_t* o) {
  for (int i = 0; i < 3; ++i)
    if (o[i] < 1) return fail(GFT_ERR_INVALID_ARGUMENT, "All shape dimensions must be positive, dimension %d is %d", i,
                              o[i]);
  return GFT_OK;
}

static int check_mips(int32_t mipmapping, int32_t num_mipmaps) {
  if (mipmapping < 0 || mipmapping > 2) return fail(GFT_ERR_INVALID_ARGUMENT, "invalid mipmapping mode");
  if (mipmapping != GFT_MIP_NONE && num_mipmaps <= 0)
    return fail(GFT_ERR_INVALID_ARGUMENT, "when using mipmaps num_mipmaps must be larger than 0.");
  return GFT_OK;
}

static int check_cameras(int64_t batch, int32_t num_M, int32_t num_cameras, int32_t separate_camera_batch) {
  if (num_M != batch) return fail(GFT_ERR_INVALID_ARGUMENT, "model matrix batch size mismatch");
  if (num_cameras < 1) return fail(GFT_ERR_INVALID_ARGUMENT, "camera batch size mismatch");
  if (!separate_camera_batch && num_cameras != batch)
    return fail(GFT_ERR_INVALID_ARGUMENT, "camera batch must match data batch when not using global sampling.");
  return GFT_OK;
}

static int check_transform_attrs(int32_t interpolation, int32_t boundary, int32_t coordinate_mode,
                                 float cell_center_offset) {
  if (interpolation != GFT_NEAREST && interpolation != GFT_LINEAR)
    return fail(GFT_ERR_INVALID_ARGUMENT, "interpolation must be NEAREST or LINEAR");
  if (boundary != GFT_BORDER && boundary != GFT_CLAMP)
    return fail(GFT_ERR_INVALID_ARGUMENT, "boundary must be BORDER or CLAMP");
  if (coordinate_mode < 0 || coordinate_mode > 3) return fail(GFT_ERR_INVALID_ARGUMENT, "Invalid coordinate_mode");
  if (cell_center_offset != 0.5f)
    return fail(GFT_ERR_INVALID_ARGUMENT, "Invalid cell center offset for given coordinate mode.");
  return GFT_OK;
}

// Camera records of `count` views. Forward coordinate modes need P^-1 and
// (V M)^-1, reverse modes V M and P (transformations.hpp:44-64); the inverses
// are formed in double precision.
static int build_sample_views(const std::vector<ViewPair>& pairs, size_t first, size_t count, const float* M,
                              int m_index_override, const float* V, const float* P, const float* F, int cm,
                              SampleLaunch& L) {
  L.num_views = (int)count;
  L.cm = cm;
  L.accumulate = 0;
  for (size_t i = 0; i < count; ++i) {
    const ViewPair& vp = pairs[first + i];
    int mi = m_index_override >= 0 ? m_index_override : vp.batch;
    Mat4d m = mat_from_colmajor(M ? M + 16 * mi : nullptr);
    Mat4d v = mat_from_colmajor(V + 16 * vp.cam);
    Mat4d p = mat_from_colmajor(P + 16 * vp.cam);
    Mat4d mv = mat_mul(v, m);
    SampleView& sv = L.view[i];
    if (cm == CM_LINDEPTH || cm == CM_TRANSFORM) {
      Mat4d mvinv, pinv;
      if (!mat_inverse(mv, mvinv)) return fail(GFT_ERR_INVALID_ARGUMENT, "model-view matrix is singular");
      if (!mat_inverse(p, pinv)) return fail(GFT_ERR_INVALID_ARGUMENT, "projection matrix is singular");
      mat_to_float(pinv, sv.A);
      mat_to_float(mvinv, sv.B);
    } else {
      mat_to_float(mv, sv.A);
      mat_to_float(p, sv.B);
    }
    sv.p22 = P[16 * vp.cam + 4 * 2 + 2];
    sv.p32 = P[16 * vp.cam + 4 * 3 + 2];
    sv.nearp = F[6 * vp.cam + 0];
    sv.farp = F[6 * vp.cam + 1];
    sv.in_batch = vp.batch;
    sv.out_slot = vp.slot;
  }
  return GFT_OK;
}

}  // namespace gft

using namespace gft;

extern "C" {

// ------------------------------------------------------------------ mips
size_t gft_mip_atlas_bytes(const int64_t shape[5], int32_t num_mipmaps) {
  Dims3 d = dims_of(shape);
  return atlas_bytes(d, (int)shape[4], usable_levels(d, num_mipmaps), (int)shape[0]);
}

int gft_build_mips(void* stream, const float* input, const int64_t shape[5], int32_t num_mipmaps, float* atlas,
                   size_t atlas_bytes_given) {
  int rc = check_grid(shape, "input");
  if (rc) return rc;
  if (num_mipmaps < 0) return fail(GFT_ERR_INVALID_ARGUMENT, "num_mipmaps must be non-negative");
  Dims3 d = dims_of(shape);
  int C = (int)shape[4];
  int levels = usable_levels(d, num_mipmaps);
  size_t need = atlas_bytes(d, C, levels, (int)shape[0]);
  if (atlas_bytes_given < need || (need && !atlas))
    return fail(GFT_ERR_WORKSPACE, "atlas of %zu bytes is smaller than the %zu bytes required", atlas_bytes_given, need);
  if (!input) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  MipTable mt = make_mip_table(d, C, levels, nvox(d) * C);
  return check_hip(build_atlas((hipStream_t)stream, input, mt, (int)shape[0], atlas), "mip generation");
}

// ------------------------------------------------------ sample transform
size_t gft_sample_grid_transform_workspace(const int64_t input_shape[5], int32_t mipmapping, int32_t num_mipmaps) {
  if (mipmapping == GFT_MIP_NONE) return 0;
  return gft_mip_atlas_bytes(input_shape, num_mipmaps);
}

int gft_sample_grid_transform_fwd(void* stream, const float* input, const int64_t input_shape[5], const float* M,
                                  int32_t num_M, const float* V, const float* P, const float* frustum,
                                  int32_t num_cameras, const int32_t output_shape[3], int32_t interpolation,
                                  int32_t boundary, int32_t mipmapping, int32_t num_mipmaps, float mip_bias,
                                  int32_t coordinate_mode, float cell_center_offset, int32_t separate_camera_batch,
                                  void* workspace, size_t workspace_bytes, float* output) {
  int rc;
  if ((rc = check_grid(input_shape, "input"))) return rc;
  if ((rc = check_out_shape(output_shape))) return rc;
  if ((rc = check_transform_attrs(interpolation, boundary, coordinate_mode, cell_center_offset))) return rc;
  if ((rc = check_mips(mipmapping, num_mipmaps))) return rc;
  if ((rc = check_cameras(input_shape[0], num_M, num_cameras, separate_camera_batch))) return rc;
  if (!input || !output || !M || !V || !P || !frustum) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  size_t need = gft_sample_grid_transform_workspace(input_shape, mipmapping, num_mipmaps);
  if (workspace_bytes < need || (need && !workspace))
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                need);
  hipStream_t st = (hipStream_t)stream;
  Dims3 din = dims_of(input_shape), dout = dims_dhw(output_shape);
  int C = (int)input_shape[4];
  int levels = mipmapping == GFT_MIP_NONE ? 0 : usable_levels(din, num_mipmaps);
  MipTable mt = make_mip_table(din, C, levels, nvox(din) * C);
  float* atlas = (float*)workspace;
  if (levels && (rc = check_hip(build_atlas(st, input, mt, (int)input_shape[0], atlas), "mip generation"))) return rc;
  SampleCfg s{interpolation, mipmapping, levels, mip_bias, cell_center_offset};
  auto pairs = view_pairs((int)input_shape[0], num_cameras, separate_camera_batch != 0);
  for (size_t first = 0; first < pairs.size(); first += kMaxSampleViews) {
    size_t cnt = std::min(pairs.size() - first, (size_t)kMaxSampleViews);
    SampleLaunch L;
    if ((rc = build_sample_views(pairs, first, cnt, M, -1, V, P, frustum, coordinate_mode, L))) return rc;
    if ((rc = check_hip(sample_transform_launch(C, boundary, false, st, input, atlas, mt, s, dout, L, output),
                        "sample transform launch")))
      return rc;
  }
  return GFT_OK;
}

size_t gft_sample_grid_transform_bwd_workspace(const int64_t input_shape[5], int32_t mipmapping, int32_t num_mipmaps,
                                               int32_t grad_mode) {
  Dims3 d = dims_of(input_shape);
  if (grad_mode == GFT_GRAD_REFERENCE) return align256(nvox(d) * sizeof(unsigned));
  if (mipmapping == GFT_MIP_NONE) return 0;
  return 2 * gft_mip_atlas_bytes(input_shape, num_mipmaps);
}

int gft_sample_grid_transform_bwd(void* stream, const float* input, const int64_t input_shape[5],
                                  const float* output_grad, const float* M, int32_t num_M, const float* V,
                                  const float* P, const float* frustum, int32_t num_cameras,
                                  const int32_t output_shape[3], int32_t interpolation, int32_t boundary,
                                  int32_t mipmapping, int32_t num_mipmaps, float mip_bias, int32_t coordinate_mode,
                                  float cell_center_offset, int32_t separate_camera_batch, int32_t grad_mode,
                                  void* workspace, size_t workspace_bytes, float* input_grad) {
  int rc;
  if ((rc = check_grid(input_shape, "input"))) return rc;
  if ((rc = check_out_shape(output_shape))) return rc;
  if ((rc = check_transform_attrs(interpolation, boundary, coordinate_mode, cell_center_offset))) return rc;
  if ((rc = check_mips(mipmapping, num_mipmaps))) return rc;
  if ((rc = check_cameras(input_shape[0], num_M, num_cameras, separate_camera_batch))) return rc;
  if (grad_mode != GFT_GRAD_REFERENCE && grad_mode != GFT_GRAD_EXACT)
    return fail(GFT_ERR_INVALID_ARGUMENT, "invalid grad_mode");
  // the reference gradient op registers mipmapping {'NONE'} only (resample_grid.cc:64-84)
  if (grad_mode == GFT_GRAD_REFERENCE && mipmapping != GFT_MIP_NONE)
    return fail(GFT_ERR_INVALID_ARGUMENT, "the REFERENCE gradient supports mipmapping NONE only");
  if (!input || !output_grad || !input_grad || !M || !V || !P || !frustum)
    return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  size_t need = gft_sample_grid_transform_bwd_workspace(input_shape, mipmapping, num_mipmaps, grad_mode);
  if (workspace_bytes < need || (need && !workspace))
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                need);
  hipStream_t st = (hipStream_t)stream;
  Dims3 din = dims_of(input_shape), dout = dims_dhw(output_shape);
  int C = (int)input_shape[4];
  const int N = (int)input_shape[0];
  const size_t vox = nvox(din), bstride = vox * C;
  auto pairs = view_pairs(N, num_cameras, separate_camera_batch != 0);
  // Gather form (no atomics, one store per voxel) for BORDER without mips in
  // the forward coordinate modes: per batch, views in launches of up to
  // kMaxSampleViews (REFERENCE needs all of a batch's views in one launch for
  // its counts).
  const bool gather = boundary == GFT_BORDER && mipmapping == GFT_MIP_NONE &&
                      (coordinate_mode == CM_LINDEPTH || coordinate_mode == CM_TRANSFORM) &&
                      (grad_mode == GFT_GRAD_REFERENCE || interpolation == GFT_LINEAR ||
                       interpolation == GFT_NEAREST) &&
                      nvox(dout) < (size_t(1) << 31);
  if (gather) {
    std::vector<int> per_batch(N, 0);
    for (const ViewPair& vp : pairs) per_batch[vp.batch]++;
    bool fits = true;
    for (int b = 0; b < N && grad_mode == GFT_GRAD_REFERENCE; ++b) fits &= per_batch[b] <= kMaxSampleViews;
    if (fits) {
      for (int b = 0; b < N; ++b) {
        std::vector<ViewPair> mine;
        for (const ViewPair& vp : pairs)
          if (vp.batch == b) mine.push_back(vp);
        for (size_t first = 0; first < mine.size(); first += kMaxSampleViews) {
          size_t n = std::min(mine.size() - first, (size_t)kMaxSampleViews);
          SampleLaunch L, Linv;
          if ((rc = build_sample_views(mine, first, n, M, -1, V, P, frustum, coordinate_mode, L))) return rc;
          if ((rc = build_sample_views(mine, first, n, M, -1, V, P, frustum, coordinate_mode ^ 1, Linv))) return rc;
          if ((rc = check_hip(sample_transform_gather_launch(C, grad_mode == GFT_GRAD_REFERENCE, interpolation, st,
                                                             output_grad, din, dout, L, Linv, first > 0,
                                                             input_grad + bstride * b),
                              "gradient gather launch")))
            return rc;
        }
      }
      return GFT_OK;
    }
  }
  if ((rc = check_hip(fill_launch(st, input_grad, bstride * N, 0.f), "zero input_grad"))) return rc;
  if (grad_mode == GFT_GRAD_REFERENCE) {
    // per batch: counted scatter of all its views, then normalisation
    unsigned* cnt = (unsigned*)workspace;
    for (int b = 0; b < N; ++b) {
      if ((rc = check_hip(fill_launch(st, (float*)cnt, vox, 0.f), "zero counter"))) return rc;
      std::vector<ViewPair> mine;
      for (const ViewPair& vp : pairs)
        if (vp.batch == b) mine.push_back(vp);
      for (size_t first = 0; first < mine.size(); first += kMaxSampleViews) {
        size_t n = std::min(mine.size() - first, (size_t)kMaxSampleViews);
        SampleLaunch L;
        if ((rc = build_sample_views(mine, first, n, M, -1, V, P, frustum, coordinate_mode, L))) return rc;
        if ((rc = check_hip(sample_transform_scatter_ref_launch(C, boundary, st, output_grad, din, dout, L,
                                                                input_grad + bstride * b, cnt),
                            "gradient scatter launch")))
          return rc;
      }
      if ((rc = check_hip(normalize_count_launch(C, st, input_grad + bstride * b, cnt, vox), "normalize launch")))
        return rc;
    }
    return GFT_OK;
  }
  // EXACT: adjoint of every tap, mips included
  int levels = mipmapping == GFT_MIP_NONE ? 0 : usable_levels(din, num_mipmaps);
  MipTable mt = make_mip_table(din, C, levels, bstride);
  float* atlas = (float*)workspace;
  float* gatlas = levels ? (float*)((char*)workspace + need / 2) : nullptr;
  if (levels) {
    if ((rc = check_hip(build_atlas(st, input, mt, N, atlas), "mip generation"))) return rc;
    if ((rc = check_hip(fill_launch(st, gatlas, mt.atlas_stride * N, 0.f), "zero mip gradients"))) return rc;
  }
  SampleCfg s{interpolation, mipmapping, levels, mip_bias, cell_center_offset};
  for (size_t first = 0; first < pairs.size(); first += kMaxSampleViews) {
    size_t n = std::min(pairs.size() - first, (size_t)kMaxSampleViews);
    SampleLaunch L;
    if ((rc = build_sample_views(pairs, first, n, M, -1, V, P, frustum, coordinate_mode, L))) return rc;
    if ((rc = check_hip(sample_transform_adjoint_launch(C, boundary, st, input, atlas, mt, s, dout, L, output_grad,
                                                        input_grad, gatlas),
                        "gradient adjoint launch")))
      return rc;
  }
  for (int m = levels; m >= 1; --m) {
    float* child = m == 1 ? input_grad : gatlas + mt.offset[m - 1];
    size_t child_stride = m == 1 ? bstride : mt.atlas_stride;
    if ((rc = check_hip(mip_collapse_launch(C, st, gatlas + mt.offset[m], mt.atlas_stride, mt.dims[m], child,
                                            child_stride, mt.dims[m - 1], N),
                        "mip gradient collapse")))
      return rc;
  }
  return GFT_OK;
}

// Gather gradient of the renderer's sampling (renderer.py:246-272): for each
// data batch i, dx[i] = sum over views j of a resample of dy[i, j] with the
// reverse coordinate mode and camera j, mips built on dy.
size_t gft_sample_grid_transform_gather_grad_workspace(const int64_t output_grad_shape[6], int32_t mipmapping,
                                                       int32_t num_mipmaps) {
  if (mipmapping == GFT_MIP_NONE) return 0;
  const int64_t s5[5] = {output_grad_shape[0] * output_grad_shape[1], output_grad_shape[2], output_grad_shape[3],
                         output_grad_shape[4], output_grad_shape[5]};
  return gft_mip_atlas_bytes(s5, num_mipmaps);
}

int gft_sample_grid_transform_gather_grad(void* stream, const float* output_grad, const int64_t output_grad_shape[6],
                                          const int32_t input_dhw[3], const float* M, int32_t num_M, const float* V,
                                          const float* P, const float* frustum, int32_t num_cameras,
                                          int32_t interpolation, int32_t boundary, int32_t mipmapping,
                                          int32_t num_mipmaps, float mip_bias, int32_t coordinate_mode,
                                          void* workspace, size_t workspace_bytes, float* input_grad) {
  int rc;
  const int64_t N = output_grad_shape[0], Vc = output_grad_shape[1];
  const int64_t s5[5] = {N * Vc, output_grad_shape[2], output_grad_shape[3], output_grad_shape[4],
                         output_grad_shape[5]};
  if ((rc = check_grid(s5, "output_grad"))) return rc;
  if ((rc = check_out_shape(input_dhw))) return rc;
  if ((rc = check_transform_attrs(interpolation, boundary, coordinate_mode, 0.5f))) return rc;
  if ((rc = check_mips(mipmapping, num_mipmaps))) return rc;
  if (num_M != N) return fail(GFT_ERR_INVALID_ARGUMENT, "model matrix batch size mismatch");
  if (num_cameras != Vc) return fail(GFT_ERR_INVALID_ARGUMENT, "output_grads views size does not match cameras.");
  if (!output_grad || !input_grad || !M || !V || !P || !frustum)
    return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  size_t need = gft_sample_grid_transform_gather_grad_workspace(output_grad_shape, mipmapping, num_mipmaps);
  if (workspace_bytes < need || (need && !workspace))
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                need);
  hipStream_t st = (hipStream_t)stream;
  Dims3 dfr = dims_of(s5), dvol = dims_dhw(input_dhw);
  int C = (int)s5[4];
  int levels = mipmapping == GFT_MIP_NONE ? 0 : usable_levels(dfr, num_mipmaps);
  MipTable mt = make_mip_table(dfr, C, levels, nvox(dfr) * C);
  float* atlas = (float*)workspace;
  if (levels && (rc = check_hip(build_atlas(st, output_grad, mt, (int)(N * Vc), atlas), "mip generation"))) return rc;
  SampleCfg s{interpolation, mipmapping, levels, mip_bias, 0.5f};
  const int rev = coordinate_mode ^ 1;  // LINDEPTH <-> LINDEPTH_REVERSE, TRANSFORM <-> TRANSFORM_REVERSE
  const size_t vstride = nvox(dvol) * C;
  for (int64_t i = 0; i < N; ++i) {
    std::vector<ViewPair> pairs;
    for (int j = 0; j < Vc; ++j) pairs.push_back({(int)(i * Vc + j), j, 0});
    for (size_t first = 0; first < pairs.size(); first += kMaxSampleViews) {
      size_t n = std::min(pairs.size() - first, (size_t)kMaxSampleViews);
      SampleLaunch L;
      if ((rc = build_sample_views(pairs, first, n, M, (int)i, V, P, frustum, rev, L))) return rc;
      L.accumulate = first > 0;
      if ((rc = check_hip(sample_transform_launch(C, boundary, true, st, output_grad, atlas, mt, s, dvol, L,
                                                  input_grad + vstride * i),
                          "gather gradient launch")))
        return rc;
    }
  }
  return GFT_OK;
}

// -------------------------------------------------------------------- LuT
static int check_lut(const int64_t* ishape, const int64_t* lshape, int32_t interpolation, int32_t boundary,
                     float cell_center_offset, int32_t separate_camera_batch, int32_t normalized_coords) {
  int rc;
  if ((rc = check_grid(ishape, "input"))) return rc;
  for (int i = 0; i < 4; ++i)
    if (lshape[i] < 1) return fail(GFT_ERR_INVALID_ARGUMENT, "Empty lut: dimension %d is %lld", i, (long long)lshape[i]);
  if (lshape[4] != 4)
    return fail(GFT_ERR_INVALID_ARGUMENT, "Invalid lut shape (NDHWC) with C=4: channel is %lld", (long long)lshape[4]);
  if (interpolation < 0 || interpolation > 3) return fail(GFT_ERR_INVALID_ARGUMENT, "invalid filter mode");
  if (boundary < 0 || boundary > 2) return fail(GFT_ERR_INVALID_ARGUMENT, "invalid boundary mode");
  if (normalized_coords && cell_center_offset != 0.5f)
    return fail(GFT_ERR_INVALID_ARGUMENT, "Invalid cell center offset must be 0.5 when using normalized coordinates.");
  if (!separate_camera_batch && lshape[0] != ishape[0])
    return fail(GFT_ERR_INVALID_ARGUMENT, "camera batch must match data batch when not using global sampling.");
  return GFT_OK;
}

size_t gft_sample_grid_lut_workspace(const int64_t input_shape[5], int32_t mipmapping, int32_t num_mipmaps) {
  if (mipmapping == GFT_MIP_NONE) return 0;
  return gft_mip_atlas_bytes(input_shape, num_mipmaps);
}

int gft_sample_grid_lut_fwd(void* stream, const float* input, const int64_t input_shape[5], const float* lookup,
                            const int64_t lookup_shape[5], int32_t interpolation, int32_t boundary, int32_t mipmapping,
                            int32_t num_mipmaps, float mip_bias, float cell_center_offset,
                            int32_t separate_camera_batch, int32_t relative_coords, int32_t normalized_coords,
                            void* workspace, size_t workspace_bytes, float* output) {
  int rc;
  if ((rc = check_lut(input_shape, lookup_shape, interpolation, boundary, cell_center_offset, separate_camera_batch,
                      normalized_coords)))
    return rc;
  if ((rc = check_mips(mipmapping, num_mipmaps))) return rc;
  if (!input || !lookup || !output) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  size_t need = gft_sample_grid_lut_workspace(input_shape, mipmapping, num_mipmaps);
  if (workspace_bytes < need || (need && !workspace))
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                need);
  hipStream_t st = (hipStream_t)stream;
  Dims3 din = dims_of(input_shape);
  int C = (int)input_shape[4];
  const int N = (int)input_shape[0];
  int levels = mipmapping == GFT_MIP_NONE ? 0 : usable_levels(din, num_mipmaps);
  MipTable mt = make_mip_table(din, C, levels, nvox(din) * C);
  float* atlas = (float*)workspace;
  if (levels && (rc = check_hip(build_atlas(st, input, mt, N, atlas), "mip generation"))) return rc;
  SampleCfg s{interpolation, mipmapping, levels, mip_bias, cell_center_offset};
  LutLaunch L{dims_of(lookup_shape), (int)lookup_shape[0], separate_camera_batch != 0, relative_coords != 0,
              normalized_coords != 0, 0};
  const int per = L.global ? L.num_cams : 1;
  const int max_pairs = 65535;
  const int batches_per_launch = std::max(1, max_pairs / per);
  for (int b0 = 0; b0 < N; b0 += batches_per_launch) {
    int nb = std::min(batches_per_launch, N - b0);
    L.batch0 = b0;
    if ((rc = check_hip(sample_lut_launch(C, boundary, st, input, atlas, mt, s, L, nb * per, lookup, output),
                        "lut sample launch")))
      return rc;
  }
  return GFT_OK;
}

size_t gft_sample_grid_lut_bwd_workspace(const int64_t input_shape[5], int32_t grad_mode) {
  if (grad_mode != GFT_GRAD_REFERENCE) return 0;
  return align256(nvox(dims_of(input_shape)) * sizeof(unsigned));
}

int gft_sample_grid_lut_bwd(void* stream, const float* input, const int64_t input_shape[5], const float* lookup,
                            const int64_t lookup_shape[5], const float* output_grad, int32_t interpolation,
                            int32_t boundary, int32_t mipmapping, float cell_center_offset,
                            int32_t separate_camera_batch, int32_t relative_coords, int32_t normalized_coords,
                            int32_t grad_mode, void* workspace, size_t workspace_bytes, float* input_grad,
                            float* lookup_grad) {
  int rc;
  if ((rc = check_lut(input_shape, lookup_shape, interpolation, boundary, cell_center_offset, separate_camera_batch,
                      normalized_coords)))
    return rc;
  // SampleGridLutGrad registers mipmapping {'NONE'} only (resample_grid.cc:113-134)
  if (mipmapping != GFT_MIP_NONE) return fail(GFT_ERR_INVALID_ARGUMENT, "lut gradient supports mipmapping NONE only");
  if (grad_mode != GFT_GRAD_REFERENCE && grad_mode != GFT_GRAD_EXACT)
    return fail(GFT_ERR_INVALID_ARGUMENT, "invalid grad_mode");
  if (!input || !lookup || !output_grad || !input_grad || !lookup_grad)
    return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  size_t need = gft_sample_grid_lut_bwd_workspace(input_shape, grad_mode);
  if (workspace_bytes < need || (need && !workspace))
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                need);
  hipStream_t st = (hipStream_t)stream;
  Dims3 din = dims_of(input_shape);
  int C = (int)input_shape[4];
  const int N = (int)input_shape[0];
  const size_t vox = nvox(din), bstride = vox * C;
  LutLaunch L{dims_of(lookup_shape), (int)lookup_shape[0], separate_camera_batch != 0, relative_coords != 0,
              normalized_coords != 0, 0};
  const size_t lut_floats = nvox(L.dout) * 4 * (size_t)L.num_cams;
  if ((rc = check_hip(fill_launch(st, input_grad, bstride * N, 0.f), "zero input_grad"))) return rc;
  if ((rc = check_hip(fill_launch(st, lookup_grad, lut_floats, 0.f), "zero lookup_grad"))) return rc;
  const bool exact = grad_mode == GFT_GRAD_EXACT;
  unsigned* cnt = exact ? nullptr : (unsigned*)workspace;
  const int views = L.global ? L.num_cams : 1;
  for (int b = 0; b < N; ++b) {
    L.batch0 = b;
    if (!exact && (rc = check_hip(fill_launch(st, (float*)cnt, vox, 0.f), "zero counter"))) return rc;
    if ((rc = check_hip(sample_lut_bwd_launch(C, boundary, exact, st, input + bstride * b, interpolation,
                                              cell_center_offset, L, views, lookup, output_grad, din,
                                              input_grad + bstride * b, cnt, lookup_grad),
                        "lut gradient launch")))
      return rc;
    if (!exact &&
        (rc = check_hip(normalize_count_launch(C, st, input_grad + bstride * b, cnt, vox), "normalize launch")))
      return rc;
  }
  return GFT_OK;
}

// -------------------------------------------------------------------- LoD
int gft_lod_transform(void* stream, const int32_t input_dhw[3], const float* MV, const float* P, const float* frustum,
                      int32_t num_cameras, const int32_t output_shape[3], int32_t inverse_transform, float* output) {
  int rc;
  if ((rc = check_out_shape(input_dhw))) return rc;
  if ((rc = check_out_shape(output_shape))) return rc;
  if (num_cameras < 1) return fail(GFT_ERR_INVALID_ARGUMENT, "camera batch size mismatch");
  if (!MV || !P || !frustum || !output) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  hipStream_t st = (hipStream_t)stream;
  int cm = inverse_transform ? CM_LINDEPTH_REVERSE : CM_LINDEPTH;
  std::vector<ViewPair> pairs;
  for (int c = 0; c < num_cameras; ++c) pairs.push_back({c, c, c});
  for (size_t first = 0; first < pairs.size(); first += kMaxSampleViews) {
    size_t n = std::min(pairs.size() - first, (size_t)kMaxSampleViews);
    SampleLaunch L;
    if ((rc = build_sample_views(pairs, first, n, nullptr, -1, MV, P, frustum, cm, L))) return rc;
    if ((rc = check_hip(lod_transform_launch(st, dims_dhw(input_dhw), dims_dhw(output_shape), L, output),
                        "lod transform launch")))
      return rc;
  }
  return GFT_OK;
}

// ------------------------------------------------- shadow transmittance
size_t gft_shadow_transmission_saved_bytes(const int32_t shadow_shape[3]) {
  if (!shadow_shape || shadow_shape[0] < 1 || shadow_shape[1] < 1 || shadow_shape[2] < 1) return 0;
  return nvox(dims_dhw(shadow_shape)) * sizeof(float);
}

size_t gft_shadow_transmission_fwd_workspace(const int64_t density_shape[5], int32_t mipmapping,
                                             int32_t num_mipmaps) {
  return gft_sample_grid_transform_workspace(density_shape, mipmapping, num_mipmaps);
}

static int check_shadow_attrs(const int32_t* shadow_shape, int32_t boundary, int32_t blending_mode) {
  int rc;
  if (!shadow_shape) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  if ((rc = check_out_shape(shadow_shape))) return rc;
  if (boundary != GFT_BORDER && boundary != GFT_CLAMP)
    return fail(GFT_ERR_INVALID_ARGUMENT, "boundary must be BORDER or CLAMP");
  if (blending_mode != GFT_BEER_LAMBERT && blending_mode != GFT_ALPHA)
    return fail(GFT_ERR_INVALID_ARGUMENT, "Unknown blend_mode: the shadow map supports BEER_LAMBERT and ALPHA");
  return GFT_OK;
}

int gft_shadow_transmission_fwd(void* stream, const float* density, const int64_t density_shape[5], const float* M,
                                const float* V, const float* P, const float* frustum, const int32_t shadow_shape[3],
                                float depth_step, int32_t interpolation, int32_t boundary, int32_t mipmapping,
                                int32_t num_mipmaps, float mip_bias, int32_t blending_mode, const float* mask,
                                void* workspace, size_t workspace_bytes, float* saved, float* transmission) {
  int rc;
  if (!density_shape) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  if ((rc = check_grid(density_shape, "density"))) return rc;
  if (density_shape[0] != 1 || density_shape[4] != 1)
    return fail(GFT_ERR_INVALID_ARGUMENT, "density must be [1,D,H,W,1], is [%lld,%lld,%lld,%lld,%lld]",
                (long long)density_shape[0], (long long)density_shape[1], (long long)density_shape[2],
                (long long)density_shape[3], (long long)density_shape[4]);
  if ((rc = check_shadow_attrs(shadow_shape, boundary, blending_mode))) return rc;
  if (interpolation != GFT_NEAREST && interpolation != GFT_LINEAR)
    return fail(GFT_ERR_INVALID_ARGUMENT, "interpolation must be NEAREST or LINEAR");
  if ((rc = check_mips(mipmapping, num_mipmaps))) return rc;
  if (!density || !M || !V || !P || !frustum || !transmission)
    return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  size_t need = gft_shadow_transmission_fwd_workspace(density_shape, mipmapping, num_mipmaps);
  if (workspace_bytes < need || (need && !workspace))
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                need);
  hipStream_t st = (hipStream_t)stream;
  Dims3 din = dims_of(density_shape);
  int levels = mipmapping == GFT_MIP_NONE ? 0 : usable_levels(din, num_mipmaps);
  MipTable mt = make_mip_table(din, 1, levels, nvox(din));
  float* atlas = (float*)workspace;
  if (levels && (rc = check_hip(build_atlas(st, density, mt, 1, atlas, depth_step), "mip generation"))) return rc;
  SampleCfg cfg{interpolation, mipmapping, levels, mip_bias, 0.5f};
  std::vector<ViewPair> pairs{{0, 0, 0}};
  SampleLaunch L;
  if ((rc = build_sample_views(pairs, 0, 1, M, -1, V, P, frustum, CM_LINDEPTH, L))) return rc;
  return check_hip(shadow_fwd_launch(boundary, interpolation, blending_mode, st, density, atlas, mt, cfg, L.view[0],
                                     dims_dhw(shadow_shape), depth_step, mask, saved, transmission),
                   "shadow transmission launch");
}

int gft_shadow_transmission_bwd(void* stream, const float* saved, const int32_t shadow_shape[3], int32_t boundary,
                                int32_t blending_mode, const float* mask, const float* transmission_grad,
                                float* frustum_grad) {
  int rc;
  if ((rc = check_shadow_attrs(shadow_shape, boundary, blending_mode))) return rc;
  if (!saved || !transmission_grad || !frustum_grad) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  return check_hip(shadow_bwd_launch(boundary, blending_mode, (hipStream_t)stream, saved, transmission_grad, mask,
                                     dims_dhw(shadow_shape), frustum_grad),
                   "shadow transmission gradient launch");
}

}  // extern "C"
