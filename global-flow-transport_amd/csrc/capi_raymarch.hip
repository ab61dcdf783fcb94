// C ABI of the fused raymarch (RaymarchGridTransform and its gradient).
// Validation mirrors RaymarchGridTransformOp / RaymarchGridTransformGradOp
// (phitest/render/cuda/src/raymarch_grid.cc:192-297, :347-481).
#include "gft_host.hpp"

namespace gft {

hipError_t raymarch_fwd_launch(int C, int fm, int bm, int bl, hipStream_t s, const float* in, float* out,
                               const RayDims& rd, const RayLaunch& L);
hipError_t raymarch_bwd_launch(int C, bool exact, int bm, int bl, hipStream_t s, const float* in, const float* o,
                               const float* og, float* g, float* w, const RayDims& rd, const RayLaunch& L);
hipError_t raymarch_bwd_walk_launch(int C, bool exact, int bl, hipStream_t s, const float* in, const float* o,
                                    const float* og, float* gsamp, int2* ranges, const RayDims& rd,
                                    const RayLaunch& L);
hipError_t raymarch_bwd_gather_launch(int C, bool exact, hipStream_t s, const float* gsamp, const int2* ranges,
                                      float* grad, const RayDims& rd, const ProjLaunch& PL);
hipError_t normalize_add_launch(int C, hipStream_t s, float* total, float* buf, float* wsum, size_t nvox);
hipError_t fill_launch(hipStream_t s, float* p, size_t n, float v);

static int validate_common(const int64_t* ishape, int32_t num_M, int32_t num_cameras, const int32_t* oshape,
                           int32_t interpolation, int32_t boundary, int32_t blending_mode, int32_t keep_dims,
                           int32_t separate_camera_batch) {
  for (int i = 0; i < 5; ++i)
    if (ishape[i] < 1) return fail(GFT_ERR_INVALID_ARGUMENT, "Empty input (NDHWC): dimension %d is %lld", i, (long long)ishape[i]);
  const int64_t C = ishape[4];
  if (C > 4) return fail(GFT_ERR_INVALID_ARGUMENT, "Invalid input shape (NDHWC): channel %lld > 4", (long long)C);
  if (C == 3) return fail(GFT_ERR_UNIMPLEMENTED, "Only 1,2 and 4 Channel supported.");
  for (int i = 0; i < 3; ++i)
    if (oshape[i] < 1)
      return fail(GFT_ERR_INVALID_ARGUMENT, "All shape dimensions must be positive, dimension %d is %d", i, oshape[i]);
  if (num_M != ishape[0]) return fail(GFT_ERR_INVALID_ARGUMENT, "model matrix batch size mismatch");
  if (num_cameras < 1) return fail(GFT_ERR_INVALID_ARGUMENT, "camera batch size mismatch");
  if (!separate_camera_batch && num_cameras != ishape[0])
    return fail(GFT_ERR_INVALID_ARGUMENT, "camera batch must match data batch when not using global sampling.");
  if (interpolation < 0 || interpolation > 3) return fail(GFT_ERR_INVALID_ARGUMENT, "invalid filter mode");
  if (boundary < 0 || boundary > 2) return fail(GFT_ERR_INVALID_ARGUMENT, "invalid boundary mode");
  if (blending_mode < 0 || blending_mode > 3) return fail(GFT_ERR_INVALID_ARGUMENT, "invalid blend mode");
  // The reference forces the output batch to 1 without keep_dims and would
  // then write past it for N > 1 (raymarch_grid.cc:256); reject instead.
  if (!keep_dims && ishape[0] != 1)
    return fail(GFT_ERR_INVALID_ARGUMENT, "keep_dims=False requires a data batch of 1 (got %lld)", (long long)ishape[0]);
  if ((int64_t)oshape[1] * oshape[2] > (int64_t)INT32_MAX)
    return fail(GFT_ERR_INVALID_ARGUMENT, "output image too large");
  return GFT_OK;
}

static int build_views(const std::vector<ViewPair>& pairs, size_t first, size_t count, const float* M, const float* V,
                       const float* P, RayLaunch& L) {
  L.num_views = (int)count;
  for (size_t i = 0; i < count; ++i) {
    const ViewPair& vp = pairs[first + i];
    Mat4d m = mat_from_colmajor(M + 16 * vp.batch);
    Mat4d v = mat_from_colmajor(V + 16 * vp.cam);
    Mat4d p = mat_from_colmajor(P + 16 * vp.cam);
    Mat4d mv = mat_mul(v, m), mvinv, pinv;
    if (!mat_inverse(mv, mvinv)) return fail(GFT_ERR_INVALID_ARGUMENT, "model-view matrix is singular");
    if (!mat_inverse(p, pinv)) return fail(GFT_ERR_INVALID_ARGUMENT, "projection matrix is singular");
    mat_to_float(pinv, L.view[i].pinv);
    mat_to_float(mvinv, L.view[i].mvinv);
    L.view[i].in_batch = vp.batch;
    L.view[i].out_slot = vp.slot;
  }
  return GFT_OK;
}

static RayDims make_dims(const int64_t* ishape, const int32_t* oshape) {
  RayDims rd;
  rd.in = Dims3{(int)ishape[3], (int)ishape[2], (int)ishape[1]};
  rd.C = (int)ishape[4];
  rd.out_d = oshape[0];
  rd.out_h = oshape[1];
  rd.out_w = oshape[2];
  rd.in_batch_stride = (size_t)ishape[1] * ishape[2] * ishape[3] * ishape[4];
  rd.out_view_stride = (size_t)oshape[1] * oshape[2] * ishape[4];
  return rd;
}

}  // namespace gft

using namespace gft;

extern "C" {

int gft_raymarch_grid_transform_fwd(void* stream, const float* input, const int64_t input_shape[5], const float* M,
                                    int32_t num_M, const float* V, const float* P, const float* frustum,
                                    int32_t num_cameras, const int32_t output_shape[3], int32_t interpolation,
                                    int32_t boundary, int32_t blending_mode, int32_t keep_dims,
                                    int32_t separate_camera_batch, float* output) {
  (void)frustum;  // the fused raymarch samples in perspective NDC depth; the frustum is unused there
  int rc = validate_common(input_shape, num_M, num_cameras, output_shape, interpolation, boundary, blending_mode,
                           keep_dims, separate_camera_batch);
  if (rc) return rc;
  if (!input || !output || !M || !V || !P) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  hipStream_t s = (hipStream_t)stream;
  RayDims rd = make_dims(input_shape, output_shape);
  auto pairs = view_pairs((int)input_shape[0], num_cameras, separate_camera_batch != 0);
  for (size_t first = 0; first < pairs.size(); first += kMaxViewsPerLaunch) {
    size_t cnt = std::min(pairs.size() - first, (size_t)kMaxViewsPerLaunch);
    RayLaunch L;
    if ((rc = build_views(pairs, first, cnt, M, V, P, L))) return rc;
    if ((rc = check_hip(raymarch_fwd_launch(rd.C, interpolation, boundary, blending_mode, s, input, output, rd, L),
                        "raymarch forward launch")))
      return rc;
  }
  return GFT_OK;
}

// BORDER walk + gather workspace for one chunk of at most kMaxViewsPerLaunch
// views of one input batch: per-sample gradients [view][S][H][W][C] and the
// per-pixel sample ranges.
static size_t walk_chunk_views(int32_t num_cameras, int32_t separate_camera_batch) {
  return (size_t)std::min<int32_t>(separate_camera_batch ? num_cameras : 1, kMaxViewsPerLaunch);
}
static size_t walk_samples_bytes(const int64_t* input_shape, const int32_t* output_shape, size_t views) {
  const size_t plane = (size_t)output_shape[1] * output_shape[2];
  return (views * output_shape[0] * plane * input_shape[4] * sizeof(float) + 255) & ~size_t(255);
}

size_t gft_raymarch_grid_transform_bwd_workspace(const int64_t input_shape[5], int32_t num_cameras,
                                                 const int32_t output_shape[3], int32_t boundary,
                                                 int32_t separate_camera_batch, int32_t grad_mode) {
  if (boundary == GFT_BORDER) {
    const size_t views = walk_chunk_views(num_cameras, separate_camera_batch);
    const size_t plane = (size_t)output_shape[1] * output_shape[2];
    return walk_samples_bytes(input_shape, output_shape, views) + views * plane * sizeof(int2);
  }
  if (grad_mode == GFT_GRAD_EXACT) return 0;
  size_t vox = (size_t)input_shape[1] * input_shape[2] * input_shape[3];
  size_t a = ((vox * input_shape[4] * sizeof(float)) + 255) & ~size_t(255);
  return a + vox * sizeof(float);
}

int gft_raymarch_grid_transform_bwd(void* stream, const float* input, const int64_t input_shape[5],
                                    const float* output, const float* output_grad, const float* M, int32_t num_M,
                                    const float* V, const float* P, const float* frustum, int32_t num_cameras,
                                    const int32_t output_shape[3], int32_t interpolation, int32_t boundary,
                                    int32_t blending_mode, int32_t keep_dims, int32_t separate_camera_batch,
                                    int32_t grad_mode, void* workspace, size_t workspace_bytes, float* input_grad) {
  (void)frustum;
  int rc = validate_common(input_shape, num_M, num_cameras, output_shape, interpolation, boundary, blending_mode,
                           keep_dims, separate_camera_batch);
  if (rc) return rc;
  if (interpolation != GFT_LINEAR)
    return fail(GFT_ERR_INVALID_ARGUMENT, "raymarch gradient supports interpolation LINEAR only");
  if (grad_mode != GFT_GRAD_REFERENCE && grad_mode != GFT_GRAD_EXACT)
    return fail(GFT_ERR_INVALID_ARGUMENT, "invalid grad_mode");
  if (!input || !output || !output_grad || !input_grad || !M || !V || !P)
    return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  size_t need = gft_raymarch_grid_transform_bwd_workspace(input_shape, num_cameras, output_shape, boundary,
                                                          separate_camera_batch, grad_mode);
  if (workspace_bytes < need || (need && !workspace))
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                need);
  hipStream_t s = (hipStream_t)stream;
  RayDims rd = make_dims(input_shape, output_shape);
  const size_t total = rd.in_batch_stride * (size_t)input_shape[0];
  if ((rc = check_hip(fill_launch(s, input_grad, total, 0.f), "zero input_grad"))) return rc;
  auto pairs = view_pairs((int)input_shape[0], num_cameras, separate_camera_batch != 0);
  const bool exact = grad_mode == GFT_GRAD_EXACT;
  if (boundary == GFT_BORDER) {
    // walk + gather, per input batch and chunk of views of that batch
    const size_t views = walk_chunk_views(num_cameras, separate_camera_batch);
    float* gsamp = (float*)workspace;
    int2* ranges = (int2*)((char*)workspace + walk_samples_bytes(input_shape, output_shape, views));
    size_t first = 0;
    while (first < pairs.size()) {
      size_t cnt = 1;
      while (first + cnt < pairs.size() && cnt < (size_t)kMaxViewsPerLaunch &&
             pairs[first + cnt].batch == pairs[first].batch)
        ++cnt;
      RayLaunch L;
      if ((rc = build_views(pairs, first, cnt, M, V, P, L))) return rc;
      ProjLaunch PL;
      PL.num_views = (int)cnt;
      PL.accumulate = 1;
      PL.inv_zmax = (float)(1.0 / (rd.out_d > 1 ? rd.out_d - 1 : 1));
      for (size_t i = 0; i < cnt; ++i) {
        const ViewPair& vp = pairs[first + i];
        Mat4d m = mat_from_colmajor(M + 16 * vp.batch), v = mat_from_colmajor(V + 16 * vp.cam);
        Mat4d p = mat_from_colmajor(P + 16 * vp.cam), mv = mat_mul(v, m), mvinv, pinv;
        if (!mat_inverse(mv, mvinv)) return fail(GFT_ERR_INVALID_ARGUMENT, "model-view matrix is singular");
        if (!mat_inverse(p, pinv)) return fail(GFT_ERR_INVALID_ARGUMENT, "projection matrix is singular");
        mat_to_float(mat_mul(p, mv), PL.view[i].pmv);
        ray_coefficients(pinv, mvinv, rd, PL.view[i]);
        PL.view[i].out_slot = (int)i;  // launch-local index into the walk's buffers
      }
      if ((rc = check_hip(raymarch_bwd_walk_launch(rd.C, exact, blending_mode, s, input, output, output_grad, gsamp,
                                                   ranges, rd, L),
                          "raymarch backward walk")))
        return rc;
      if ((rc = check_hip(raymarch_bwd_gather_launch(rd.C, exact, s, gsamp, ranges,
                                                     input_grad + rd.in_batch_stride * pairs[first].batch, rd, PL),
                          "raymarch backward gather")))
        return rc;
      first += cnt;
    }
    return GFT_OK;
  }
  if (exact) {
    for (size_t first = 0; first < pairs.size(); first += kMaxViewsPerLaunch) {
      size_t cnt = std::min(pairs.size() - first, (size_t)kMaxViewsPerLaunch);
      RayLaunch L;
      if ((rc = build_views(pairs, first, cnt, M, V, P, L))) return rc;
      if ((rc = check_hip(raymarch_bwd_launch(rd.C, true, boundary, blending_mode, s, input, output, output_grad,
                                              input_grad, nullptr, rd, L),
                          "raymarch backward launch")))
        return rc;
    }
    return GFT_OK;
  }
  // REFERENCE (CLAMP / WRAP): per camera scatter into a buffer with weights,
  // then normalise by the weight sum and add (raymarch_grid.cu.cc:504-570).
  const size_t vox = (size_t)input_shape[1] * input_shape[2] * input_shape[3];
  float* buf = (float*)workspace;
  float* wsum = (float*)((char*)workspace + (((vox * rd.C * sizeof(float)) + 255) & ~size_t(255)));
  if ((rc = check_hip(fill_launch(s, buf, vox * rd.C, 0.f), "zero buffer"))) return rc;
  if ((rc = check_hip(fill_launch(s, wsum, vox, 0.f), "zero weights"))) return rc;
  for (size_t i = 0; i < pairs.size(); ++i) {
    RayLaunch L;
    if ((rc = build_views(pairs, i, 1, M, V, P, L))) return rc;
    if ((rc = check_hip(raymarch_bwd_launch(rd.C, false, boundary, blending_mode, s, input, output, output_grad, buf,
                                            wsum, rd, L),
                        "raymarch backward launch")))
      return rc;
    if ((rc = check_hip(normalize_add_launch(rd.C, s, input_grad + rd.in_batch_stride * pairs[i].batch, buf, wsum, vox),
                        "normalize launch")))
      return rc;
  }
  return GFT_OK;
}

}  // extern "C"
