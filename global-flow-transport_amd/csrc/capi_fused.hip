// C ABI of the fused transport/render path: MacCormack advection with the
// MC_SMOOTH clamp (VelocityGrid.warp, data_structures.py:870-905) and the
// ambient-light density render with the L2 image loss
// (Renderer.render_density, renderer.py:603-651; loss_dens_target,
// optimization.py:396-412).
#include <mutex>

#include "gft_host.hpp"

namespace gft {


hipError_t advect_fwd_launch(hipStream_t s, const void* dens, const void* vel, bool bf16, float* f, float* out,
                             uint8_t* mask, const AdvectDims& a);
hipError_t advect_bwd_launch(hipStream_t s, const void* dens, const void* vel, bool bf16, const float* f,
                             const uint8_t* mask, const float* gout, float* gf_acc, float* gd, float* gv,
                             const AdvectDims& a);
hipError_t render_density_fwd_launch(int bm, hipStream_t s, const float* dens, float* img, const RayDims& rd,
                                     const RayLaunch& L, const RenderParams& rp, const RenderSaved& sv);
size_t render_tiles_per_view(const RayDims& rd);
hipError_t render_density_pixel_launch(bool l2, hipStream_t s, const float* img, const float* second,
                                       const RenderSaved& sv, const RayDims& rd, const RayLaunch& L,
                                       const RenderParams& rp);
hipError_t render_density_gather_launch(hipStream_t s, const RenderSaved& sv, float* grad, const RayDims& rd,
                                        const ProjLaunch& PL, float* loss, float loss_scale);
hipError_t render_density_scatter_launch(int bm, bool l2, hipStream_t s, const float* dens, const float* img,
                                         const float* second, float* grad, float* loss, const RayDims& rd,
                                         const RayLaunch& L, const RenderParams& rp);
hipError_t fill_launch(hipStream_t s, float* p, size_t n, float v);
hipError_t quad_copy_launch(hipStream_t s, int bm, const float* lin, float* out, Dims3 d, const float* fac = nullptr);
size_t quad_copy_bf16_bytes(Dims3 d);
hipError_t quad_copy_bf16_launch(hipStream_t s, const float* lin, void* out, Dims3 d);
hipError_t render_density_fwd_bf16_launch(hipStream_t s, const void* quad, float* img, const RayDims& rd,
                                          const RayLaunch& L, const RenderParams& rp, const RenderSaved& sv);
hipError_t render_density_lit_fwd_launch(hipStream_t s, const float* dq, const float* eq, float* img,
                                         const RayDims& rd, const RayLaunch& L, const RenderParams& rp,
                                         const RenderSaved& sv);
hipError_t render_density_lit_gather_launch(hipStream_t s, const RenderSaved& sv, const float* dens,
                                            const float* lam, float* grad, float* lam_grad, const RayDims& rd,
                                            const ProjLaunch& PL, float* loss, float loss_scale);
size_t quad_copy_floats(Dims3 d);

// Bytes of the corner-quad density copy the BORDER render kernels read.
static size_t quad_bytes(const int64_t* shape) {
  return (quad_copy_floats(Dims3{(int)shape[3], (int)shape[2], (int)shape[1]}) * sizeof(float) + 255) & ~size_t(255);
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

static int check_scalar_grid(const int64_t* shape) {
  for (int i = 0; i < 5; ++i)
    if (shape[i] < 1) return fail(GFT_ERR_INVALID_ARGUMENT, "Empty input (NDHWC): dimension %d is %lld", i, (long long)shape[i]);
  if (shape[0] != 1) return fail(GFT_ERR_INVALID_ARGUMENT, "fused transport/render expects a data batch of 1");
  if (shape[4] != 1) return fail(GFT_ERR_UNIMPLEMENTED, "fused transport/render expects a 1-channel density grid");
  return GFT_OK;
}

static int check_advect_mode(int32_t order, int32_t clamp, float dt) {
  if (order != 1 && order != 2) return fail(GFT_ERR_INVALID_ARGUMENT, "Unsupported warp order '%d'", order);
  if (clamp < 0 || clamp > 2) return fail(GFT_ERR_INVALID_ARGUMENT, "invalid warp clamp %d", clamp);
  if (!std::isfinite(dt)) return fail(GFT_ERR_INVALID_ARGUMENT, "dt must be finite");
  return GFT_OK;
}

static int ray_views(const std::vector<ViewPair>& pairs, size_t first, size_t count, const float* M, const float* V,
                     const float* P, RayLaunch& L, ProjLaunch* PL = nullptr, const RayDims* rd = nullptr) {
  L.num_views = (int)count;
  if (PL) PL->num_views = (int)count;
  for (size_t i = 0; i < count; ++i) {
    const ViewPair& vp = pairs[first + i];
    Mat4d mv = mat_mul(mat_from_colmajor(V + 16 * vp.cam), mat_from_colmajor(M + 16 * vp.batch)), mvinv, pinv;
    Mat4d p = mat_from_colmajor(P + 16 * vp.cam);
    if (!mat_inverse(mv, mvinv)) return fail(GFT_ERR_INVALID_ARGUMENT, "model-view matrix is singular");
    if (!mat_inverse(p, pinv)) return fail(GFT_ERR_INVALID_ARGUMENT, "projection matrix is singular");
    mat_to_float(pinv, L.view[i].pinv);
    mat_to_float(mvinv, L.view[i].mvinv);
    L.view[i].in_batch = vp.batch;
    L.view[i].out_slot = vp.slot;
    if (PL) {
      mat_to_float(mat_mul(p, mv), PL->view[i].pmv);
      ray_coefficients(pinv, mvinv, *rd, PL->view[i]);
      PL->view[i].out_slot = vp.slot;
      PL->inv_zmax = (float)(1.0 / (rd->out_d > 1 ? rd->out_d - 1 : 1));
    }
  }
  return GFT_OK;
}

static RayDims render_dims(const int64_t* shape, const int32_t* output_shape) {
  RayDims rd;
  rd.in = Dims3{(int)shape[3], (int)shape[2], (int)shape[1]};
  rd.C = 1;
  rd.out_d = output_shape[0];
  rd.out_h = output_shape[1];
  rd.out_w = output_shape[2];
  rd.in_batch_stride = (size_t)shape[1] * shape[2] * shape[3];
  rd.out_view_stride = (size_t)rd.out_h * rd.out_w * 4;
  return rd;
}

static int check_render_args(const int64_t* shape, int32_t num_M, int32_t num_cameras, const int32_t* output_shape,
                             int32_t boundary) {
  int rc;
  if ((rc = check_scalar_grid(shape))) return rc;
  if (num_M != 1) return fail(GFT_ERR_INVALID_ARGUMENT, "model matrix batch size mismatch");
  if (num_cameras < 1) return fail(GFT_ERR_INVALID_ARGUMENT, "camera batch size mismatch");
  for (int i = 0; i < 3; ++i)
    if (output_shape[i] < 1)
      return fail(GFT_ERR_INVALID_ARGUMENT, "All shape dimensions must be positive, dimension %d is %d", i, output_shape[i]);
  if (boundary < 0 || boundary > 2) return fail(GFT_ERR_INVALID_ARGUMENT, "invalid boundary mode");
  if ((int64_t)shape[1] * shape[2] * shape[3] >= ((int64_t)1 << 31))
    return fail(GFT_ERR_INVALID_ARGUMENT, "grid too large (2^31 voxels)");
  return GFT_OK;
}

}  // namespace gft

using namespace gft;

extern "C" {

size_t gft_advect_saved_bytes(const int64_t shape[5]) {
  size_t vox = (size_t)shape[1] * shape[2] * shape[3];
  return align256(vox * sizeof(float)) + align256(vox);
}

}  // extern "C"

namespace gft {

static int advect_fwd_common(void* stream, const void* input, const int64_t shape[5], const void* velocity, bool bf16,
                             float dt, int32_t order, int32_t clamp, float* output, void* saved, size_t saved_bytes) {
  int rc;
  if ((rc = check_scalar_grid(shape)) || (rc = check_advect_mode(order, clamp, dt))) return rc;
  if (!input || !velocity || !output) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  size_t need = gft_advect_saved_bytes(shape);
  if (!saved || saved_bytes < need)
    return fail(GFT_ERR_WORKSPACE, "saved buffer of %zu bytes is smaller than the %zu bytes required", saved_bytes, need);
  size_t vox = (size_t)shape[1] * shape[2] * shape[3];
  AdvectDims a{Dims3{(int)shape[3], (int)shape[2], (int)shape[1]}, dt, order, clamp};
  float* f = (float*)saved;
  uint8_t* mask = (uint8_t*)saved + align256(vox * sizeof(float));
  return check_hip(advect_fwd_launch((hipStream_t)stream, input, velocity, bf16, f, output, mask, a),
                   "advect forward");
}

static int advect_bwd_common(void* stream, const void* input, const int64_t shape[5], const void* velocity, bool bf16,
                             float dt, int32_t order, int32_t clamp, const void* saved, const float* output_grad,
                             float* input_grad, float* velocity_grad, void* workspace, size_t workspace_bytes) {
  int rc;
  if ((rc = check_scalar_grid(shape)) || (rc = check_advect_mode(order, clamp, dt))) return rc;
  if (!input || !velocity || !saved || !output_grad || !input_grad || !velocity_grad)
    return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  size_t need = gft_advect_bwd_workspace(shape);
  if (order == 2 && (!workspace || workspace_bytes < need))
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes, need);
  hipStream_t s = (hipStream_t)stream;
  size_t vox = (size_t)shape[1] * shape[2] * shape[3];
  AdvectDims a{Dims3{(int)shape[3], (int)shape[2], (int)shape[1]}, dt, order, clamp};
  const float* f = (const float*)saved;
  const uint8_t* mask = (const uint8_t*)saved + align256(vox * sizeof(float));
  float* gf_acc = (float*)workspace;
  if (order == 2) {
    if ((rc = check_hip(fill_launch(s, gf_acc, vox, 0.f), "clear gf"))) return rc;
  } else {
    if ((rc = check_hip(fill_launch(s, input_grad, vox, 0.f), "clear input_grad"))) return rc;
  }
  return check_hip(
      advect_bwd_launch(s, input, velocity, bf16, f, mask, output_grad, gf_acc, input_grad, velocity_grad, a),
      "advect backward");
}

}  // namespace gft

extern "C" {

int gft_advect_fwd(void* stream, const float* input, const int64_t shape[5], const float* velocity, float dt,
                   int32_t order, int32_t clamp, float* output, void* saved, size_t saved_bytes) {
  return advect_fwd_common(stream, input, shape, velocity, false, dt, order, clamp, output, saved, saved_bytes);
}

int gft_advect_fwd_bf16(void* stream, const uint16_t* input, const int64_t shape[5], const uint16_t* velocity,
                        float dt, int32_t order, int32_t clamp, float* output, void* saved, size_t saved_bytes) {
  return advect_fwd_common(stream, input, shape, velocity, true, dt, order, clamp, output, saved, saved_bytes);
}

size_t gft_advect_bwd_workspace(const int64_t shape[5]) {
  return align256((size_t)shape[1] * shape[2] * shape[3] * sizeof(float));
}

int gft_advect_bwd(void* stream, const float* input, const int64_t shape[5], const float* velocity, float dt,
                   int32_t order, int32_t clamp, const void* saved, const float* output_grad, float* input_grad,
                   float* velocity_grad, void* workspace, size_t workspace_bytes) {
  return advect_bwd_common(stream, input, shape, velocity, false, dt, order, clamp, saved, output_grad, input_grad,
                           velocity_grad, workspace, workspace_bytes);
}

int gft_advect_bwd_bf16(void* stream, const uint16_t* input, const int64_t shape[5], const uint16_t* velocity,
                        float dt, int32_t order, int32_t clamp, const void* saved, const float* output_grad,
                        float* input_grad, float* velocity_grad, void* workspace, size_t workspace_bytes) {
  return advect_bwd_common(stream, input, shape, velocity, true, dt, order, clamp, saved, output_grad, input_grad,
                           velocity_grad, workspace, workspace_bytes);
}

}  // extern "C"

// ------------------------------------------------ no-allocation advection adjoint
// gft_advect_bwd_noalloc keeps the MacCormack adjoint's accumulator of the
// forward-trace gradient (one float per voxel, the workspace of gft_advect_bwd)
// in this file-scope array instead of caller memory, so the call allocates
// nothing. Every process that loads the library holds it (64 MiB) on each GPU it
// runs on until the process ends. The accumulator is cleared by a fill kernel at
// the start of every call; concurrent calls on one GPU share it and race.
static __device__ __align__(256) float scratch[GFT_ADVECT_NOALLOC_MAX_VOXELS];

namespace gft {

static constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mutex;
static void* g_scratch_ptr[kMaxScratchDevices];  // device address of `scratch` on each device

// Device address of `scratch` on the current device, looked up once per device
// with that device current. Host code never passes the symbol itself to a kernel.
static int scratch_on_current_device(float** out) {
  int dev = 0;
  int rc;
  if ((rc = check_hip(hipGetDevice(&dev), "hipGetDevice"))) return rc;
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(GFT_ERR_RUNTIME, "device %d beyond the %d devices the no-allocation scratch tracks", dev,
                kMaxScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if ((rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))) return rc;
    if (!p) return fail(GFT_ERR_RUNTIME, "hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[dev] = p;
  }
  *out = (float*)g_scratch_ptr[dev];
  return GFT_OK;
}

}  // namespace gft

extern "C" {

int gft_advect_bwd_noalloc(void* stream, const float* input, const int64_t shape[5], const float* velocity, float dt,
                           int32_t order, int32_t clamp, const void* saved, const float* output_grad,
                           float* input_grad, float* velocity_grad) {
  int rc;
  if ((rc = check_scalar_grid(shape))) return rc;
  const size_t vox = (size_t)shape[1] * shape[2] * shape[3];
  if (vox > (size_t)GFT_ADVECT_NOALLOC_MAX_VOXELS)
    return fail(GFT_ERR_INVALID_ARGUMENT,
                "grid of %zu voxels exceeds the %d voxels of the no-allocation scratch; use gft_advect_bwd with a "
                "workspace",
                vox, (int)GFT_ADVECT_NOALLOC_MAX_VOXELS);
  float* ws = nullptr;
  if ((rc = scratch_on_current_device(&ws))) return rc;
  return advect_bwd_common(stream, input, shape, velocity, false, dt, order, clamp, saved, output_grad, input_grad,
                           velocity_grad, ws, (size_t)GFT_ADVECT_NOALLOC_MAX_VOXELS * sizeof(float));
}

size_t gft_render_density_fwd_workspace(const int64_t shape[5], int32_t boundary) {
  return boundary == GFT_BORDER || boundary == GFT_CLAMP ? quad_bytes(shape) : 0;
}

size_t gft_render_density_saved_bytes(const int64_t shape[5], int32_t num_cameras, const int32_t output_shape[3],
                                      int32_t boundary) {
  if (boundary != GFT_BORDER) return 0;
  const RayDims rd = render_dims(shape, output_shape);
  const size_t plane = (size_t)output_shape[1] * output_shape[2];
  const size_t nsamples = render_waves_per_view(output_shape[2], output_shape[1]) * kSampleRun * output_shape[0];
  return align256(nsamples * num_cameras * sizeof(float)) + align256(plane * num_cameras * sizeof(float4)) +
         align256(render_tiles_per_view(rd) * num_cameras * sizeof(float));
}

// the saved buffer: per-sample values, per-pixel records,
// per-tile loss partials (the layout of RenderSaved)
static RenderSaved saved_parts(void* saved, const int64_t* shape, const int32_t* output_shape, int32_t num_cameras) {
  const RayDims rd = render_dims(shape, output_shape);
  const size_t plane = (size_t)output_shape[1] * output_shape[2];
  const size_t nsamples = render_waves_per_view(output_shape[2], output_shape[1]) * kSampleRun * output_shape[0];
  char* p = (char*)saved;
  RenderSaved sv{};
  sv.samples = (float*)p;
  p += align256(nsamples * num_cameras * sizeof(float));
  sv.records = (float4*)p;
  p += align256(plane * num_cameras * sizeof(float4));
  sv.partials = (float*)p;
  sv.n_partials = (int)(render_tiles_per_view(rd) * num_cameras);
  return sv;
}

int gft_render_density_fwd(void* stream, const float* density, const int64_t shape[5], const float* M, int32_t num_M,
                           const float* V, const float* P, const float* frustum, int32_t num_cameras,
                           const int32_t output_shape[3], const float light[3], float depth_step, int32_t boundary,
                           void* workspace, size_t workspace_bytes, void* saved, size_t saved_bytes, float* images) {
  (void)frustum;
  int rc;
  if ((rc = check_render_args(shape, num_M, num_cameras, output_shape, boundary))) return rc;
  if (!density || !images || !M || !V || !P || !light) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  const size_t need = gft_render_density_fwd_workspace(shape, boundary);
  if (need && (!workspace || workspace_bytes < need))
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                need);
  const size_t need_saved = gft_render_density_saved_bytes(shape, num_cameras, output_shape, boundary);
  if (saved && saved_bytes < need_saved)
    return fail(GFT_ERR_WORKSPACE, "saved buffer of %zu bytes is smaller than the %zu bytes required", saved_bytes,
                need_saved);
  RenderSaved sv{nullptr, nullptr, nullptr};
  if (saved && need_saved) sv = saved_parts(saved, shape, output_shape, num_cameras);
  const RayDims rd = render_dims(shape, output_shape);
  const RenderParams rp{{light[0], light[1], light[2]}, depth_step, 0.f};
  auto pairs = view_pairs(1, num_cameras, true);
  if (boundary == GFT_BORDER || boundary == GFT_CLAMP) {
    if ((rc = check_hip(quad_copy_launch((hipStream_t)stream, boundary == GFT_CLAMP ? B_CLAMP : B_BORDER, density,
                                         (float*)workspace, rd.in),
                        "quad copy")))
      return rc;
    density = (const float*)workspace;
  }
  for (size_t first = 0; first < pairs.size(); first += kMaxViewsPerLaunch) {
    size_t cnt = std::min(pairs.size() - first, (size_t)kMaxViewsPerLaunch);
    RayLaunch L;
    if ((rc = ray_views(pairs, first, cnt, M, V, P, L))) return rc;
    if ((rc = check_hip(render_density_fwd_launch(boundary, (hipStream_t)stream, density, images, rd, L, rp, sv),
                        "render forward")))
      return rc;
  }
  return GFT_OK;
}

int gft_render_density_fwd_bf16(void* stream, const float* density, const int64_t shape[5], const float* M,
                                int32_t num_M, const float* V, const float* P, const float* frustum,
                                int32_t num_cameras, const int32_t output_shape[3], const float light[3],
                                float depth_step, int32_t boundary, void* workspace, size_t workspace_bytes,
                                void* saved, size_t saved_bytes, float* images) {
  (void)frustum;
  int rc;
  if ((rc = check_render_args(shape, num_M, num_cameras, output_shape, boundary))) return rc;
  if (boundary != GFT_BORDER)
    return fail(GFT_ERR_UNIMPLEMENTED, "the bfloat16 density render supports the BORDER boundary only");
  if (!density || !images || !M || !V || !P || !light) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  const RayDims rd = render_dims(shape, output_shape);
  const size_t need = align256(quad_copy_bf16_bytes(rd.in));
  if (!workspace || workspace_bytes < need)
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                need);
  const size_t need_saved = gft_render_density_saved_bytes(shape, num_cameras, output_shape, boundary);
  if (saved && saved_bytes < need_saved)
    return fail(GFT_ERR_WORKSPACE, "saved buffer of %zu bytes is smaller than the %zu bytes required", saved_bytes,
                need_saved);
  RenderSaved sv{};
  if (saved) sv = saved_parts(saved, shape, output_shape, num_cameras);
  const RenderParams rp{{light[0], light[1], light[2]}, depth_step, 0.f};
  hipStream_t s = (hipStream_t)stream;
  if ((rc = check_hip(quad_copy_bf16_launch(s, density, workspace, rd.in), "bfloat16 quad copy"))) return rc;
  auto pairs = view_pairs(1, num_cameras, true);
  for (size_t first = 0; first < pairs.size(); first += kMaxViewsPerLaunch) {
    size_t cnt = std::min(pairs.size() - first, (size_t)kMaxViewsPerLaunch);
    RayLaunch L;
    if ((rc = ray_views(pairs, first, cnt, M, V, P, L))) return rc;
    if ((rc = check_hip(render_density_fwd_bf16_launch(s, workspace, images, rd, L, rp, sv),
                        "bfloat16 render forward")))
      return rc;
  }
  return GFT_OK;
}

size_t gft_render_density_bwd_workspace(const int64_t shape[5], int32_t num_cameras, const int32_t output_shape[3],
                                        int32_t boundary) {
  (void)num_cameras;
  (void)output_shape;
  // the corner-quad copy of the scatter walk (CLAMP, and BORDER without a saved buffer)
  return boundary == GFT_CLAMP || boundary == GFT_BORDER ? quad_bytes(shape) : 0;
}

static int render_bwd_common(void* stream, const float* density, const int64_t shape[5], const float* M,
                             int32_t num_M, const float* V, const float* P, int32_t num_cameras,
                             const int32_t output_shape[3], const float light[3], float depth_step, int32_t boundary,
                             const float* images, const float* second, bool l2, float loss_scale,
                             int32_t loss_kind, const float* backgrounds, void* saved, size_t saved_bytes,
                             void* workspace, size_t workspace_bytes, float* density_grad, float* loss) {
  int rc;
  if ((rc = check_render_args(shape, num_M, num_cameras, output_shape, boundary))) return rc;
  if (loss_kind != GFT_LOSS_SE && loss_kind != GFT_LOSS_AE) return fail(GFT_ERR_INVALID_ARGUMENT, "invalid loss_kind");
  if (!density || !images || !second || !density_grad || (l2 && !loss) || !M || !V || !P || !light)
    return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  const size_t need = gft_render_density_saved_bytes(shape, num_cameras, output_shape, boundary);
  if (saved && saved_bytes < need)
    return fail(GFT_ERR_WORKSPACE, "saved buffer of %zu bytes is smaller than the %zu bytes required", saved_bytes,
                need);
  hipStream_t s = (hipStream_t)stream;
  const RayDims rd = render_dims(shape, output_shape);
  const RenderParams rp{{light[0], light[1], light[2]}, depth_step, loss_scale, loss_kind, backgrounds};
  auto pairs = view_pairs(1, num_cameras, true);
  if (boundary == GFT_BORDER && saved) {
    const RenderSaved sv = saved_parts(saved, shape, output_shape, num_cameras);
    for (size_t first = 0; first < pairs.size(); first += kMaxViewsPerLaunch) {
      size_t cnt = std::min(pairs.size() - first, (size_t)kMaxViewsPerLaunch);
      RayLaunch L;
      ProjLaunch PL;
      if ((rc = ray_views(pairs, first, cnt, M, V, P, L, &PL, &rd))) return rc;
      PL.accumulate = first > 0;
      if ((rc = check_hip(render_density_pixel_launch(l2, s, images, second, sv, rd, L, rp),
                          "render backward pixel pass")))
        return rc;
      // the last chunk's gather also sums the per-tile loss partials of every view in a fixed order
      const bool lastc = first + cnt == pairs.size();
      if ((rc = check_hip(render_density_gather_launch(s, sv, density_grad, rd, PL, (l2 && lastc) ? loss : nullptr,
                                                       loss_scale),
                          "render backward gather")))
        return rc;
    }
    return GFT_OK;
  }
  const size_t need_ws = gft_render_density_bwd_workspace(shape, num_cameras, output_shape, boundary);
  if (need_ws && (!workspace || workspace_bytes < need_ws))
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                need_ws);
  if (boundary == GFT_CLAMP || boundary == GFT_BORDER) {
    const int bm = boundary == GFT_CLAMP ? B_CLAMP : B_BORDER;
    if ((rc = check_hip(quad_copy_launch(s, bm, density, (float*)workspace, rd.in), "quad copy"))) return rc;
    density = (const float*)workspace;
  }
  if (l2 && (rc = check_hip(fill_launch(s, loss, 1, 0.f), "clear loss"))) return rc;
  if ((rc = check_hip(fill_launch(s, density_grad, rd.in_batch_stride, 0.f), "clear density_grad"))) return rc;
  for (size_t first = 0; first < pairs.size(); first += kMaxViewsPerLaunch) {
    size_t cnt = std::min(pairs.size() - first, (size_t)kMaxViewsPerLaunch);
    RayLaunch L;
    if ((rc = ray_views(pairs, first, cnt, M, V, P, L))) return rc;
    if ((rc = check_hip(render_density_scatter_launch(boundary, l2, s, density, images, second, density_grad, loss, rd,
                                                      L, rp),
                        "render backward")))
      return rc;
  }
  return GFT_OK;
}

int gft_render_density_bwd_l2(void* stream, const float* density, const int64_t shape[5], const float* M,
                              int32_t num_M, const float* V, const float* P, const float* frustum,
                              int32_t num_cameras, const int32_t output_shape[3], const float light[3],
                              float depth_step, int32_t boundary, const float* images, const float* targets,
                              float loss_scale, void* saved, size_t saved_bytes, void* workspace,
                              size_t workspace_bytes, float* density_grad, float* loss) {
  (void)frustum;
  return render_bwd_common(stream, density, shape, M, num_M, V, P, num_cameras, output_shape, light, depth_step,
                           boundary, images, targets, true, loss_scale, GFT_LOSS_SE, nullptr, saved, saved_bytes,
                           workspace, workspace_bytes, density_grad, loss);
}

int gft_render_density_bwd_loss(void* stream, const float* density, const int64_t shape[5], const float* M,
                                int32_t num_M, const float* V, const float* P, const float* frustum,
                                int32_t num_cameras, const int32_t output_shape[3], const float light[3],
                                float depth_step, int32_t boundary, const float* images, const float* targets,
                                const float* backgrounds, int32_t loss_kind, float loss_scale, void* saved,
                                size_t saved_bytes, void* workspace, size_t workspace_bytes, float* density_grad,
                                float* loss) {
  (void)frustum;
  return render_bwd_common(stream, density, shape, M, num_M, V, P, num_cameras, output_shape, light, depth_step,
                           boundary, images, targets, true, loss_scale, loss_kind, backgrounds, saved, saved_bytes,
                           workspace, workspace_bytes, density_grad, loss);
}

int gft_render_density_bwd(void* stream, const float* density, const int64_t shape[5], const float* M, int32_t num_M,
                           const float* V, const float* P, const float* frustum, int32_t num_cameras,
                           const int32_t output_shape[3], const float light[3], float depth_step, int32_t boundary,
                           const float* images, const float* images_grad, void* saved, size_t saved_bytes,
                           void* workspace, size_t workspace_bytes, float* density_grad) {
  (void)frustum;
  return render_bwd_common(stream, density, shape, M, num_M, V, P, num_cameras, output_shape, light, depth_step,
                           boundary, images, images_grad, false, 0.f, GFT_LOSS_SE, nullptr, saved, saved_bytes,
                           workspace, workspace_bytes, density_grad, nullptr);
}

// ------------------------------------------------------------- spot-lit render
size_t gft_render_density_lit_fwd_workspace(const int64_t shape[5], int32_t boundary) {
  return boundary == GFT_BORDER ? 2 * quad_bytes(shape) : 0;
}

size_t gft_render_density_lit_saved_bytes(const int64_t shape[5], int32_t num_cameras, const int32_t output_shape[3],
                                          int32_t boundary) {
  if (boundary != GFT_BORDER) return 0;
  const size_t nsamples = render_waves_per_view(output_shape[2], output_shape[1]) * kSampleRun * output_shape[0];
  return gft_render_density_saved_bytes(shape, num_cameras, output_shape, boundary) +
         align256(nsamples * num_cameras * sizeof(float));
}

// the lit saved buffer: the ambient render's layout followed by the u_i planes
static RenderSaved saved_parts_lit(void* saved, const int64_t* shape, const int32_t* output_shape,
                                   int32_t num_cameras) {
  RenderSaved sv = saved_parts(saved, shape, output_shape, num_cameras);
  sv.samples2 = (float*)((char*)saved + gft_render_density_saved_bytes(shape, num_cameras, output_shape, GFT_BORDER));
  return sv;
}

static int check_lit_args(const int64_t* shape, int32_t num_M, int32_t num_cameras, const int32_t* output_shape,
                          int32_t boundary, const float* color, float ambient) {
  int rc;
  if ((rc = check_render_args(shape, num_M, num_cameras, output_shape, boundary))) return rc;
  if (boundary != GFT_BORDER)
    return fail(GFT_ERR_UNIMPLEMENTED, "the lit density render supports the BORDER boundary only");
  if (!color) return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  if (!std::isfinite(ambient)) return fail(GFT_ERR_INVALID_ARGUMENT, "ambient must be finite");
  return GFT_OK;
}

int gft_render_density_lit_fwd(void* stream, const float* density, const float* light_factor,
                               const int64_t shape[5], const float* M, int32_t num_M, const float* V,
                               const float* P, const float* frustum, int32_t num_cameras,
                               const int32_t output_shape[3], const float light_color[3], float ambient,
                               float depth_step, int32_t boundary, void* workspace, size_t workspace_bytes,
                               void* saved, size_t saved_bytes, float* images) {
  (void)frustum;
  int rc;
  if ((rc = check_lit_args(shape, num_M, num_cameras, output_shape, boundary, light_color, ambient))) return rc;
  if (!density || !light_factor || !images || !M || !V || !P)
    return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  const size_t need = gft_render_density_lit_fwd_workspace(shape, boundary);
  if (!workspace || workspace_bytes < need)
    return fail(GFT_ERR_WORKSPACE, "workspace of %zu bytes is smaller than the %zu bytes required", workspace_bytes,
                need);
  const size_t need_saved = gft_render_density_lit_saved_bytes(shape, num_cameras, output_shape, boundary);
  if (saved && saved_bytes < need_saved)
    return fail(GFT_ERR_WORKSPACE, "saved buffer of %zu bytes is smaller than the %zu bytes required", saved_bytes,
                need_saved);
  RenderSaved sv{nullptr, nullptr, nullptr, nullptr};
  if (saved) sv = saved_parts_lit(saved, shape, output_shape, num_cameras);
  hipStream_t s = (hipStream_t)stream;
  const RayDims rd = render_dims(shape, output_shape);
  RenderParams rp{{light_color[0], light_color[1], light_color[2]}, depth_step, 0.f};
  rp.ambient = ambient;
  float* dq = (float*)workspace;
  float* eq = (float*)((char*)workspace + quad_bytes(shape));
  if ((rc = check_hip(quad_copy_launch(s, B_BORDER, density, dq, rd.in), "quad copy"))) return rc;
  if ((rc = check_hip(quad_copy_launch(s, B_BORDER, density, eq, rd.in, light_factor), "lit quad copy"))) return rc;
  auto pairs = view_pairs(1, num_cameras, true);
  for (size_t first = 0; first < pairs.size(); first += kMaxViewsPerLaunch) {
    size_t cnt = std::min(pairs.size() - first, (size_t)kMaxViewsPerLaunch);
    RayLaunch L;
    if ((rc = ray_views(pairs, first, cnt, M, V, P, L))) return rc;
    if ((rc = check_hip(render_density_lit_fwd_launch(s, dq, eq, images, rd, L, rp, sv), "lit render forward")))
      return rc;
  }
  return GFT_OK;
}

static int render_lit_bwd_common(void* stream, const float* density, const float* light_factor,
                                 const int64_t shape[5], const float* M, int32_t num_M, const float* V,
                                 const float* P, int32_t num_cameras, const int32_t output_shape[3],
                                 const float light_color[3], float ambient, float depth_step, int32_t boundary,
                                 const float* images, const float* second, bool l2, float loss_scale,
                                 int32_t loss_kind, const float* backgrounds, void* saved, size_t saved_bytes,
                                 float* density_grad, float* light_factor_grad, float* loss) {
  int rc;
  if ((rc = check_lit_args(shape, num_M, num_cameras, output_shape, boundary, light_color, ambient))) return rc;
  if (loss_kind != GFT_LOSS_SE && loss_kind != GFT_LOSS_AE) return fail(GFT_ERR_INVALID_ARGUMENT, "invalid loss_kind");
  if (!density || !light_factor || !images || !second || !density_grad || !light_factor_grad || (l2 && !loss) || !M ||
      !V || !P)
    return fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  const size_t need = gft_render_density_lit_saved_bytes(shape, num_cameras, output_shape, boundary);
  if (!saved || saved_bytes < need)
    return fail(GFT_ERR_WORKSPACE, "saved buffer of %zu bytes is smaller than the %zu bytes required", saved_bytes,
                need);
  hipStream_t s = (hipStream_t)stream;
  const RayDims rd = render_dims(shape, output_shape);
  RenderParams rp{{light_color[0], light_color[1], light_color[2]}, depth_step, loss_scale, loss_kind, backgrounds};
  rp.ambient = ambient;
  const RenderSaved sv = saved_parts_lit(saved, shape, output_shape, num_cameras);
  auto pairs = view_pairs(1, num_cameras, true);
  for (size_t first = 0; first < pairs.size(); first += kMaxViewsPerLaunch) {
    size_t cnt = std::min(pairs.size() - first, (size_t)kMaxViewsPerLaunch);
    RayLaunch L;
    ProjLaunch PL;
    if ((rc = ray_views(pairs, first, cnt, M, V, P, L, &PL, &rd))) return rc;
    PL.accumulate = first > 0;
    if ((rc = check_hip(render_density_pixel_launch(l2, s, images, second, sv, rd, L, rp),
                        "lit render backward pixel pass")))
      return rc;
    const bool lastc = first + cnt == pairs.size();
    if ((rc = check_hip(render_density_lit_gather_launch(s, sv, density, light_factor, density_grad,
                                                         light_factor_grad, rd, PL, (l2 && lastc) ? loss : nullptr,
                                                         loss_scale),
                        "lit render backward gather")))
      return rc;
  }
  return GFT_OK;
}

int gft_render_density_lit_bwd_loss(void* stream, const float* density, const float* light_factor,
                                    const int64_t shape[5], const float* M, int32_t num_M, const float* V,
                                    const float* P, const float* frustum, int32_t num_cameras,
                                    const int32_t output_shape[3], const float light_color[3], float ambient,
                                    float depth_step, int32_t boundary, const float* images, const float* targets,
                                    const float* backgrounds, int32_t loss_kind, float loss_scale, void* saved,
                                    size_t saved_bytes, float* density_grad, float* light_factor_grad, float* loss) {
  (void)frustum;
  return render_lit_bwd_common(stream, density, light_factor, shape, M, num_M, V, P, num_cameras, output_shape,
                               light_color, ambient, depth_step, boundary, images, targets, true, loss_scale,
                               loss_kind, backgrounds, saved, saved_bytes, density_grad, light_factor_grad, loss);
}

int gft_render_density_lit_bwd(void* stream, const float* density, const float* light_factor,
                               const int64_t shape[5], const float* M, int32_t num_M, const float* V,
                               const float* P, const float* frustum, int32_t num_cameras,
                               const int32_t output_shape[3], const float light_color[3], float ambient,
                               float depth_step, int32_t boundary, const float* images, const float* images_grad,
                               void* saved, size_t saved_bytes, float* density_grad, float* light_factor_grad) {
  (void)frustum;
  return render_lit_bwd_common(stream, density, light_factor, shape, M, num_M, V, P, num_cameras, output_shape,
                               light_color, ambient, depth_step, boundary, images, images_grad, false, 0.f,
                               GFT_LOSS_SE, nullptr, saved, saved_bytes, density_grad, light_factor_grad, nullptr);
}

}  // extern "C"
