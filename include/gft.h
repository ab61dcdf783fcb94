/*
 * gft: C ABI of the MI355X (gfx950) rendering/transport kernels.
 *
 * Every entry point takes device pointers for grids and images, host pointers
 * for matrices, frustum parameters and shapes (the reference passes these as
 * HostMemory op inputs), a hipStream_t passed as void*, and returns a
 * gft_status. On failure gft_last_error() returns a thread-local message.
 * Nothing calls exit(); there is no global device state; calls are
 * stream-ordered and re-entrant. Scratch memory is supplied by the caller
 * through (workspace, workspace_bytes) sized by the matching *_workspace()
 * query.
 *
 * Layouts follow the reference: grids are NDHWC float32 with C in {1,2,4};
 * matrices are 4x4 float32 in column-major storage (the reference's
 * numpy ``.transpose()`` of row-major matrices, renderer.py:108-112);
 * frustum parameters are [near, far, left, right, top, bottom]
 * (camera.py:61-62).
 */
#ifndef GFT_H
#define GFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GFT_OK = 0,
  GFT_ERR_INVALID_ARGUMENT = 1, /* TF InvalidArgument in the reference */
  GFT_ERR_UNIMPLEMENTED = 2,    /* TF Unimplemented */
  GFT_ERR_RUNTIME = 3,          /* HIP runtime failure (reference: exit(10)) */
  GFT_ERR_WORKSPACE = 4         /* workspace smaller than the query returned */
} gft_status;

/* op attribute enums, same spelling as the reference's string attrs */
enum { GFT_NEAREST = 0, GFT_LINEAR = 1, GFT_MIN = 2, GFT_MAX = 3 };             /* interpolation */
enum { GFT_BORDER = 0, GFT_CLAMP = 1, GFT_WRAP = 2 };                          /* boundary */
enum { GFT_BEER_LAMBERT = 0, GFT_ALPHA = 1, GFT_ALPHA_ADDITIVE = 2, GFT_ADDITIVE = 3 }; /* blending_mode */
enum { GFT_MIP_NONE = 0, GFT_MIP_NEAREST = 1, GFT_MIP_LINEAR = 2 };             /* mipmapping */
enum {
  GFT_TRANSFORM_LINDEPTH = 0,
  GFT_TRANSFORM_LINDEPTH_REVERSE = 1,
  GFT_TRANSFORM = 2,
  GFT_TRANSFORM_REVERSE = 3
}; /* coordinate_mode */
enum { GFT_GRAD_REFERENCE = 0, GFT_GRAD_EXACT = 1 }; /* gradient flavour */
enum { GFT_LOSS_SE = 0, GFT_LOSS_AE = 1 };           /* folded image loss: squared / absolute error */

const char* gft_last_error(void);
int gft_abi_version(void);

/*
 * RaymarchGridTransform forward.
 * Replaces RaymarchGridTransformOp::Compute (phitest/render/cuda/src/
 * raymarch_grid.cc:192-297) and RaymarchGridKernel<GPUDevice,T,C>
 * (raymarch_grid.hpp:12-20, raymarch_grid.cu.cc:222-312).
 *   input        [N,D,H,W,C] device
 *   M            [num_M,4,4] host, num_M == N
 *   V, P         [num_cameras,4,4] host; frustum [num_cameras,6] host
 *   output_shape DHW (D = samples per ray)
 *   output       keep_dims: [N,V',1,H,W,C]; else [1,V',H,W,C] (N must be 1)
 *                with V' = num_cameras if separate_camera_batch else 1.
 */
int gft_raymarch_grid_transform_fwd(void* stream, const float* input, const int64_t input_shape[5], const float* M,
                                    int32_t num_M, const float* V, const float* P, const float* frustum,
                                    int32_t num_cameras, const int32_t output_shape[3], int32_t interpolation,
                                    int32_t boundary, int32_t blending_mode, int32_t keep_dims,
                                    int32_t separate_camera_batch, float* output);

/*
 * RaymarchGridTransformGrad.
 * Replaces RaymarchGridTransformGradOp::Compute (raymarch_grid.cc:347-481) and
 * RaymarchGridGradKernel (raymarch_grid.hpp:45-53, raymarch_grid.cu.cc:451-594).
 * interpolation must be LINEAR (the reference's grad op attr, :66).
 * grad_mode REFERENCE reproduces the per-camera weight normalisation;
 * EXACT is the true adjoint. input_grad is fully overwritten. BORDER runs
 * without atomics (a walk stores each sample's gradient, a voxel-block
 * gather sums them in a fixed order) and needs per-sample workspace for up to
 * 16 views of one batch; CLAMP/WRAP scatter with atomics.
 */
size_t gft_raymarch_grid_transform_bwd_workspace(const int64_t input_shape[5], int32_t num_cameras,
                                                 const int32_t output_shape[3], int32_t boundary,
                                                 int32_t separate_camera_batch, int32_t grad_mode);
int gft_raymarch_grid_transform_bwd(void* stream, const float* input, const int64_t input_shape[5],
                                    const float* output, const float* output_grad, const float* M, int32_t num_M,
                                    const float* V, const float* P, const float* frustum, int32_t num_cameras,
                                    const int32_t output_shape[3], int32_t interpolation, int32_t boundary,
                                    int32_t blending_mode, int32_t keep_dims, int32_t separate_camera_batch,
                                    int32_t grad_mode, void* workspace, size_t workspace_bytes, float* input_grad);

/*
 * Fused advection of a scalar grid [1,D,H,W,1] by a centred velocity
 * [1,D,H,W,3] (cells per unit time): semi-Lagrangian (order 1) or MacCormack
 * (order 2) with clamp 0 = NONE, 1 = MC_SMOOTH, 2 = NEGATIVE, CLAMP boundary,
 * cell-centre offset 0. Replaces the SampleGridLut composition of
 * VelocityGrid.warp (phitest/render/data_structures.py:870-905) over
 * SampleGridLuTKernelLauncher (resample_grid.cu.cc:498-660).
 * The forward writes the intermediates the backward needs into `saved`
 * (gft_advect_saved_bytes). The backward is the exact adjoint w.r.t. the
 * scalar and the velocity; input_grad and velocity_grad are overwritten.
 */
size_t gft_advect_saved_bytes(const int64_t shape[5]);
int gft_advect_fwd(void* stream, const float* input, const int64_t shape[5], const float* velocity, float dt,
                   int32_t order, int32_t clamp, float* output, void* saved, size_t saved_bytes);
size_t gft_advect_bwd_workspace(const int64_t shape[5]);
int gft_advect_bwd(void* stream, const float* input, const int64_t shape[5], const float* velocity, float dt,
                   int32_t order, int32_t clamp, const void* saved, const float* output_grad, float* input_grad,
                   float* velocity_grad, void* workspace, size_t workspace_bytes);
/*
 * The same advection with the scalar and the velocity stored in bfloat16
 * (uint16_t bit patterns, the upper half of an IEEE float): values are widened
 * to float on load and everything else (the output, `saved`, the gradients)
 * is float, as in gft_advect_fwd / gft_advect_bwd. Same buffer sizes.
 */
int gft_advect_fwd_bf16(void* stream, const uint16_t* input, const int64_t shape[5], const uint16_t* velocity,
                        float dt, int32_t order, int32_t clamp, float* output, void* saved, size_t saved_bytes);
int gft_advect_bwd_bf16(void* stream, const uint16_t* input, const int64_t shape[5], const uint16_t* velocity,
                        float dt, int32_t order, int32_t clamp, const void* saved, const float* output_grad,
                        float* input_grad, float* velocity_grad, void* workspace, size_t workspace_bytes);

/*
 * gft_advect_bwd without a workspace argument: the accumulator the MacCormack
 * adjoint needs (one float per voxel) lives in a device array inside the
 * library, GFT_ADVECT_NOALLOC_MAX_VOXELS floats (64 MiB) per GPU, held from
 * the library's load until the process ends. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc); the call makes no
 * allocation. Grids above the limit are rejected (GFT_ERR_INVALID_ARGUMENT).
 * Calls that overlap on one GPU share the array and race; it keeps the last
 * call's contents, and each call clears what it uses first.
 */
#define GFT_ADVECT_NOALLOC_MAX_VOXELS (1 << 24)
int gft_advect_bwd_noalloc(void* stream, const float* input, const int64_t shape[5], const float* velocity, float dt,
                           int32_t order, int32_t clamp, const void* saved, const float* output_grad,
                           float* input_grad, float* velocity_grad);

/*
 * Fused density render: BEER_LAMBERT raymarch of a density [1,D,H,W,1] lit
 * by an ambient light (rgb emission = light * density * depth_step, optical
 * depth = density * depth_step), images [1,V,H,W,4] = (r, g, b, optical
 * depth). Replaces Renderer.render_density with an ambient Light
 * (renderer.py:603-651, lighting.py:23-26) over raymarch_grid_transform
 * (raymarch_grid.cu.cc:107-143) without the materialised 4-channel grid.
 * BORDER and CLAMP read a corner-quad copy of the density (one float4 of x/y
 * corner neighbours per cell) that the call writes into the workspace
 * (gft_render_density_fwd_workspace; 0 bytes for WRAP). With BORDER a
 * non-null `saved` buffer (gft_render_density_saved_bytes) receives every
 * ray's sample range and per-sample extinction for the backward; it may be
 * null when no backward follows.
 */
size_t gft_render_density_fwd_workspace(const int64_t shape[5], int32_t boundary);
size_t gft_render_density_saved_bytes(const int64_t shape[5], int32_t num_cameras, const int32_t output_shape[3],
                                      int32_t boundary);
int gft_render_density_fwd(void* stream, const float* density, const int64_t shape[5], const float* M, int32_t num_M,
                           const float* V, const float* P, const float* frustum, int32_t num_cameras,
                           const int32_t output_shape[3], const float light[3], float depth_step, int32_t boundary,
                           void* workspace, size_t workspace_bytes, void* saved, size_t saved_bytes, float* images);
/*
 * gft_render_density_fwd with the density's corner-quad copy stored in
 * bfloat16 (rounded to nearest even; the workspace of
 * gft_render_density_fwd_workspace suffices): the render samples the density
 * at bfloat16 precision, as stored densities are in the bfloat16 precision
 * variant, with half the bytes per trilinear tap; interpolation, the
 * transmittance and E stay float32. BORDER only. The backward is
 * gft_render_density_bwd_l2 / _bwd_loss / _bwd with the `saved` buffer this
 * forward filled (the rounding is passed through as the identity).
 */
int gft_render_density_fwd_bf16(void* stream, const float* density, const int64_t shape[5], const float* M,
                                int32_t num_M, const float* V, const float* P, const float* frustum,
                                int32_t num_cameras, const int32_t output_shape[3], const float light[3],
                                float depth_step, int32_t boundary, void* workspace, size_t workspace_bytes,
                                void* saved, size_t saved_bytes, float* images);
/*
 * Backward of gft_render_density_fwd with the L2 image loss folded in:
 * loss = loss_scale * sum (rgb - target)^2 (targets [1,V,H,W,3]); writes the
 * scalar loss (device float) and the exact density gradient (overwritten).
 * Replaces loss_dens_target (optimization.py:396-412) + the raymarch gradient
 * (raymarch_grid.cu.cc:343-372). BORDER with the `saved` buffer that the
 * forward filled from the same density and cameras reduces the gradient with
 * a voxel-block gather (it rewrites the per-pixel records and loss partials
 * that follow the per-sample values in `saved`) and is deterministic. BORDER
 * with saved = NULL (the forward then needs none either), CLAMP and WRAP walk
 * every ray from far to near, recomputing the transmittance, and scatter with
 * an LDS pre-reduction and float atomics (summation order varies between
 * runs). The walk of BORDER and CLAMP needs a workspace of
 * gft_render_density_bwd_workspace bytes (the call writes the density's
 * corner-quad copy there); WRAP needs none.
 */
size_t gft_render_density_bwd_workspace(const int64_t shape[5], int32_t num_cameras, const int32_t output_shape[3],
                                        int32_t boundary);
int gft_render_density_bwd_l2(void* stream, const float* density, const int64_t shape[5], const float* M,
                              int32_t num_M, const float* V, const float* P, const float* frustum,
                              int32_t num_cameras, const int32_t output_shape[3], const float light[3],
                              float depth_step, int32_t boundary, const float* images, const float* targets,
                              float loss_scale, void* saved, size_t saved_bytes, void* workspace,
                              size_t workspace_bytes, float* density_grad, float* loss);
/*
 * As gft_render_density_bwd_l2 with the reference's image loss choices:
 * loss = loss_scale * sum_c f(rgb_c + bkg_c * exp(-depth) - target_c), f the
 * squared (GFT_LOSS_SE) or absolute (GFT_LOSS_AE, gradient sign(e), 0 at 0)
 * error. backgrounds [1,V,H,W,3] (device) or NULL: with a background the loss
 * is loss_dens_target_raw (optimization.py:414-430), without it
 * loss_dens_target (optimization.py:396-412); the reference's default loss
 * function for both is AE (optimization.py:114-116) and its mean over views
 * is loss_scale = scale / V.
 */
int gft_render_density_bwd_loss(void* stream, const float* density, const int64_t shape[5], const float* M,
                                int32_t num_M, const float* V, const float* P, const float* frustum,
                                int32_t num_cameras, const int32_t output_shape[3], const float light[3],
                                float depth_step, int32_t boundary, const float* images, const float* targets,
                                const float* backgrounds, int32_t loss_kind, float loss_scale, void* saved,
                                size_t saved_bytes, void* workspace, size_t workspace_bytes, float* density_grad,
                                float* loss);
/* Backward of gft_render_density_fwd from an image gradient [1,V,H,W,4]. */
int gft_render_density_bwd(void* stream, const float* density, const int64_t shape[5], const float* M, int32_t num_M,
                           const float* V, const float* P, const float* frustum, int32_t num_cameras,
                           const int32_t output_shape[3], const float light[3], float depth_step, int32_t boundary,
                           const float* images, const float* images_grad, void* saved, size_t saved_bytes,
                           void* workspace, size_t workspace_bytes, float* density_grad);

/*
 * Fused density render under a spot or point light plus ambient light:
 * Renderer.render_density with [SpotLight, Light] (renderer.py:573-651,
 * lighting.py:23-147) whose light grid is d * (c_spot lambda + c_ambient),
 * with the two colours proportional: light_color = c, c_spot = c * (the
 * intensity folded into lambda), c_ambient = c * ambient. light_factor
 * [1,D,H,W,1] (device) is lambda per cell: the spot intensity times the shadow
 * transmittance resampled to the grid (SpotLight._render_shadow_map) times the
 * PointLight falloff. The raymarch is the BEER_LAMBERT one of
 * gft_render_density_fwd without the materialised 4-channel grid: images
 * [1,V,H,W,4] = (c * E, optical depth). BORDER only (other boundaries return
 * GFT_ERR_UNIMPLEMENTED). The workspace (gft_render_density_lit_fwd_workspace)
 * receives the corner-quad copies of d and d * lambda; a non-null `saved`
 * (gft_render_density_lit_saved_bytes) receives two values per sample for the
 * backward, which is deterministic (no atomics) and writes the exact
 * gradients w.r.t. the density and the light factor (both overwritten); the
 * caller chains the latter through the shadow map and the falloff.
 */
size_t gft_render_density_lit_fwd_workspace(const int64_t shape[5], int32_t boundary);
size_t gft_render_density_lit_saved_bytes(const int64_t shape[5], int32_t num_cameras, const int32_t output_shape[3],
                                          int32_t boundary);
int gft_render_density_lit_fwd(void* stream, const float* density, const float* light_factor,
                               const int64_t shape[5], const float* M, int32_t num_M, const float* V,
                               const float* P, const float* frustum, int32_t num_cameras,
                               const int32_t output_shape[3], const float light_color[3], float ambient,
                               float depth_step, int32_t boundary, void* workspace, size_t workspace_bytes,
                               void* saved, size_t saved_bytes, float* images);
/* backward with the image loss folded in, as gft_render_density_bwd_loss */
int gft_render_density_lit_bwd_loss(void* stream, const float* density, const float* light_factor,
                                    const int64_t shape[5], const float* M, int32_t num_M, const float* V,
                                    const float* P, const float* frustum, int32_t num_cameras,
                                    const int32_t output_shape[3], const float light_color[3], float ambient,
                                    float depth_step, int32_t boundary, const float* images, const float* targets,
                                    const float* backgrounds, int32_t loss_kind, float loss_scale, void* saved,
                                    size_t saved_bytes, float* density_grad, float* light_factor_grad, float* loss);
/* backward from an image gradient [1,V,H,W,4] */
int gft_render_density_lit_bwd(void* stream, const float* density, const float* light_factor,
                               const int64_t shape[5], const float* M, int32_t num_M, const float* V,
                               const float* P, const float* frustum, int32_t num_cameras,
                               const int32_t output_shape[3], const float light_color[3], float ambient,
                               float depth_step, int32_t boundary, const float* images, const float* images_grad,
                               void* saved, size_t saved_bytes, float* density_grad, float* light_factor_grad);

/*
 * Mip pyramid of a grid batch [N,D,H,W,C]: levels 1..L (L = num_mipmaps,
 * reduced while a level would have an empty dimension) of pairwise means,
 * level m dims = base dims >> m. Layout: per batch a 256-byte aligned block
 * holding levels 1..L back to back, each NDHWC. Replaces MipAtlas3D and
 * kGenerateMip3D (resample_grid.cu.cc:46-236, setupMipAtlas
 * resample_grid.cc:201-220) without the per-call pinned pointer table.
 */
size_t gft_mip_atlas_bytes(const int64_t shape[5], int32_t num_mipmaps);
int gft_build_mips(void* stream, const float* input, const int64_t shape[5], int32_t num_mipmaps, float* atlas,
                   size_t atlas_bytes);

/*
 * SampleGridTransform forward. Replaces SampleGridTransformOp::Compute
 * (resample_grid.cc:268-447) and SampleGridKernelLauncher / kSample3D
 * (resample_grid.cu.cc:240-438).
 *   input  [N,D,H,W,C]; output [N,V',Do,Ho,Wo,C], V' = num_cameras if
 *   separate_camera_batch else 1 (camera i samples batch i).
 *   interpolation NEAREST|LINEAR, boundary BORDER|CLAMP, coordinate_mode one
 *   of the four transform modes, cell_center_offset must be 0.5.
 *   The workspace (gft_sample_grid_transform_workspace) holds the mip atlas.
 */
size_t gft_sample_grid_transform_workspace(const int64_t input_shape[5], int32_t mipmapping, int32_t num_mipmaps);
int gft_sample_grid_transform_fwd(void* stream, const float* input, const int64_t input_shape[5], const float* M,
                                  int32_t num_M, const float* V, const float* P, const float* frustum,
                                  int32_t num_cameras, const int32_t output_shape[3], int32_t interpolation,
                                  int32_t boundary, int32_t mipmapping, int32_t num_mipmaps, float mip_bias,
                                  int32_t coordinate_mode, float cell_center_offset, int32_t separate_camera_batch,
                                  void* workspace, size_t workspace_bytes, float* output);

/*
 * SampleGridTransformGrad. Replaces SampleGridTransformGradOp::Compute
 * (resample_grid.cc:469-625) and SampleGridGradKernelLauncher
 * (resample_grid.cu.cc:761-1018). grad_mode REFERENCE: trilinear scatter
 * with a per-voxel counter and division by the count per data batch
 * (mipmapping must be NONE, as the reference op registers); EXACT: the true
 * adjoint of the forward, mip levels included. input_grad is overwritten.
 * BORDER without mips in the forward coordinate modes runs in gather form
 * (per voxel block, no global atomics; REFERENCE when a batch has at most 8
 * views); the other cases scatter with atomics. The workspace query covers
 * both.
 */
size_t gft_sample_grid_transform_bwd_workspace(const int64_t input_shape[5], int32_t mipmapping, int32_t num_mipmaps,
                                               int32_t grad_mode);
int gft_sample_grid_transform_bwd(void* stream, const float* input, const int64_t input_shape[5],
                                  const float* output_grad, const float* M, int32_t num_M, const float* V,
                                  const float* P, const float* frustum, int32_t num_cameras,
                                  const int32_t output_shape[3], int32_t interpolation, int32_t boundary,
                                  int32_t mipmapping, int32_t num_mipmaps, float mip_bias, int32_t coordinate_mode,
                                  float cell_center_offset, int32_t separate_camera_batch, int32_t grad_mode,
                                  void* workspace, size_t workspace_bytes, float* input_grad);

/*
 * Gather gradient of the renderer's sampling: replaces the custom gradient
 * of Renderer._sample_transform / _sample_camera_transform
 * (renderer.py:200-227, 246-272), i.e. per data batch i
 *   input_grad[i] = sum_j sample_grid_transform(output_grad[i, j], M[i],
 *                   camera j, reverse coordinate mode, mips on output_grad)
 * fused into one pass with the sum over views. output_grad
 * [N,V,Do,Ho,Wo,C]; input_grad [N,D,H,W,C] with DHW = input_dhw;
 * coordinate_mode is the forward op's mode; mip_bias already includes the
 * gradient bias add.
 */
size_t gft_sample_grid_transform_gather_grad_workspace(const int64_t output_grad_shape[6], int32_t mipmapping,
                                                       int32_t num_mipmaps);
int gft_sample_grid_transform_gather_grad(void* stream, const float* output_grad, const int64_t output_grad_shape[6],
                                          const int32_t input_dhw[3], const float* M, int32_t num_M, const float* V,
                                          const float* P, const float* frustum, int32_t num_cameras,
                                          int32_t interpolation, int32_t boundary, int32_t mipmapping,
                                          int32_t num_mipmaps, float mip_bias, int32_t coordinate_mode,
                                          void* workspace, size_t workspace_bytes, float* input_grad);

/*
 * SampleGridLut forward. Replaces SampleGridLutOp::Compute
 * (resample_grid.cc:650-802) and SampleGridLuTKernelLauncher /
 * kSampleGrid3DLuT (resample_grid.cu.cc:465-660).
 *   lookup [V,Do,Ho,Wo,4] = (x, y, z, lod); position = xyz (* out dims if
 *   normalized) (+ cell index if relative). interpolation NEAREST|LINEAR|
 *   MIN|MAX, boundary BORDER|CLAMP|WRAP. output [N,V',Do,Ho,Wo,C].
 */
size_t gft_sample_grid_lut_workspace(const int64_t input_shape[5], int32_t mipmapping, int32_t num_mipmaps);
int gft_sample_grid_lut_fwd(void* stream, const float* input, const int64_t input_shape[5], const float* lookup,
                            const int64_t lookup_shape[5], int32_t interpolation, int32_t boundary, int32_t mipmapping,
                            int32_t num_mipmaps, float mip_bias, float cell_center_offset,
                            int32_t separate_camera_batch, int32_t relative_coords, int32_t normalized_coords,
                            void* workspace, size_t workspace_bytes, float* output);

/*
 * SampleGridLutGrad. Replaces SampleGridLutGradOp::Compute
 * (resample_grid.cc:828-979) and SampleGridLuTGradKernelLauncher /
 * kScatterLuT3DGradients (resample_grid.cu.cc:1051-1251). REFERENCE: lookup
 * gradient (LINEAR only) written per view, input gradient a counted
 * trilinear scatter for every filter; EXACT: lookup gradient summed over
 * batches (scaled for normalised coordinates), input gradient the adjoint of
 * the filter. mipmapping must be NONE. Both outputs are overwritten.
 */
size_t gft_sample_grid_lut_bwd_workspace(const int64_t input_shape[5], int32_t grad_mode);
int gft_sample_grid_lut_bwd(void* stream, const float* input, const int64_t input_shape[5], const float* lookup,
                            const int64_t lookup_shape[5], const float* output_grad, int32_t interpolation,
                            int32_t boundary, int32_t mipmapping, float cell_center_offset,
                            int32_t separate_camera_batch, int32_t relative_coords, int32_t normalized_coords,
                            int32_t grad_mode, void* workspace, size_t workspace_bytes, float* input_grad,
                            float* lookup_grad);

/*
 * LodTransform: per camera the sampling position and LoD of every output
 * cell, output [num_cameras,Do,Ho,Wo,4]. Replaces LodTransformOp::Compute
 * (resample_grid.cc:986-1105) and kComputeLoD3D (resample_grid.cu.cc:1288-1300).
 */
int gft_lod_transform(void* stream, const int32_t input_dhw[3], const float* MV, const float* P, const float* frustum,
                      int32_t num_cameras, const int32_t output_shape[3], int32_t inverse_transform, float* output);

/*
 * ReduceGridBlend / ReduceGridBlendGrad: blend along depth, output
 * [N,H,W,C] or every prefix [N,D,H,W,C] with keep_dims. Replace
 * ReduceGridBlendOp / ReduceGridBlendGradOp (reduce_blend.cc:60-242) and
 * kBlend / kBlendGrad (reduce_blend.cu.cc:34-227). REFERENCE rejects
 * ALPHA_ADDITIVE gradients as the reference launcher does.
 */
int gft_reduce_grid_blend_fwd(void* stream, const float* input, const int64_t shape[5], int32_t blending_mode,
                              int32_t keep_dims, float* output);
int gft_reduce_grid_blend_bwd(void* stream, const float* input, const int64_t shape[5], const float* output,
                              const float* output_grad, int32_t blending_mode, int32_t keep_dims, int32_t grad_mode,
                              float* input_grad);

/*
 * Fused spot-light shadow transmittance. Replaces the op sequence of
 * SpotLight._render_shadow_map up to the reverse resample (phitest/render/
 * lighting.py:106-141): sample_camera of the density into the light frustum
 * with the depth step on channel 0 (renderer.py:365-385, SampleGridTransform
 * TRANSFORM_LINDEPTH, resample_grid.cc:268-447), reduce_grid_blend
 * keep_dims (a running sum on one channel, reduce_blend.cc:60-124), the
 * one-cell depth shift padded by 0 (BORDER) or the first cell (CLAMP), then
 * exp(-D) (BEER_LAMBERT) or 1 - clip(D, 0, 1) (ALPHA), times the cone mask.
 *   density       [1,D,H,W,1] device; M, V, P [4,4] host; frustum [6] host
 *   shadow_shape  [Ds,Hs,Ws]; transmission [Ds,Hs,Ws] device
 *   mask          [Hs,Ws] device, or NULL for no cone mask
 *   workspace     gft_shadow_transmission_fwd_workspace() bytes (the mip
 *                 pyramid of density * depth_step; 0 without mips)
 *   saved         NULL, or gft_shadow_transmission_saved_bytes() of device
 *                 memory that receives the prefix sums D for the backward
 * interpolation NEAREST or LINEAR and boundary BORDER or CLAMP (the sampling
 * op's own limits); mipmapping, num_mipmaps and mip_bias as in
 * gft_sample_grid_transform_fwd.
 *
 * The backward is the reverse of the shift, transmittance and running sum:
 * frustum_grad [Ds,Hs,Ws] receives dL/d(step-scaled frustum sample), i.e. the
 * gradient the op sequence hands to SampleGridTransform's gradient; the
 * caller passes it, as a [1,1,Ds,Hs,Ws,1] output gradient, to
 * gft_sample_grid_transform_bwd or gft_sample_grid_transform_gather_grad
 * and multiplies the result by depth_step.
 */
size_t gft_shadow_transmission_saved_bytes(const int32_t shadow_shape[3]);
size_t gft_shadow_transmission_fwd_workspace(const int64_t density_shape[5], int32_t mipmapping,
                                             int32_t num_mipmaps);
int gft_shadow_transmission_fwd(void* stream, const float* density, const int64_t density_shape[5], const float* M,
                                const float* V, const float* P, const float* frustum, const int32_t shadow_shape[3],
                                float depth_step, int32_t interpolation, int32_t boundary, int32_t mipmapping,
                                int32_t num_mipmaps, float mip_bias, int32_t blending_mode, const float* mask,
                                void* workspace, size_t workspace_bytes, float* saved, float* transmission);
int gft_shadow_transmission_bwd(void* stream, const float* saved, const int32_t shadow_shape[3], int32_t boundary,
                                int32_t blending_mode, const float* mask, const float* transmission_grad,
                                float* frustum_grad);

/*
 * Divergence loss of a staggered velocity: replaces loss_vel_divergence
 * (optimization.py:824-838) over VelocityGrid.divergence
 * (data_structures.py:929-936) with its SE loss function:
 *   div = wx dx(vel_x) + wy dy(vel_y) + wz dz(vel_z) per centred cell,
 *   loss = loss_scale * sum div^2 (device float).
 * x [D,H,W+1], y [D,H+1,W], z [D+1,H,W] (device, one channel) for
 * centered_shape [D,H,W]; world_scale host [3] (x, y, z). The gradients
 * (all three or none) are overwritten, or added to when accumulate != 0.
 * Deterministic (no atomics).
 */
size_t gft_divergence_loss_workspace(const int32_t centered_shape[3]);
int gft_divergence_loss(void* stream, const float* x, const float* y, const float* z,
                        const int32_t centered_shape[3], const float world_scale[3], float loss_scale,
                        int32_t accumulate, void* workspace, size_t workspace_bytes, float* grad_x, float* grad_y,
                        float* grad_z, float* loss);

/*
 * One Adam update of up to 8 parameter tensors in a single launch, with the
 * update tf.train.AdamOptimizer documents (the reference's optimizer,
 * reconstruct_sequence.py:1899-1905): lr_t = lr sqrt(1-b2^t)/(1-b1^t),
 * m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2, p -= lr_t m / (sqrt(v) + eps).
 * Host arrays of num_tensors device pointers (params, grads, first and second
 * moments) and element counts; step = t >= 1, the update count including this
 * one. Parameters and moments are updated in place.
 */
int gft_adam_step(void* stream, int32_t num_tensors, float* const* params, const float* const* grads,
                  float* const* exp_avg, float* const* exp_avg_sq, const int64_t* sizes, float learning_rate,
                  float beta1, float beta2, float epsilon, int64_t step);
/*
 * As gft_adam_step with the update count on the device: t = *step_counter + 1
 * is read by the update kernel and the counter is incremented after it, so a
 * captured graph of the optimisation step advances t on every replay.
 * step_counter: one device int64, 0 before the first update.
 */
int gft_adam_step_device(void* stream, int32_t num_tensors, float* const* params, const float* const* grads,
                         float* const* exp_avg, float* const* exp_avg_sq, const int64_t* sizes, float learning_rate,
                         float beta1, float beta2, float epsilon, int64_t* step_counter);

/*
 * Batched density-reconstruction step of many independent reconstructions of
 * uneven grid sizes and view counts: per reconstruction the reference's
 * non-fused render of a density under an ambient light (Renderer.render_density
 * with fused=False, renderer.py:555-568,603-651: SampleGridTransform
 * TRANSFORM_LINDEPTH with LINEAR mipmapping, the LoD of every sample from its
 * frustum footprint, levels 0..num_mipmaps <= 3, resample_grid.cc:268-447 and
 * sampling.hpp:114-123; then _blend_grid BEER_LAMBERT, renderer.py:488-499),
 * the mean squared rgb error times loss_scale against the targets, and the
 * exact gradient w.r.t. the density through blend, resample and mips. Replaces
 * the per-(batch, camera) launch loops of SampleGridKernelLauncher
 * (resample_grid.cu.cc:372-436) and their gradients, for all reconstructions
 * at once: per mip level m one pass building the zero-bordered level m+1 and
 * level m's corner-quad copy (level 0 is the density itself), one persistent
 * forward launch whose resident workgroups take the (reconstruction, view,
 * 16x16 tile) items from a device work queue (rays clipped to the grid box;
 * it keeps each sample's c_k = exp(-D_k) + E_{k-1} and each pixel's gradient
 * record in the workspace), voxel-block gathers of every sample's gradient:
 * one per coarser level (3, then 2) for the views whose samples reach it,
 * then levels 0 and 1 together, whose store applies the mip adjoint (the
 * density gradient) and sums each reconstruction's loss. `brick` picks the
 * gather's voxel block per reconstruction: 0 = 16x16x2, 1 = 8x8x8,
 * 2 = 16x8x4. The workspace holds 4 bytes per frustum sample besides the
 * pyramids. No atomics in global memory besides the queue counter (zeroed
 * by the run's first launch): images, losses and gradients repeat bit for
 * bit. Grids with (H+2)(D+2) >= 2^24 are rejected. Runs that overlap on one
 * workspace race.
 * gft_batch_render_prepare writes the reconstruction table into the workspace
 * (a set-up call that waits for its copy); gft_batch_render_run only launches
 * kernels (capturable in a HIP graph) and reads the host array for launch sizes.
 */
typedef struct {
  int32_t grid_dhw[3];       /* D, H, W of the density */
  const float* density;      /* [1,D,H,W,1] device */
  const float* targets;      /* [V,H,W,3] device */
  float* images;             /* [V,H,W,4] device: (rgb, optical depth) */
  float* density_grad;       /* [1,D,H,W,1] device, overwritten */
  float* loss;               /* one device float */
  const float* M;            /* [4,4] host, column-major */
  const float* V;            /* [num_views,4,4] host */
  const float* P;            /* [num_views,4,4] host */
  const float* frustum;      /* [num_views,6] host */
  int32_t num_views;         /* 1..8 */
  int32_t output_shape[3];   /* S, H, W of every view */
  float depth_step;
  float light[3];            /* ambient colour * intensity */
  float loss_scale;
  int32_t brick;             /* 0, 1 or 2 (above) */
} gft_batch_recon;
size_t gft_batch_render_workspace(int32_t n, const gft_batch_recon* recons, int32_t num_mipmaps);
int gft_batch_render_prepare(void* stream, int32_t n, const gft_batch_recon* recons, int32_t num_mipmaps,
                             void* workspace, size_t workspace_bytes);
int gft_batch_render_run(void* stream, int32_t n, const gft_batch_recon* recons, int32_t num_mipmaps,
                         void* workspace, size_t workspace_bytes);

/*
 * Kernel timers (instrumentation; no counterpart in the reference, which
 * has none). gft_timer_start("a,b" or "*") makes every later launch of the
 * named kernels on the calling thread record a HIP event before and after
 * it on its stream; gft_timer_read synchronises those events and returns
 * the launch count and summed duration of one name; gft_timer_names lists
 * the names seen (returns the full length); gft_timer_stop frees the events.
 * Names: advect_sl, advect_mc, advect_o1_out, advect_bwd_a, advect_bwd_b,
 * render_fwd, render_bwd_pixel (with the folded loss sum), render_bwd_gather, render_bwd_scatter, quad_copy, fill,
 * render_lit_fwd, render_lit_bwd_gather, batch_levels, batch_fwd, batch_gather,
 * shadow_fwd, shadow_bwd, divergence, divergence_grad, adam, transform_grad_gather, lut_grad.
 */
int gft_timer_start(const char* kernels);
int gft_timer_read(const char* kernel, int32_t* launches, double* total_ms);
int gft_timer_names(char* buf, size_t buf_bytes);
void gft_timer_stop(void);

#ifdef __cplusplus
}
#endif

#endif /* GFT_H */
