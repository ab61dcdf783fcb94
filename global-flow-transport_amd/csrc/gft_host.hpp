// Host-side helpers of the C ABI: thread-local error reporting, shape
// validation in the reference's wording, and camera-record packing.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "gft_device.hpp"
#include "../../include/gft.h"

namespace gft {

void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);
int check_hip(hipError_t e, const char* what);

// Column-major 4x4 in double; m[4*col + row]
struct Mat4d {
  double m[16];
};
Mat4d mat_from_colmajor(const float* a);
Mat4d mat_identity();
Mat4d mat_mul(const Mat4d& a, const Mat4d& b);
bool mat_inverse(const Mat4d& a, Mat4d& out);
void mat_to_float(const Mat4d& a, float* out);

// The (batch, camera) pairs an op visits, in the reference's launch order
// (batch-major; all cameras per batch when separate_camera_batch, camera b for
// batch b otherwise).
struct ViewPair {
  int batch, cam, slot;
};
std::vector<ViewPair> view_pairs(int N, int num_cameras, bool separate_camera_batch);

// Per-view projective coefficients of the raymarch's pixel rays (near0, far0,
// gx, gy of ProjView) from the inverse projection and inverse model-view.
void ray_coefficients(const Mat4d& pinv, const Mat4d& mvinv, const RayDims& rd, ProjView& pv);

}  // namespace gft
