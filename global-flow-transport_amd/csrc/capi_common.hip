// Shared host code of the C ABI: errors, 4x4 algebra in double precision and
// the view-pair enumeration.
#include "gft_host.hpp"

namespace gft {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return GFT_OK;
  return fail(GFT_ERR_RUNTIME, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

Mat4d mat_from_colmajor(const float* a) {
  Mat4d r;
  for (int i = 0; i < 16; ++i) r.m[i] = a ? (double)a[i] : ((i % 5) == 0 ? 1.0 : 0.0);
  return r;
}
Mat4d mat_identity() { return mat_from_colmajor(nullptr); }

Mat4d mat_mul(const Mat4d& a, const Mat4d& b) {
  Mat4d r;
  for (int c = 0; c < 4; ++c)
    for (int row = 0; row < 4; ++row) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += a.m[4 * k + row] * b.m[4 * c + k];
      r.m[4 * c + row] = s;
    }
  return r;
}

// Gauss-Jordan elimination with partial pivoting in double precision.
bool mat_inverse(const Mat4d& a, Mat4d& out) {
  double w[4][8];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      w[r][c] = a.m[4 * c + r];
      w[r][c + 4] = (r == c) ? 1.0 : 0.0;
    }
  for (int col = 0; col < 4; ++col) {
    int piv = col;
    for (int r = col + 1; r < 4; ++r)
      if (std::fabs(w[r][col]) > std::fabs(w[piv][col])) piv = r;
    if (w[piv][col] == 0.0) return false;
    if (piv != col)
      for (int c = 0; c < 8; ++c) std::swap(w[piv][c], w[col][c]);
    double inv = 1.0 / w[col][col];
    for (int c = 0; c < 8; ++c) w[col][c] *= inv;
    for (int r = 0; r < 4; ++r) {
      if (r == col) continue;
      double f = w[r][col];
      if (f == 0.0) continue;
      for (int c = 0; c < 8; ++c) w[r][c] -= f * w[col][c];
    }
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) out.m[4 * c + r] = w[r][c + 4];
  return true;
}

void mat_to_float(const Mat4d& a, float* out) {
  for (int i = 0; i < 16; ++i) out[i] = (float)a.m[i];
}

std::vector<ViewPair> view_pairs(int N, int num_cameras, bool separate_camera_batch) {
  std::vector<ViewPair> v;
  for (int b = 0; b < N; ++b) {
    if (separate_camera_batch)
      for (int c = 0; c < num_cameras; ++c) v.push_back({b, c, b * num_cameras + c});
    else
      v.push_back({b, b, b});
  }
  return v;
}

}  // namespace gft

extern "C" {

const char* gft_last_error(void) { return gft::g_last_error.c_str(); }

int gft_abi_version(void) { return 1; }

}  // extern "C"
