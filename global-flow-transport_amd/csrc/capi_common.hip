// Shared host code of the C ABI: errors, 4x4 algebra in double precision and
// the view-pair enumeration.
#include <algorithm>

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

// Homogeneous object-space point of the raymarch's index->object map at pixel
// (px, py) and depth index iz as g0 + px gx + py gy: ndc = (2 (i + 0.5) / n - 1)
// per axis, h = Pinv ndc, object = MVinv (h.xyz / h.w) = (A h).xyz / h.w with A
// the affine part of MVinv (its last row is not used, as in the device map).
void ray_coefficients(const Mat4d& pinv, const Mat4d& mvinv, const RayDims& rd, ProjView& pv) {
  const double W = rd.out_w, H = rd.out_h, S = rd.out_d;
  const double zmax = rd.out_d > 1 ? rd.out_d - 1 : 1;
  auto hcol = [&](int c, double* out) {
    for (int r = 0; r < 4; ++r) out[r] = pinv.m[4 * c + r];  // column-major
  };
  auto affine = [&](const double* h, float* out) {
    for (int r = 0; r < 3; ++r)
      out[r] = (float)(mvinv.m[r] * h[0] + mvinv.m[4 + r] * h[1] + mvinv.m[8 + r] * h[2] + mvinv.m[12 + r] * h[3]);
    out[3] = (float)h[3];
  };
  double c0[4], c1[4], c2[4], c3[4];
  hcol(0, c0);
  hcol(1, c1);
  hcol(2, c2);
  hcol(3, c3);
  double hx[4], hy[4], hn[4], hf[4];
  const double zn = 2.0 * 0.5 / S - 1.0, zf = 2.0 * (zmax + 0.5) / S - 1.0;
  for (int r = 0; r < 4; ++r) {
    hx[r] = c0[r] * (2.0 / W);
    hy[r] = c1[r] * (2.0 / H);
    const double base = c0[r] * (1.0 / W - 1.0) + c1[r] * (1.0 / H - 1.0) + c3[r];
    hn[r] = base + c2[r] * zn;
    hf[r] = base + c2[r] * zf;
  }
  affine(hx, pv.gx);
  affine(hy, pv.gy);
  affine(hn, pv.near0);
  affine(hf, pv.far0);
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

// ------------------------------------------------------------ kernel timers
// Per thread: the set of timed names ("*" = all) and, per name, the event
// pairs recorded since gft_timer_start.
struct TimerState {
  std::string names;
  bool all = false;
  std::vector<std::string> order;
  std::vector<std::vector<std::pair<hipEvent_t, hipEvent_t>>> events;
  std::vector<hipEvent_t> pending;
};
static thread_local TimerState g_timer;

static int timer_slot(const char* name) {
  for (size_t i = 0; i < g_timer.order.size(); ++i)
    if (g_timer.order[i] == name) return (int)i;
  g_timer.order.push_back(name);
  g_timer.events.emplace_back();
  g_timer.pending.push_back(nullptr);
  return (int)g_timer.order.size() - 1;
}

bool timer_active(const char* name) {
  if (g_timer.all) return true;
  if (g_timer.names.empty()) return false;
  const std::string n(name);
  size_t pos = 0;
  while (pos <= g_timer.names.size()) {
    size_t end = g_timer.names.find(',', pos);
    if (end == std::string::npos) end = g_timer.names.size();
    if (g_timer.names.compare(pos, end - pos, n) == 0) return true;
    pos = end + 1;
  }
  return false;
}

void timer_record(hipStream_t s, const char* name, bool begin) {
  int i = timer_slot(name);
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return;
  if (hipEventRecord(e, s) != hipSuccess) {
    (void)hipEventDestroy(e);
    return;
  }
  if (begin) {
    if (g_timer.pending[i]) (void)hipEventDestroy(g_timer.pending[i]);
    g_timer.pending[i] = e;
  } else if (g_timer.pending[i]) {
    g_timer.events[i].push_back({g_timer.pending[i], e});
    g_timer.pending[i] = nullptr;
  } else {
    (void)hipEventDestroy(e);
  }
}

static void timer_release() {
  for (auto& v : g_timer.events)
    for (auto& p : v) {
      (void)hipEventDestroy(p.first);
      (void)hipEventDestroy(p.second);
    }
  for (hipEvent_t e : g_timer.pending)
    if (e) (void)hipEventDestroy(e);
  g_timer = TimerState();
}

}  // namespace gft

extern "C" {

int gft_timer_start(const char* kernels) {
  gft::timer_release();
  if (!kernels) return gft::fail(GFT_ERR_INVALID_ARGUMENT, "null kernel list");
  gft::g_timer.names = kernels;
  gft::g_timer.all = gft::g_timer.names == "*";
  return GFT_OK;
}

int gft_timer_read(const char* kernel, int32_t* launches, double* total_ms) {
  if (!kernel || !launches || !total_ms) return gft::fail(GFT_ERR_INVALID_ARGUMENT, "null pointer argument");
  *launches = 0;
  *total_ms = 0.0;
  for (size_t i = 0; i < gft::g_timer.order.size(); ++i) {
    if (gft::g_timer.order[i] != kernel) continue;
    for (auto& p : gft::g_timer.events[i]) {
      int rc = gft::check_hip(hipEventSynchronize(p.second), "timer event");
      if (rc) return rc;
      float ms = 0.f;
      if ((rc = gft::check_hip(hipEventElapsedTime(&ms, p.first, p.second), "timer elapsed"))) return rc;
      *total_ms += ms;
      *launches += 1;
    }
  }
  return GFT_OK;
}

int gft_timer_names(char* buf, size_t buf_bytes) {
  std::string all;
  for (size_t i = 0; i < gft::g_timer.order.size(); ++i) {
    if (i) all += ",";
    all += gft::g_timer.order[i];
  }
  if (buf && buf_bytes) {
    size_t n = std::min(all.size(), buf_bytes - 1);
    all.copy(buf, n);
    buf[n] = 0;
  }
  return (int)all.size();
}

void gft_timer_stop(void) { gft::timer_release(); }


const char* gft_last_error(void) { return gft::g_last_error.c_str(); }

int gft_abi_version(void) { return 4; }

}  // extern "C"
