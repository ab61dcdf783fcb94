// Oracle for the rendering/transport hot path: a plain C++ CPU restatement of
// the reference op semantics (tum-pbs/Global-Flow-Transport, phitest/render).
//
// TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the reference results or as the
// reported CPU baseline. The product path never links or calls it.
//
// Every routine is templated on the arithmetic type T (float = the reference's
// FP32 arithmetic, double = the FP64 oracle). Data arrays cross the C ABI as
// double (float inputs convert exactly); matrices arrive as float like the
// reference's HostMemory float32 inputs.
//
// Two gradient modes are restated:
//   REFERENCE: the reference's normalised gradients (per-batch hit count for
//              resampling and LuT, per-camera trilinear weight for raymarch),
//              including its BORDER band and its ALPHA blend-gradient quirk.
//   EXACT:     the true adjoint of the forward op (what finite differences see).
//
// Source files restated (paths relative to the reference root):
//   phitest/render/cuda/src/vectormath.hpp        column-major float4x4, inverse
//   phitest/render/cuda/src/transformations.hpp   v1 IDX<->OS incl. linear depth
//   phitest/render/cuda/src/transformations_v2.hpp v2 IDXtoOS (raymarch)
//   phitest/render/cuda/src/sampling.hpp          v1 sampler, LoD, mips, LuT
//   phitest/render/cuda/src/sampling_v2.hpp       v2 sampler + scatter
//   phitest/render/cuda/src/blending.hpp          blend recurrences and grads
//   phitest/render/cuda/src/resample_grid.cu.cc   mip gen, sample/scatter kernels
//   phitest/render/cuda/src/raymarch_grid.cu.cc   raymarch fwd/bwd kernels
//   phitest/render/cuda/src/reduce_blend.cu.cc    depth blend fwd/bwd
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

enum Filter { F_NEAREST = 0, F_LINEAR = 1, F_MIN = 2, F_MAX = 3 };
enum Boundary { B_BORDER = 0, B_CLAMP = 1, B_WRAP = 2 };
enum Blend { BL_BEER_LAMBERT = 0, BL_ALPHA = 1, BL_ALPHA_ADDITIVE = 2, BL_ADDITIVE = 3 };
enum Mip { MIP_NONE = 0, MIP_NEAREST = 1, MIP_LINEAR = 2 };
enum Coord { CM_LINDEPTH = 0, CM_LINDEPTH_REVERSE = 1, CM_TRANSFORM = 2, CM_TRANSFORM_REVERSE = 3 };
enum GradMode { GM_REFERENCE = 0, GM_EXACT = 1 };

// FLOAT_MAX of vectormath.hpp:12-14 (a double literal rounded to float there).
static const float kFloatMax = 3.40282e+38f;

template <class T> struct v3 { T x, y, z; };
template <class T> struct v4 { T x, y, z, w; };
template <class T> inline v3<T> mk3(T x, T y, T z) { return {x, y, z}; }
template <class T> inline v4<T> mk4(T x, T y, T z, T w) { return {x, y, z, w}; }
template <class T> inline v4<T> add(v4<T> a, v4<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
template <class T> inline v4<T> scl(v4<T> a, T s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
template <class T> inline v3<T> add3(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline v3<T> sub3(v3<T> a, v3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline v3<T> scl3(v3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline T len3(v3<T> a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// Column-major 4x4 (vectormath.hpp:285-312): 16 floats read as 4 columns.
template <class T> struct m4 { v4<T> c[4]; };
template <class T> m4<T> m4_from(const float* a) {
  m4<T> m;
  for (int j = 0; j < 4; ++j) m.c[j] = mk4<T>(a[4 * j], a[4 * j + 1], a[4 * j + 2], a[4 * j + 3]);
  return m;
}
template <class T> m4<T> m4_ident() {
  m4<T> m;
  for (int j = 0; j < 4; ++j) m.c[j] = mk4<T>(j == 0, j == 1, j == 2, j == 3);
  return m;
}
// matmul(m, v) = c0*v.x + c1*v.y + c2*v.z + c3*v.w (vectormath.hpp:411-413)
template <class T> v4<T> mul(const m4<T>& m, v4<T> v) {
  return add(add(add(scl(m.c[0], v.x), scl(m.c[1], v.y)), scl(m.c[2], v.z)), scl(m.c[3], v.w));
}
template <class T> m4<T> mul(const m4<T>& a, const m4<T>& b) {
  m4<T> r;
  for (int j = 0; j < 4; ++j) r.c[j] = mul(a, b.c[j]);
  return r;
}
template <class T> inline T el(const m4<T>& m, int col, int row) {
  const v4<T>& c = m.c[col];
  return row == 0 ? c.x : row == 1 ? c.y : row == 2 ? c.z : c.w;
}
// 4x4 inverse by the adjugate built from 2x2 sub-determinants, evaluated in T
// (the reference computes it on the host in FP32, vectormath.hpp:353-408 and
// glm::inverse in transformations.hpp:56-61).
template <class T> m4<T> inverse(const m4<T>& m) {
  T a[4][4];  // a[row][col]
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) a[r][c] = el(m, c, r);
  T s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
  T s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
  T s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
  T s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
  T s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
  T s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
  T c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
  T c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
  T c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
  T c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
  T c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
  T c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
  T det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  T id = T(1) / det;
  T b[4][4];
  b[0][0] = (a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3) * id;
  b[0][1] = (-a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3) * id;
  b[0][2] = (a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3) * id;
  b[0][3] = (-a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3) * id;
  b[1][0] = (-a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1) * id;
  b[1][1] = (a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1) * id;
  b[1][2] = (-a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1) * id;
  b[1][3] = (a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1) * id;
  b[2][0] = (a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0) * id;
  b[2][1] = (-a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0) * id;
  b[2][2] = (a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0) * id;
  b[2][3] = (-a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0) * id;
  b[3][0] = (-a[1][0] * c3 + a[1][1] * c1 - a[1][2] * c0) * id;
  b[3][1] = (a[0][0] * c3 - a[0][1] * c1 + a[0][2] * c0) * id;
  b[3][2] = (-a[3][0] * s3 + a[3][1] * s1 - a[3][2] * s0) * id;
  b[3][3] = (a[2][0] * s3 - a[2][1] * s1 + a[2][2] * s0) * id;
  m4<T> r;
  for (int c = 0; c < 4; ++c) r.c[c] = mk4<T>(b[0][c], b[1][c], b[2][c], b[3][c]);
  return r;
}

// Per (batch, camera) transformation set (transformations.hpp:44-64,
// transformations_v2.hpp:90-110): MV = V*M and the four inverses.
template <class T> struct Xf {
  m4<T> P, MV, Pinv, MVinv;
  T nearp, farp;
};
template <class T> Xf<T> make_xf(const float* M, const float* V, const float* P, const float* F) {
  Xf<T> x;
  m4<T> Mm = M ? m4_from<T>(M) : m4_ident<T>();
  m4<T> Vm = V ? m4_from<T>(V) : m4_ident<T>();
  x.P = m4_from<T>(P);
  x.MV = mul(Vm, Mm);
  x.Pinv = inverse(x.P);
  x.MVinv = inverse(x.MV);
  x.nearp = F[0];
  x.farp = F[1];
  return x;
}

// ---------------------------------------------------------------- transforms
// v2 (transformations_v2.hpp:166-179): idx -> ndc -> vs (divide by w) -> os
template <class T> v3<T> idx_to_os(const Xf<T>& x, v3<T> idx, v3<T> dinv) {
  v4<T> ndc = mk4<T>(idx.x * dinv.x * 2 - 1, idx.y * dinv.y * 2 - 1, idx.z * dinv.z * 2 - 1, 1);
  v4<T> vs = mul(x.Pinv, ndc);
  vs = mk4<T>(vs.x / vs.w, vs.y / vs.w, vs.z / vs.w, T(1));
  v4<T> os = mul(x.MVinv, vs);
  return mk3(os.x, os.y, os.z);
}
// v1 linear-depth variant (transformations.hpp:135-145)
template <class T> v3<T> idx_to_os_lindepth(const Xf<T>& x, v3<T> idx, v3<T> dinv) {
  v4<T> ndc = mk4<T>(idx.x * dinv.x * 2 - 1, idx.y * dinv.y * 2 - 1, idx.z * dinv.z * 2 - 1, 1);
  T t = idx.z * dinv.z;
  T z = -(x.nearp + (x.farp - x.nearp) * t);
  ndc.z = -el(x.P, 2, 2) - el(x.P, 3, 2) / z;
  v4<T> vs = mul(x.Pinv, ndc);
  vs = mk4<T>(vs.x / vs.w, vs.y / vs.w, vs.z / vs.w, T(1));
  v4<T> os = mul(x.MVinv, vs);
  return mk3(os.x, os.y, os.z);
}
// os -> idx (transformations.hpp:86-104)
template <class T> v3<T> os_to_idx(const Xf<T>& x, v3<T> os, v3<T> dims, T* ndcz = nullptr) {
  v4<T> vs = mul(x.MV, mk4<T>(os.x, os.y, os.z, 1));
  v4<T> ndc = mul(x.P, vs);
  ndc = mk4<T>(ndc.x / ndc.w, ndc.y / ndc.w, ndc.z / ndc.w, T(1));
  if (ndcz) *ndcz = ndc.z;
  return mk3<T>((ndc.x * T(0.5) + T(0.5)) * dims.x, (ndc.y * T(0.5) + T(0.5)) * dims.y,
                (ndc.z * T(0.5) + T(0.5)) * dims.z);
}
// os -> idx with linear depth (transformations.hpp:106-117)
template <class T> v3<T> os_to_idx_lindepth(const Xf<T>& x, v3<T> os, v3<T> dims) {
  T ndcz;
  v3<T> idx = os_to_idx(x, os, dims, &ndcz);
  T z = -el(x.P, 3, 2) / (ndcz + el(x.P, 2, 2));
  T t = -(z + x.nearp) / (x.farp - x.nearp);
  idx.z = t * dims.z;
  return idx;
}

// --------------------------------------------------------------- grid access
struct Dims {
  int x, y, z;
};
inline size_t flat3(int x, int y, int z, const Dims& d) { return (size_t(z) * d.y + y) * d.x + x; }
inline int pmod(int v, int m) {
  int r = v % m;
  return r < 0 ? r + m : r;
}
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <class T> inline T lerp(T a, T b, T t) { return a + t * (b - a); }
template <class T> inline T fract(T v) { return v - std::floor(v); }

// Trilinear read with cell centres at integer positions (sampling_v2.hpp:37-124,
// sampling.hpp:158-252; identical arithmetic in both versions).
template <class T>
void read_linear(const T* buf, const Dims& d, int C, int bm, T px, T py, T pz, T* out) {
  for (int c = 0; c < C; ++c) out[c] = 0;
  T wx = fract(px), wy = fract(py), wz = fract(pz);
  int cx = (int)std::ceil(px), cy = (int)std::ceil(py), cz = (int)std::ceil(pz);
  int fx = (int)std::floor(px), fy = (int)std::floor(py), fz = (int)std::floor(pz);
  bool vx[2] = {true, true}, vy[2] = {true, true}, vz[2] = {true, true};
  if (bm == B_BORDER) {
    T qx = px + T(0.5), qy = py + T(0.5), qz = pz + T(0.5);
    if (!(0 <= qx && qx < d.x && 0 <= qy && qy < d.y && 0 <= qz && qz < d.z)) return;
    vx[0] = -1 < fx && fx < d.x; vx[1] = -1 < cx && cx < d.x;
    vy[0] = -1 < fy && fy < d.y; vy[1] = -1 < cy && cy < d.y;
    vz[0] = -1 < fz && fz < d.z; vz[1] = -1 < cz && cz < d.z;
  } else if (bm == B_CLAMP) {
    fx = clampi(fx, 0, d.x - 1); cx = clampi(cx, 0, d.x - 1);
    fy = clampi(fy, 0, d.y - 1); cy = clampi(cy, 0, d.y - 1);
    fz = clampi(fz, 0, d.z - 1); cz = clampi(cz, 0, d.z - 1);
  } else {
    fx = pmod(fx, d.x); cx = pmod(cx, d.x);
    fy = pmod(fy, d.y); cy = pmod(cy, d.y);
    fz = pmod(fz, d.z); cz = pmod(cz, d.z);
  }
  int ix[2] = {fx, cx}, iy[2] = {fy, cy}, iz[2] = {fz, cz};
  for (int c = 0; c < C; ++c) {
    T v[2][2][2];
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i)
          v[k][j][i] = (vx[i] && vy[j] && vz[k]) ? buf[flat3(ix[i], iy[j], iz[k], d) * C + c] : T(0);
    T v00 = lerp(v[0][0][0], v[0][0][1], wx);
    T v01 = lerp(v[0][1][0], v[0][1][1], wx);
    T v0 = lerp(v00, v01, wy);
    T v10 = lerp(v[1][0][0], v[1][0][1], wx);
    T v11 = lerp(v[1][1][0], v[1][1][1], wx);
    T v1 = lerp(v10, v11, wy);
    out[c] = lerp(v0, v1, wz);
  }
}

// Nearest read: index = trunc(p + 0.5) (C cast), sampling_v2.hpp:125-140,
// sampling.hpp:356-371. Returns false (data 0) when BORDER rejects it.
inline bool nearest_index(int bm, const Dims& d, double px, double py, double pz, int& ix, int& iy, int& iz) {
  ix = (int)(px + 0.5); iy = (int)(py + 0.5); iz = (int)(pz + 0.5);
  if (bm == B_BORDER) {
    return 0 <= ix && ix < d.x && 0 <= iy && iy < d.y && 0 <= iz && iz < d.z;
  } else if (bm == B_CLAMP) {
    ix = clampi(ix, 0, d.x - 1); iy = clampi(iy, 0, d.y - 1); iz = clampi(iz, 0, d.z - 1);
  } else {
    ix = pmod(ix, d.x); iy = pmod(iy, d.y); iz = pmod(iz, d.z);
  }
  return true;
}
template <class T>
void read_nearest(const T* buf, const Dims& d, int C, int bm, T px, T py, T pz, T* out) {
  int ix, iy, iz;
  // the +0.5 and the truncation happen in T in the reference
  ix = (int)(px + T(0.5)); iy = (int)(py + T(0.5)); iz = (int)(pz + T(0.5));
  bool ok = true;
  if (bm == B_BORDER) ok = 0 <= ix && ix < d.x && 0 <= iy && iy < d.y && 0 <= iz && iz < d.z;
  else if (bm == B_CLAMP) { ix = clampi(ix, 0, d.x - 1); iy = clampi(iy, 0, d.y - 1); iz = clampi(iz, 0, d.z - 1); }
  else { ix = pmod(ix, d.x); iy = pmod(iy, d.y); iz = pmod(iz, d.z); }
  for (int c = 0; c < C; ++c) out[c] = ok ? buf[flat3(ix, iy, iz, d) * C + c] : T(0);
}

// v1 MIN/MAX of the 8 corners (sampling.hpp:372-491): BORDER outside -> 0, an
// invalid corner starts the extremum at 0, valid corners via fminf/fmaxf.
template <class T>
void read_minmax_v1(const T* buf, const Dims& d, int C, int bm, bool is_max, T px, T py, T pz, T* out,
                    int* arg = nullptr) {
  int cx = (int)std::ceil(px), cy = (int)std::ceil(py), cz = (int)std::ceil(pz);
  int fx = (int)std::floor(px), fy = (int)std::floor(py), fz = (int)std::floor(pz);
  bool vx[2] = {true, true}, vy[2] = {true, true}, vz[2] = {true, true};
  T init = is_max ? T(-kFloatMax) : T(kFloatMax);
  if (arg) for (int c = 0; c < C; ++c) arg[c] = -1;
  if (bm == B_BORDER) {
    T qx = px + T(0.5), qy = py + T(0.5), qz = pz + T(0.5);
    if (!(0 <= qx && qx < d.x && 0 <= qy && qy < d.y && 0 <= qz && qz < d.z)) {
      for (int c = 0; c < C; ++c) out[c] = 0;
      return;
    }
    vx[0] = -1 < fx && fx < d.x; vx[1] = -1 < cx && cx < d.x;
    vy[0] = -1 < fy && fy < d.y; vy[1] = -1 < cy && cy < d.y;
    vz[0] = -1 < fz && fz < d.z; vz[1] = -1 < cz && cz < d.z;
    if (!(vx[0] && vx[1] && vy[0] && vy[1] && vz[0] && vz[1])) init = 0;
  } else if (bm == B_CLAMP) {
    fx = clampi(fx, 0, d.x - 1); cx = clampi(cx, 0, d.x - 1);
    fy = clampi(fy, 0, d.y - 1); cy = clampi(cy, 0, d.y - 1);
    fz = clampi(fz, 0, d.z - 1); cz = clampi(cz, 0, d.z - 1);
  } else {
    fx = pmod(fx, d.x); cx = pmod(cx, d.x);
    fy = pmod(fy, d.y); cy = pmod(cy, d.y);
    fz = pmod(fz, d.z); cz = pmod(cz, d.z);
  }
  int ix[2] = {fx, cx}, iy[2] = {fy, cy}, iz[2] = {fz, cz};
  for (int c = 0; c < C; ++c) {
    T m = init;
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i) {
          if (!(vx[i] && vy[j] && vz[k])) continue;
          size_t f = flat3(ix[i], iy[j], iz[k], d);
          T v = buf[f * C + c];
          bool better = is_max ? (v > m) : (v < m);
          if (better) {
            m = v;
            if (arg) arg[c] = (int)f;
          }
        }
    out[c] = m;
  }
}

// v2 MIN/MAX with extrema tracked through lerp(ext, v, v<ext) starting from
// +-FLOAT_MAX (sampling_v2.hpp:149-253); reproduced literally.
template <class T>
void read_minmax_v2(const T* buf, const Dims& d, int C, int bm, bool is_max, T px, T py, T pz, T* out) {
  int cx = (int)std::ceil(px), cy = (int)std::ceil(py), cz = (int)std::ceil(pz);
  int fx = (int)std::floor(px), fy = (int)std::floor(py), fz = (int)std::floor(pz);
  bool vx[2] = {true, true}, vy[2] = {true, true}, vz[2] = {true, true};
  for (int c = 0; c < C; ++c) out[c] = is_max ? T(-kFloatMax) : T(kFloatMax);
  if (bm == B_BORDER) {
    T qx = px + T(0.5), qy = py + T(0.5), qz = pz + T(0.5);
    if (!(0 <= qx && qx < d.x && 0 <= qy && qy < d.y && 0 <= qz && qz < d.z)) return;
    vx[0] = -1 < fx && fx < d.x; vx[1] = -1 < cx && cx < d.x;
    vy[0] = -1 < fy && fy < d.y; vy[1] = -1 < cy && cy < d.y;
    vz[0] = -1 < fz && fz < d.z; vz[1] = -1 < cz && cz < d.z;
  } else if (bm == B_CLAMP) {
    fx = clampi(fx, 0, d.x - 1); cx = clampi(cx, 0, d.x - 1);
    fy = clampi(fy, 0, d.y - 1); cy = clampi(cy, 0, d.y - 1);
    fz = clampi(fz, 0, d.z - 1); cz = clampi(cz, 0, d.z - 1);
  } else {
    fx = pmod(fx, d.x); cx = pmod(cx, d.x);
    fy = pmod(fy, d.y); cy = pmod(cy, d.y);
    fz = pmod(fz, d.z); cz = pmod(cz, d.z);
  }
  int ix[2] = {fx, cx}, iy[2] = {fy, cy}, iz[2] = {fz, cz};
  // corner read order of the reference macro sequence: (f,f,f) (c,f,f) (f,c,f) (c,c,f) (f,f,c) ...
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        if (!(vx[i] && vy[j] && vz[k])) continue;
        size_t f = flat3(ix[i], iy[j], iz[k], d);
        for (int c = 0; c < C; ++c) {
          T v = buf[f * C + c];
          T t = is_max ? T(out[c] < v) : T(v < out[c]);
          out[c] = out[c] + t * (v - out[c]);
        }
      }
}

// v2 sample3D (sampling_v2.hpp:358-372): position has +0.5 cell-centre offset
template <class T>
void sample_v2(const T* buf, const Dims& d, int C, int fm, int bm, v3<T> p, T* out) {
  T px = p.x - T(0.5), py = p.y - T(0.5), pz = p.z - T(0.5);
  if (fm == F_LINEAR) read_linear(buf, d, C, bm, px, py, pz, out);
  else if (fm == F_NEAREST) read_nearest(buf, d, C, bm, px, py, pz, out);
  else read_minmax_v2(buf, d, C, bm, fm == F_MAX, px, py, pz, out);
}

// Trilinear scatter: weights of the 8 corners for position p (offset removed).
// mode REFERENCE: BORDER accepts p in (-1, dims) and checks each corner index
// (sampling_v2.hpp:459-486, resample_grid.cu.cc:761-812); EXACT: BORDER accepts
// only p+0.5 in [0, dims), matching the forward read.
template <class T, class F>
void scatter_linear(const Dims& d, int bm, bool exact, T px, T py, T pz, F&& add) {
  if (bm == B_BORDER) {
    if (exact) {
      T qx = px + T(0.5), qy = py + T(0.5), qz = pz + T(0.5);
      if (!(0 <= qx && qx < d.x && 0 <= qy && qy < d.y && 0 <= qz && qz < d.z)) return;
    } else {
      if (!(T(-1) < px && px < d.x && T(-1) < py && py < d.y && T(-1) < pz && pz < d.z)) return;
    }
  }
  T cwx = fract(px), cwy = fract(py), cwz = fract(pz);
  T fwx = 1 - cwx, fwy = 1 - cwy, fwz = 1 - cwz;
  int cx = (int)std::ceil(px), cy = (int)std::ceil(py), cz = (int)std::ceil(pz);
  int fx = (int)std::floor(px), fy = (int)std::floor(py), fz = (int)std::floor(pz);
  if (bm == B_CLAMP) {
    fx = clampi(fx, 0, d.x - 1); cx = clampi(cx, 0, d.x - 1);
    fy = clampi(fy, 0, d.y - 1); cy = clampi(cy, 0, d.y - 1);
    fz = clampi(fz, 0, d.z - 1); cz = clampi(cz, 0, d.z - 1);
  } else if (bm == B_WRAP) {
    fx = pmod(fx, d.x); cx = pmod(cx, d.x);
    fy = pmod(fy, d.y); cy = pmod(cy, d.y);
    fz = pmod(fz, d.z); cz = pmod(cz, d.z);
  }
  int ix[2] = {fx, cx}, iy[2] = {fy, cy}, iz[2] = {fz, cz};
  T wx[2] = {fwx, cwx}, wy[2] = {fwy, cwy}, wz[2] = {fwz, cwz};
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        if (bm == B_BORDER && !(0 <= ix[i] && ix[i] < d.x && 0 <= iy[j] && iy[j] < d.y && 0 <= iz[k] && iz[k] < d.z))
          continue;
        add(flat3(ix[i], iy[j], iz[k], d), wx[i] * wy[j] * wz[k]);
      }
}

// ------------------------------------------------------------------ blending
// Blend recurrences (blending.hpp:9-227). C==1 and ADDITIVE are plain sums.
template <class T> void blend(int mode, int C, T* acc, const T* c) {
  if (C == 1 || mode == BL_ADDITIVE) {
    for (int i = 0; i < C; ++i) acc[i] = acc[i] + c[i];
    return;
  }
  int a = C - 1;
  if (mode == BL_BEER_LAMBERT) {
    T dd = std::max(acc[a] + c[a], T(0));
    T t = std::exp(-dd);
    for (int i = 0; i < a; ++i) acc[i] = acc[i] + t * c[i];
    acc[a] = dd;
  } else if (mode == BL_ALPHA) {
    T t = 1 - acc[a];
    for (int i = 0; i < a; ++i) acc[i] = acc[i] + t * c[i];
    acc[a] = acc[a] + t * c[a];
  } else {  // ALPHA_ADDITIVE
    T t = 1 - c[a];
    for (int i = 0; i < a; ++i) acc[i] = acc[i] + t * c[i];
    acc[a] = acc[a] + c[a];
  }
}
// Reverse step: g = running output gradient (updated to the gradient w.r.t. the
// previous accumulator), acc = accumulator after this cell (updated to before).
// Writes the gradient w.r.t. the cell into gc. REFERENCE keeps the ALPHA quirk
// of blending.hpp:128-142; EXACT uses the true derivative.
template <class T> void blend_grad(int mode, bool exact, int C, T* g, const T* c, T* acc, T* gc) {
  if (C == 1 || mode == BL_ADDITIVE) {
    for (int i = 0; i < C; ++i) gc[i] = g[i];
    return;
  }
  int a = C - 1;
  T s = 0;
  for (int i = 0; i < a; ++i) s += c[i] * g[i];
  if (mode == BL_BEER_LAMBERT) {
    T e = std::exp(-acc[a]);
    acc[a] = std::max(acc[a] - c[a], T(0));
    T d_dacc = g[a] - e * s;
    for (int i = 0; i < a; ++i) gc[i] = e * g[i];
    gc[a] = d_dacc;
    g[a] = d_dacc;
  } else if (mode == BL_ALPHA) {
    T tprev = (1 - acc[a]) / (1 - c[a]);
    if (exact) {
      acc[a] = 1 - tprev;
      for (int i = 0; i < a; ++i) gc[i] = tprev * g[i];
      gc[a] = tprev * g[a];
      g[a] = (1 - c[a]) * g[a] - s;
    } else {
      acc[a] = tprev;
      T t = 1 - acc[a];
      T d_dacc = (1 - c[a]) * g[a] - s;
      for (int i = 0; i < a; ++i) gc[i] = t * g[i];
      gc[a] = t * g[a];
      g[a] = d_dacc;
    }
  } else {  // ALPHA_ADDITIVE
    acc[a] = acc[a] - c[a];
    T t = 1 - c[a];
    for (int i = 0; i < a; ++i) gc[i] = t * g[i];
    gc[a] = g[a] - s;
  }
}

// ------------------------------------------------------------ view iteration
struct View {
  int batch, cam, out_slot;
};
static std::vector<View> make_views(int N, int nCam, int global) {
  std::vector<View> v;
  for (int b = 0; b < N; ++b) {
    if (global)
      for (int c = 0; c < nCam; ++c) v.push_back({b, c, b * nCam + c});
    else
      v.push_back({b, b, b});
  }
  return v;
}

static int nthreads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
static int tid() {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

// ------------------------------------------------------------------ raymarch
// raymarch_grid.cu.cc:107-143 (fwd), :343-372 + :451-574 (bwd)
template <class T> struct RayGeo {
  v3<T> nearp, step;
};
template <class T> RayGeo<T> ray_geo(const Xf<T>& xf, int x, int y, const Dims& dout) {
  v3<T> dinv = mk3<T>(T(1) / dout.x, T(1) / dout.y, T(1) / dout.z);
  T zMax = (T)std::max(dout.z - 1, 1);
  v3<T> n = idx_to_os(xf, mk3<T>(x + T(0.5), y + T(0.5), T(0.5)), dinv);
  v3<T> f = idx_to_os(xf, mk3<T>(x + T(0.5), y + T(0.5), zMax + T(0.5)), dinv);
  RayGeo<T> g;
  g.nearp = n;
  g.step = mk3<T>((f.x - n.x) / zMax, (f.y - n.y) / zMax, (f.z - n.z) / zMax);
  return g;
}

template <class T>
void raymarch_fwd(const double* input, const int64_t* ishape, const float* M, const float* V, const float* P,
                  const float* F, int nCam, const int64_t* out_dhw, int fm, int bm, int bl, int global,
                  double* output) {
  int N = (int)ishape[0], C = (int)ishape[4];
  Dims din{(int)ishape[3], (int)ishape[2], (int)ishape[1]};
  Dims dout{(int)out_dhw[2], (int)out_dhw[1], (int)out_dhw[0]};
  size_t vox = size_t(din.x) * din.y * din.z;
  std::vector<T> in(vox * C * N);
  for (size_t i = 0; i < in.size(); ++i) in[i] = (T)input[i];
  auto views = make_views(N, nCam, global);
  size_t px = size_t(dout.x) * dout.y;
  for (const View& vw : views) {
    Xf<T> xf = make_xf<T>(M + 16 * vw.batch, V + 16 * vw.cam, P + 16 * vw.cam, F + 6 * vw.cam);
    const T* buf = in.data() + vox * C * vw.batch;
    double* out = output + px * C * vw.out_slot;
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < dout.y; ++y)
      for (int x = 0; x < dout.x; ++x) {
        RayGeo<T> g = ray_geo(xf, x, y, dout);
        v3<T> p = g.nearp;
        T acc[4] = {0, 0, 0, 0}, s[4];
        for (int i = 0; i < dout.z; ++i) {
          sample_v2(buf, din, C, fm, bm, p, s);
          blend(bl, C, acc, s);
          p = add3(p, g.step);
        }
        for (int c = 0; c < C; ++c) out[(size_t(y) * dout.x + x) * C + c] = acc[c];
      }
  }
}

template <class T>
void raymarch_bwd(int gm, const double* input, const int64_t* ishape, const double* output, const double* out_grad,
                  const float* M, const float* V, const float* P, const float* F, int nCam, const int64_t* out_dhw,
                  int bm, int bl, int global, double* in_grad) {
  const bool exact = gm == GM_EXACT;
  int N = (int)ishape[0], C = (int)ishape[4];
  Dims din{(int)ishape[3], (int)ishape[2], (int)ishape[1]};
  Dims dout{(int)out_dhw[2], (int)out_dhw[1], (int)out_dhw[0]};
  size_t vox = size_t(din.x) * din.y * din.z;
  std::vector<T> in(vox * C * N);
  for (size_t i = 0; i < in.size(); ++i) in[i] = (T)input[i];
  std::vector<T> total(vox * C * N, T(0));
  auto views = make_views(N, nCam, global);
  size_t px = size_t(dout.x) * dout.y;
  int nt = nthreads();
  std::vector<std::vector<T>> tg(nt, std::vector<T>(vox * C));
  std::vector<std::vector<T>> tw(nt, std::vector<T>(exact ? 0 : vox));
  for (const View& vw : views) {
    Xf<T> xf = make_xf<T>(M + 16 * vw.batch, V + 16 * vw.cam, P + 16 * vw.cam, F + 6 * vw.cam);
    const T* buf = in.data() + vox * C * vw.batch;
    const double* o = output + px * C * vw.out_slot;
    const double* og = out_grad + px * C * vw.out_slot;
    for (int t = 0; t < nt; ++t) {
      std::fill(tg[t].begin(), tg[t].end(), T(0));
      std::fill(tw[t].begin(), tw[t].end(), T(0));
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < dout.y; ++y) {
      T* gbuf = tg[tid()].data();
      T* wbuf = exact ? nullptr : tw[tid()].data();
      std::vector<v3<T>> pos(exact ? dout.z : 0);
      for (int x = 0; x < dout.x; ++x) {
        RayGeo<T> g = ray_geo(xf, x, y, dout);
        size_t pi = (size_t(y) * dout.x + x) * C;
        T acc[4], gr[4], s[4], gc[4];
        for (int c = 0; c < C; ++c) {
          acc[c] = (T)o[pi + c];
          gr[c] = (T)og[pi + c];
        }
        // REFERENCE walks from the far point subtracting the step (fp32
        // incremental positions); EXACT uses the forward's positions.
        v3<T> pfar = g.nearp;
        if (exact) {
          v3<T> p = g.nearp;
          for (int i = 0; i < dout.z; ++i) {
            pos[i] = p;
            p = add3(p, g.step);
          }
        } else {
          T zMax = (T)std::max(dout.z - 1, 1);
          v3<T> dinv = mk3<T>(T(1) / dout.x, T(1) / dout.y, T(1) / dout.z);
          pfar = idx_to_os(xf, mk3<T>(x + T(0.5), y + T(0.5), zMax + T(0.5)), dinv);
        }
        v3<T> p = pfar;
        for (int i = dout.z - 1; i >= 0; --i) {
          v3<T> q = exact ? pos[i] : p;
          sample_v2(buf, din, C, F_LINEAR, bm, q, s);
          blend_grad(bl, exact, C, gr, s, acc, gc);
          scatter_linear(din, bm, exact, q.x - T(0.5), q.y - T(0.5), q.z - T(0.5), [&](size_t f, T w) {
            for (int c = 0; c < C; ++c) gbuf[f * C + c] += gc[c] * w;
            if (wbuf) wbuf[f] += w;
          });
          if (!exact) p = sub3(p, g.step);
        }
      }
    }
    T* tot = total.data() + vox * C * vw.batch;
#pragma omp parallel for
    for (long long v = 0; v < (long long)vox; ++v) {
      T w = 0;
      T gsum[4] = {0, 0, 0, 0};
      for (int t = 0; t < nt; ++t) {
        if (!exact) w += tw[t][v];
        for (int c = 0; c < C; ++c) gsum[c] += tg[t][v * C + c];
      }
      if (exact) {
        for (int c = 0; c < C; ++c) tot[v * C + c] += gsum[c];
      } else if (w > 0) {
        T inv = T(1) / w;
        for (int c = 0; c < C; ++c) tot[v * C + c] += gsum[c] * inv;
      }
    }
  }
  for (size_t i = 0; i < total.size(); ++i) in_grad[i] = (double)total[i];
}

// ------------------------------------------------------------- reduce blend
// reduce_blend.cu.cc:34-79
template <class T>
void reduce_blend_fwd(const double* input, const int64_t* ishape, int bl, int keep, double* output) {
  int N = (int)ishape[0], D = (int)ishape[1], H = (int)ishape[2], W = (int)ishape[3], C = (int)ishape[4];
  for (int n = 0; n < N; ++n)
#pragma omp parallel for
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        T acc[4] = {0, 0, 0, 0}, c[4];
        for (int z = 0; z < D; ++z) {
          size_t ii = ((((size_t)n * D + z) * H + y) * W + x) * C;
          for (int k = 0; k < C; ++k) c[k] = (T)input[ii + k];
          blend(bl, C, acc, c);
          if (keep)
            for (int k = 0; k < C; ++k) output[ii + k] = acc[k];
        }
        if (!keep) {
          size_t oi = (((size_t)n * H + y) * W + x) * C;
          for (int k = 0; k < C; ++k) output[oi + k] = acc[k];
        }
      }
}
template <class T>
void reduce_blend_bwd(int gm, const double* input, const int64_t* ishape, const double* output,
                      const double* out_grad, int bl, int keep, double* in_grad) {
  bool exact = gm == GM_EXACT;
  int N = (int)ishape[0], D = (int)ishape[1], H = (int)ishape[2], W = (int)ishape[3], C = (int)ishape[4];
  for (int n = 0; n < N; ++n)
#pragma omp parallel for
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        T g[4] = {0, 0, 0, 0}, prev[4] = {0, 0, 0, 0}, c[4], gc[4];
        if (!keep) {
          size_t oi = (((size_t)n * H + y) * W + x) * C;
          for (int k = 0; k < C; ++k) {
            g[k] = (T)out_grad[oi + k];
            prev[k] = (T)output[oi + k];
          }
        }
        for (int z = D - 1; z >= 0; --z) {
          size_t ii = ((((size_t)n * D + z) * H + y) * W + x) * C;
          if (keep)
            for (int k = 0; k < C; ++k) {
              g[k] += (T)out_grad[ii + k];
              prev[k] = (T)output[ii + k];
            }
          for (int k = 0; k < C; ++k) c[k] = (T)input[ii + k];
          blend_grad(bl, exact, C, g, c, prev, gc);
          for (int k = 0; k < C; ++k) in_grad[ii + k] = gc[k];
        }
      }
}

// ------------------------------------------------------------ mip pyramid
// kGenerateMip3D (resample_grid.cu.cc:46-72): pairwise means in z, y, x order.
template <class T> struct Pyramid {
  std::vector<Dims> dims;
  std::vector<std::vector<T>> lv;  // level 0 = base copy
};
template <class T> Pyramid<T> build_pyramid(const T* base, Dims d, int C, int levels) {
  Pyramid<T> p;
  p.dims.push_back(d);
  p.lv.emplace_back(base, base + size_t(d.x) * d.y * d.z * C);
  for (int m = 0; m < levels; ++m) {
    Dims s = p.dims.back();
    Dims o{s.x >> 1, s.y >> 1, s.z >> 1};
    std::vector<T> out(size_t(o.x) * o.y * o.z * C);
    const std::vector<T>& in = p.lv.back();
    for (int k = 0; k < o.z; ++k)
      for (int j = 0; j < o.y; ++j)
        for (int i = 0; i < o.x; ++i)
          for (int c = 0; c < C; ++c) {
            T ya[2];
            for (int xi = 0; xi < 2; ++xi) {
              T za[2];
              for (int yi = 0; yi < 2; ++yi) {
                T a = in[flat3(2 * i + xi, 2 * j + yi, 2 * k, s) * C + c];
                T b = in[flat3(2 * i + xi, 2 * j + yi, 2 * k + 1, s) * C + c];
                za[yi] = (a + b) * T(0.5);
              }
              ya[xi] = (za[0] + za[1]) * T(0.5);
            }
            out[flat3(i, j, k, o) * C + c] = (ya[0] + ya[1]) * T(0.5);
          }
    p.dims.push_back(o);
    p.lv.push_back(std::move(out));
  }
  return p;
}
// setupMipAtlas (resample_grid.cc:201-220): levels that would be empty are dropped
inline int clamp_mip_levels(Dims d, int levels) {
  for (int m = 1; m < levels + 1; ++m) {
    long long e = (long long)(d.x >> m) * (d.y >> m) * (d.z >> m);
    if (e <= 0) return m - 1;
  }
  return levels;
}

// --------------------------------------------------- v1 sample positions
template <class T> struct Pos4 {
  v3<T> p;
  T lod;
};
// samplePos3D (sampling.hpp:59-136) for the four coordinate modes, with and
// without LoD. dims_in/dims_out are the sampled-from / sampled-to grid dims.
template <class T> Pos4<T> sample_pos(const Xf<T>& xf, int cm, bool lod, v3<T> ic, Dims din, Dims dout) {
  v3<T> oinv = mk3<T>(T(1) / dout.x, T(1) / dout.y, T(1) / dout.z);
  v3<T> idims = mk3<T>((T)din.x, (T)din.y, (T)din.z);
  auto fwdmap = [&](v3<T> q) -> v3<T> {
    switch (cm) {
      case CM_LINDEPTH: return idx_to_os_lindepth(xf, q, oinv);
      case CM_TRANSFORM: return idx_to_os(xf, q, oinv);
      case CM_LINDEPTH_REVERSE: return os_to_idx_lindepth(xf, q, idims);
      default: return os_to_idx(xf, q, idims);
    }
  };
  Pos4<T> r;
  r.p = fwdmap(ic);
  r.lod = 0;
  if (lod) {
    v3<T> dx = sub3(r.p, fwdmap(add3(ic, mk3<T>(1, 0, 0))));
    v3<T> dy = sub3(r.p, fwdmap(add3(ic, mk3<T>(0, 1, 0))));
    v3<T> dz = sub3(r.p, fwdmap(add3(ic, mk3<T>(0, 0, 1))));
    T px, py, pz;
    if (cm == CM_LINDEPTH || cm == CM_TRANSFORM) {  // Jacobian rows
      px = len3(mk3(dx.x, dy.x, dz.x));
      py = len3(mk3(dx.y, dy.y, dz.y));
      pz = len3(mk3(dx.z, dy.z, dz.z));
    } else {  // Jacobian columns
      px = len3(dx);
      py = len3(dy);
      pz = len3(dz);
    }
    T dmax = std::max(px, std::max(py, pz));
    r.lod = std::max(T(0), std::log2(dmax));
  }
  return r;
}
// calcLoD (sampling.hpp:21-56): central differences, Jacobian rows for both modes
template <class T> T calc_lod(const Xf<T>& xf, int cm, v3<T> ic, Dims din, Dims dout) {
  v3<T> oinv = mk3<T>(T(1) / dout.x, T(1) / dout.y, T(1) / dout.z);
  v3<T> idims = mk3<T>((T)din.x, (T)din.y, (T)din.z);
  auto fm = [&](v3<T> q) -> v3<T> {
    return cm == CM_LINDEPTH ? idx_to_os_lindepth(xf, q, oinv) : os_to_idx_lindepth(xf, q, idims);
  };
  v3<T> dx = scl3(sub3(fm(add3(ic, mk3<T>(-1, 0, 0))), fm(add3(ic, mk3<T>(1, 0, 0)))), T(0.5));
  v3<T> dy = scl3(sub3(fm(add3(ic, mk3<T>(0, -1, 0))), fm(add3(ic, mk3<T>(0, 1, 0)))), T(0.5));
  v3<T> dz = scl3(sub3(fm(add3(ic, mk3<T>(0, 0, -1))), fm(add3(ic, mk3<T>(0, 0, 1)))), T(0.5));
  T px = len3(mk3(dx.x, dy.x, dz.x));
  T py = len3(mk3(dx.y, dy.y, dz.y));
  T pz = len3(mk3(dx.z, dy.z, dz.z));
  T dmax = std::max(px, std::max(py, pz));
  return std::max(T(0), std::log2(dmax));
}

// Sampler for the resample and LuT ops (Sampler2::sample3D sampling.hpp:586-605
// and the free sample3D sampling.hpp:720-740): level select + optional lerp.
struct SamplerCfg {
  int fm, bm, mm, levels;
  double bias, offset;
  bool v2_wrap_zero;  // Sampler2 knows only BORDER/CLAMP: WRAP reads 0
};
template <class T> struct LevelTap {
  int level;
  T w;
};
template <class T> int level_taps(const SamplerCfg& s, T lod, LevelTap<T>* taps) {
  if (s.mm == MIP_NONE) {
    taps[0] = {0, T(1)};
    return 1;
  }
  T l = std::min(std::max(lod + (T)s.bias, T(0)), (T)s.levels);
  int fl = std::min((int)(l + (s.mm == MIP_NEAREST ? T(0.5) : T(0))), s.levels);
  if (s.mm == MIP_LINEAR) {
    int cl = std::min((int)std::ceil(l), s.levels);
    if (cl > fl) {
      T fr = fract(l);
      taps[0] = {fl, T(1) - fr};
      taps[1] = {cl, fr};
      return 2;
    }
  }
  taps[0] = {fl, T(1)};
  return 1;
}
template <class T>
void fetch_v1(const SamplerCfg& s, const T* buf, Dims d, int C, v3<T> p, T* out) {
  T px = p.x - (T)s.offset, py = p.y - (T)s.offset, pz = p.z - (T)s.offset;
  if (s.v2_wrap_zero && s.bm == B_WRAP) {
    for (int c = 0; c < C; ++c) out[c] = 0;
    return;
  }
  if (s.fm == F_LINEAR) read_linear(buf, d, C, s.bm, px, py, pz, out);
  else if (s.fm == F_NEAREST) read_nearest(buf, d, C, s.bm, px, py, pz, out);
  else read_minmax_v1(buf, d, C, s.bm, s.fm == F_MAX, px, py, pz, out);
}
template <class T>
void sample_v1(const SamplerCfg& s, const Pyramid<T>& pyr, int C, Pos4<T> pos, T* out) {
  LevelTap<T> taps[2];
  int n = level_taps(s, pos.lod, taps);
  T a[4], b[4];
  {
    int m = taps[0].level;
    T sc = T(1) / T(1 << m);
    fetch_v1(s, pyr.lv[m].data(), pyr.dims[m], C, scl3(pos.p, sc), a);
  }
  if (n == 2) {
    int m = taps[1].level;
    T sc = T(1) / T(1 << m);
    fetch_v1(s, pyr.lv[m].data(), pyr.dims[m], C, scl3(pos.p, sc), b);
    for (int c = 0; c < C; ++c) a[c] = lerp(a[c], b[c], taps[1].w);
  }
  for (int c = 0; c < C; ++c) out[c] = a[c];
}

// ------------------------------------------------------- sample transform
// resample_grid.cu.cc:240-252, :289-438
template <class T>
void sample_transform_fwd(const double* input, const int64_t* ishape, const float* M, const float* V,
                          const float* P, const float* F, int nCam, const int64_t* out_dhw, int fm, int bm,
                          int mm, int levels, double bias, int cm, int global, double* output) {
  int N = (int)ishape[0], C = (int)ishape[4];
  Dims din{(int)ishape[3], (int)ishape[2], (int)ishape[1]};
  Dims dout{(int)out_dhw[2], (int)out_dhw[1], (int)out_dhw[0]};
  size_t vox = size_t(din.x) * din.y * din.z, ovox = size_t(dout.x) * dout.y * dout.z;
  SamplerCfg s{fm, bm, mm, mm == MIP_NONE ? 0 : clamp_mip_levels(din, levels), bias, 0.5, true};
  auto views = make_views(N, nCam, global);
  int lastb = -1;
  Pyramid<T> pyr;
  for (const View& vw : views) {
    if (vw.batch != lastb) {
      std::vector<T> base(vox * C);
      for (size_t i = 0; i < base.size(); ++i) base[i] = (T)input[vox * C * vw.batch + i];
      pyr = build_pyramid(base.data(), din, C, s.levels);
      lastb = vw.batch;
    }
    Xf<T> xf = make_xf<T>(M + 16 * vw.batch, V + 16 * vw.cam, P + 16 * vw.cam, F + 6 * vw.cam);
    double* out = output + ovox * C * vw.out_slot;
#pragma omp parallel for schedule(dynamic, 2)
    for (int z = 0; z < dout.z; ++z)
      for (int y = 0; y < dout.y; ++y)
        for (int x = 0; x < dout.x; ++x) {
          Pos4<T> pos = sample_pos(xf, cm, mm != MIP_NONE, mk3<T>(x + T(0.5), y + T(0.5), z + T(0.5)), din, dout);
          T o[4];
          sample_v1(s, pyr, C, pos, o);
          for (int c = 0; c < C; ++c) out[flat3(x, y, z, dout) * C + c] = o[c];
        }
  }
}

// Raw sample positions and LoD of every output cell centre for one camera
// (samplePos3D, sampling.hpp:59-136): what sample_transform_fwd feeds its
// sampler, exported so the coordinate maps can be checked on their own.
template <class T>
void sample_positions(const float* M, const float* V, const float* P, const float* F, int nCam,
                      const int64_t* in_dhw, const int64_t* out_dhw, int cm, int lod, double* out) {
  Dims din{(int)in_dhw[2], (int)in_dhw[1], (int)in_dhw[0]};
  Dims dout{(int)out_dhw[2], (int)out_dhw[1], (int)out_dhw[0]};
  size_t ovox = size_t(dout.x) * dout.y * dout.z;
  for (int c = 0; c < nCam; ++c) {
    Xf<T> xf = make_xf<T>(M, V + 16 * c, P + 16 * c, F + 6 * c);
    double* o = out + ovox * 4 * c;
    for (int z = 0; z < dout.z; ++z)
      for (int y = 0; y < dout.y; ++y)
        for (int x = 0; x < dout.x; ++x) {
          Pos4<T> pos = sample_pos(xf, cm, lod != 0, mk3<T>(x + T(0.5), y + T(0.5), z + T(0.5)), din, dout);
          double* q = o + flat3(x, y, z, dout) * 4;
          q[0] = pos.p.x;
          q[1] = pos.p.y;
          q[2] = pos.p.z;
          q[3] = pos.lod;
        }
  }
}

// Adjoint helpers: scatter one tap of a v1 fetch into a level gradient.
template <class T, class Add>
void fetch_v1_adjoint(const SamplerCfg& s, const T* buf, Dims d, int C, v3<T> p, const T* g, T scale, Add&& add) {
  T px = p.x - (T)s.offset, py = p.y - (T)s.offset, pz = p.z - (T)s.offset;
  if (s.v2_wrap_zero && s.bm == B_WRAP) return;
  if (s.fm == F_LINEAR) {
    scatter_linear(d, s.bm, true, px, py, pz, [&](size_t f, T w) {
      for (int c = 0; c < C; ++c) add(f, c, g[c] * w * scale);
    });
  } else if (s.fm == F_NEAREST) {
    int ix = (int)(px + T(0.5)), iy = (int)(py + T(0.5)), iz = (int)(pz + T(0.5));
    if (s.bm == B_BORDER) {
      if (!(0 <= ix && ix < d.x && 0 <= iy && iy < d.y && 0 <= iz && iz < d.z)) return;
    } else if (s.bm == B_CLAMP) {
      ix = clampi(ix, 0, d.x - 1); iy = clampi(iy, 0, d.y - 1); iz = clampi(iz, 0, d.z - 1);
    } else {
      ix = pmod(ix, d.x); iy = pmod(iy, d.y); iz = pmod(iz, d.z);
    }
    size_t f = flat3(ix, iy, iz, d);
    for (int c = 0; c < C; ++c) add(f, c, g[c] * scale);
  } else {
    T o[4];
    int arg[4];
    read_minmax_v1(buf, d, C, s.bm, s.fm == F_MAX, px, py, pz, o, arg);
    for (int c = 0; c < C; ++c)
      if (arg[c] >= 0) add((size_t)arg[c], c, g[c] * scale);
  }
}
// mip adjoint: children receive 1/8 of the parent gradient, coarse to fine
template <class T> void collapse_pyramid_grads(std::vector<std::vector<T>>& gl, const std::vector<Dims>& dims, int C) {
  for (int m = (int)gl.size() - 1; m >= 1; --m) {
    Dims o = dims[m], s = dims[m - 1];
    for (int k = 0; k < o.z; ++k)
      for (int j = 0; j < o.y; ++j)
        for (int i = 0; i < o.x; ++i)
          for (int c = 0; c < C; ++c) {
            T g = gl[m][flat3(i, j, k, o) * C + c] * T(0.125);
            for (int dz = 0; dz < 2; ++dz)
              for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) gl[m - 1][flat3(2 * i + dx, 2 * j + dy, 2 * k + dz, s) * C + c] += g;
          }
  }
}

// REFERENCE: kScatter3DGradients + kNormalize3DGradients per batch
// (resample_grid.cu.cc:761-861, :865-1018). EXACT: adjoint incl. mips.
template <class T>
void sample_transform_bwd(int gm, const double* input, const int64_t* ishape, const double* out_grad,
                          const float* M, const float* V, const float* P, const float* F, int nCam,
                          const int64_t* out_dhw, int fm, int bm, int mm, int levels, double bias, int cm,
                          int global, double* in_grad) {
  bool exact = gm == GM_EXACT;
  int N = (int)ishape[0], C = (int)ishape[4];
  Dims din{(int)ishape[3], (int)ishape[2], (int)ishape[1]};
  Dims dout{(int)out_dhw[2], (int)out_dhw[1], (int)out_dhw[0]};
  size_t vox = size_t(din.x) * din.y * din.z, ovox = size_t(dout.x) * dout.y * dout.z;
  auto views = make_views(N, nCam, global);
  std::fill(in_grad, in_grad + vox * C * N, 0.0);
  SamplerCfg s{fm, bm, exact ? mm : MIP_NONE, 0, bias, 0.5, true};
  if (exact && mm != MIP_NONE) s.levels = clamp_mip_levels(din, levels);
  for (int b = 0; b < N; ++b) {
    std::vector<T> base(vox * C);
    for (size_t i = 0; i < base.size(); ++i) base[i] = (T)input[vox * C * b + i];
    Pyramid<T> pyr = build_pyramid(base.data(), din, C, s.levels);
    std::vector<std::vector<T>> gl(pyr.lv.size());
    for (size_t m = 0; m < gl.size(); ++m) gl[m].assign(pyr.lv[m].size(), T(0));
    std::vector<T> cnt(vox, T(0));
    for (const View& vw : views) {
      if (vw.batch != b) continue;
      Xf<T> xf = make_xf<T>(M + 16 * vw.batch, V + 16 * vw.cam, P + 16 * vw.cam, F + 6 * vw.cam);
      const double* og = out_grad + ovox * C * vw.out_slot;
      // sequential accumulation keeps this simple and deterministic
      for (int z = 0; z < dout.z; ++z)
        for (int y = 0; y < dout.y; ++y)
          for (int x = 0; x < dout.x; ++x) {
            T g[4];
            for (int c = 0; c < C; ++c) g[c] = (T)og[flat3(x, y, z, dout) * C + c];
            Pos4<T> pos = sample_pos(xf, cm, exact && mm != MIP_NONE, mk3<T>(x + T(0.5), y + T(0.5), z + T(0.5)),
                                     din, dout);
            if (!exact) {
              scatter_linear(din, bm, false, pos.p.x - T(0.5), pos.p.y - T(0.5), pos.p.z - T(0.5),
                             [&](size_t f, T w) {
                               for (int c = 0; c < C; ++c) gl[0][f * C + c] += g[c] * w;
                               cnt[f] += 1;
                             });
            } else {
              LevelTap<T> taps[2];
              int n = level_taps(s, pos.lod, taps);
              for (int t = 0; t < n; ++t) {
                int m = taps[t].level;
                T sc = T(1) / T(1 << m);
                fetch_v1_adjoint(s, pyr.lv[m].data(), pyr.dims[m], C, scl3(pos.p, sc), g, taps[t].w,
                                 [&](size_t f, int c, T v) { gl[m][f * C + c] += v; });
              }
            }
          }
    }
    if (exact) {
      collapse_pyramid_grads(gl, pyr.dims, C);
      for (size_t i = 0; i < vox * C; ++i) in_grad[vox * C * b + i] = gl[0][i];
    } else {
      for (size_t v = 0; v < vox; ++v)
        for (int c = 0; c < C; ++c) {
          T g = gl[0][v * C + c];
          if (cnt[v] > 0) g = g * (T(1) / cnt[v]);
          in_grad[vox * C * b + v * C + c] = g;
        }
    }
  }
}

// ------------------------------------------------------------------- LuT
// kSampleGrid3DLuT (resample_grid.cu.cc:465-484)
template <class T>
Pos4<T> lut_pos(const double* lut, size_t idx, int x, int y, int z, Dims dout, int relative, int normalized) {
  Pos4<T> r;
  r.p = mk3<T>((T)lut[idx * 4], (T)lut[idx * 4 + 1], (T)lut[idx * 4 + 2]);
  r.lod = (T)lut[idx * 4 + 3];
  if (normalized) r.p = mk3<T>(r.p.x * dout.x, r.p.y * dout.y, r.p.z * dout.z);
  if (relative) r.p = mk3<T>(r.p.x + x, r.p.y + y, r.p.z + z);
  return r;
}
template <class T>
void sample_lut_fwd(const double* input, const int64_t* ishape, const double* lut, const int64_t* lshape, int fm,
                    int bm, int mm, int levels, double bias, double offset, int global, int relative, int normalized,
                    double* output) {
  int N = (int)ishape[0], C = (int)ishape[4];
  int nCam = (int)lshape[0];
  Dims din{(int)ishape[3], (int)ishape[2], (int)ishape[1]};
  Dims dout{(int)lshape[3], (int)lshape[2], (int)lshape[1]};
  size_t vox = size_t(din.x) * din.y * din.z, ovox = size_t(dout.x) * dout.y * dout.z;
  SamplerCfg s{fm, bm, mm, mm == MIP_NONE ? 0 : clamp_mip_levels(din, levels), bias, offset, false};
  auto views = make_views(N, nCam, global);
  int lastb = -1;
  Pyramid<T> pyr;
  for (const View& vw : views) {
    if (vw.batch != lastb) {
      std::vector<T> base(vox * C);
      for (size_t i = 0; i < base.size(); ++i) base[i] = (T)input[vox * C * vw.batch + i];
      pyr = build_pyramid(base.data(), din, C, s.levels);
      lastb = vw.batch;
    }
    const double* L = lut + ovox * 4 * vw.cam;
    double* out = output + ovox * C * vw.out_slot;
#pragma omp parallel for schedule(dynamic, 2)
    for (int z = 0; z < dout.z; ++z)
      for (int y = 0; y < dout.y; ++y)
        for (int x = 0; x < dout.x; ++x) {
          size_t f = flat3(x, y, z, dout);
          Pos4<T> pos = lut_pos<T>(L, f, x, y, z, dout, relative, normalized);
          T o[4];
          sample_v1(s, pyr, C, pos, o);
          for (int c = 0; c < C; ++c) out[f * C + c] = o[c];
        }
  }
}

// Spatial derivative of the trilinear read (read3DGrad, sampling.hpp:257-352).
// REFERENCE uses floor/ceil corners (zero derivative at integer positions);
// EXACT uses floor/floor+1 (the right derivative) with the same boundary rule.
template <class T>
void read_grad(const T* buf, Dims d, int C, int bm, bool exact, T px, T py, T pz, T* gx, T* gy, T* gz) {
  for (int c = 0; c < C; ++c) gx[c] = gy[c] = gz[c] = 0;
  T wx = fract(px), wy = fract(py), wz = fract(pz);
  int fx = (int)std::floor(px), fy = (int)std::floor(py), fz = (int)std::floor(pz);
  int cx = exact ? fx + 1 : (int)std::ceil(px), cy = exact ? fy + 1 : (int)std::ceil(py),
      cz = exact ? fz + 1 : (int)std::ceil(pz);
  bool vx[2] = {true, true}, vy[2] = {true, true}, vz[2] = {true, true};
  if (bm == B_BORDER) {
    T qx = px + T(0.5), qy = py + T(0.5), qz = pz + T(0.5);
    if (!(0 <= qx && qx < d.x && 0 <= qy && qy < d.y && 0 <= qz && qz < d.z)) return;
    vx[0] = -1 < fx && fx < d.x; vx[1] = -1 < cx && cx < d.x;
    vy[0] = -1 < fy && fy < d.y; vy[1] = -1 < cy && cy < d.y;
    vz[0] = -1 < fz && fz < d.z; vz[1] = -1 < cz && cz < d.z;
  } else if (bm == B_CLAMP) {
    fx = clampi(fx, 0, d.x - 1); cx = clampi(cx, 0, d.x - 1);
    fy = clampi(fy, 0, d.y - 1); cy = clampi(cy, 0, d.y - 1);
    fz = clampi(fz, 0, d.z - 1); cz = clampi(cz, 0, d.z - 1);
  } else {
    fx = pmod(fx, d.x); cx = pmod(cx, d.x);
    fy = pmod(fy, d.y); cy = pmod(cy, d.y);
    fz = pmod(fz, d.z); cz = pmod(cz, d.z);
  }
  int ix[2] = {fx, cx}, iy[2] = {fy, cy}, iz[2] = {fz, cz};
  for (int c = 0; c < C; ++c) {
    T v[2][2][2];  // [z][y][x]
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int i = 0; i < 2; ++i)
          v[k][j][i] = (vx[i] && vy[j] && vz[k]) ? buf[flat3(ix[i], iy[j], iz[k], d) * C + c] : T(0);
    gx[c] = lerp(lerp(v[0][0][1] - v[0][0][0], v[0][1][1] - v[0][1][0], wy),
                 lerp(v[1][0][1] - v[1][0][0], v[1][1][1] - v[1][1][0], wy), wz);
    gy[c] = lerp(lerp(v[0][1][0] - v[0][0][0], v[0][1][1] - v[0][0][1], wx),
                 lerp(v[1][1][0] - v[1][0][0], v[1][1][1] - v[1][0][1], wx), wz);
    gz[c] = lerp(lerp(v[1][0][0] - v[0][0][0], v[1][0][1] - v[0][0][1], wx),
                 lerp(v[1][1][0] - v[0][1][0], v[1][1][1] - v[0][1][1], wx), wy);
  }
}

// REFERENCE: kScatterLuT3DGradients + count normalisation per batch
// (resample_grid.cu.cc:1051-1238); the lut gradient is written, not summed,
// per (batch, camera). EXACT: adjoint w.r.t. input and lookup (summed over
// batches, scaled by the output dims for normalised coordinates).
template <class T>
void sample_lut_bwd(int gm, const double* input, const int64_t* ishape, const double* lut, const int64_t* lshape,
                    const double* out_grad, int fm, int bm, double offset, int global, int relative, int normalized,
                    double* in_grad, double* lut_grad) {
  bool exact = gm == GM_EXACT;
  int N = (int)ishape[0], C = (int)ishape[4];
  int nCam = (int)lshape[0];
  Dims din{(int)ishape[3], (int)ishape[2], (int)ishape[1]};
  Dims dout{(int)lshape[3], (int)lshape[2], (int)lshape[1]};
  size_t vox = size_t(din.x) * din.y * din.z, ovox = size_t(dout.x) * dout.y * dout.z;
  SamplerCfg s{fm, bm, MIP_NONE, 0, 0.0, offset, false};
  auto views = make_views(N, nCam, global);
  std::fill(in_grad, in_grad + vox * C * N, 0.0);
  std::fill(lut_grad, lut_grad + ovox * 4 * nCam, 0.0);
  for (int b = 0; b < N; ++b) {
    std::vector<T> base(vox * C);
    for (size_t i = 0; i < base.size(); ++i) base[i] = (T)input[vox * C * b + i];
    std::vector<T> gacc(vox * C, T(0)), cnt(vox, T(0));
    for (const View& vw : views) {
      if (vw.batch != b) continue;
      const double* L = lut + ovox * 4 * vw.cam;
      double* LG = lut_grad + ovox * 4 * vw.cam;
      const double* og = out_grad + ovox * C * vw.out_slot;
      for (int z = 0; z < dout.z; ++z)
        for (int y = 0; y < dout.y; ++y)
          for (int x = 0; x < dout.x; ++x) {
            size_t f = flat3(x, y, z, dout);
            Pos4<T> pos = lut_pos<T>(L, f, x, y, z, dout, relative, normalized);
            T g[4];
            for (int c = 0; c < C; ++c) g[c] = (T)og[f * C + c];
            T px = pos.p.x - (T)offset, py = pos.p.y - (T)offset, pz = pos.p.z - (T)offset;
            if (fm == F_LINEAR) {
              T gx[4], gy[4], gz[4];
              read_grad(base.data(), din, C, bm, exact, px, py, pz, gx, gy, gz);
              T lx = 0, ly = 0, lz = 0;
              for (int c = 0; c < C; ++c) {
                lx += g[c] * gx[c];
                ly += g[c] * gy[c];
                lz += g[c] * gz[c];
              }
              if (exact) {
                if (normalized) {
                  lx *= dout.x;
                  ly *= dout.y;
                  lz *= dout.z;
                }
                LG[f * 4 + 0] += lx;
                LG[f * 4 + 1] += ly;
                LG[f * 4 + 2] += lz;
              } else {
                LG[f * 4 + 0] = lx;
                LG[f * 4 + 1] = ly;
                LG[f * 4 + 2] = lz;
                LG[f * 4 + 3] = 0;
              }
            }
            if (!exact) {
              scatter_linear(din, bm, false, px, py, pz, [&](size_t v, T w) {
                for (int c = 0; c < C; ++c) gacc[v * C + c] += g[c] * w;
                cnt[v] += 1;
              });
            } else {
              fetch_v1_adjoint(s, base.data(), din, C, pos.p, g, T(1),
                               [&](size_t v, int c, T val) { gacc[v * C + c] += val; });
            }
          }
    }
    for (size_t v = 0; v < vox; ++v)
      for (int c = 0; c < C; ++c) {
        T g = gacc[v * C + c];
        if (!exact && cnt[v] > 0) g = g * (T(1) / cnt[v]);
        in_grad[vox * C * b + v * C + c] = g;
      }
  }
}

// ------------------------------------------------------------------- LoD
// kComputeLoD3D (resample_grid.cu.cc:1288-1300): M = identity, V = MV
template <class T>
void lod_transform(const int64_t* in_dhw, const float* MV, const float* P, const float* F, int nCam,
                   const int64_t* out_dhw, int inverse, double* output) {
  Dims din{(int)in_dhw[2], (int)in_dhw[1], (int)in_dhw[0]};
  Dims dout{(int)out_dhw[2], (int)out_dhw[1], (int)out_dhw[0]};
  size_t ovox = size_t(dout.x) * dout.y * dout.z;
  int cm = inverse ? CM_LINDEPTH_REVERSE : CM_LINDEPTH;
  for (int b = 0; b < nCam; ++b) {
    Xf<T> xf = make_xf<T>(nullptr, MV + 16 * b, P + 16 * b, F + 6 * b);
    double* out = output + ovox * 4 * b;
#pragma omp parallel for
    for (int z = 0; z < dout.z; ++z)
      for (int y = 0; y < dout.y; ++y)
        for (int x = 0; x < dout.x; ++x) {
          v3<T> ic = mk3<T>(x + T(0.5), y + T(0.5), z + T(0.5));
          Pos4<T> p = sample_pos(xf, cm, false, ic, din, dout);
          T lod = calc_lod(xf, cm, ic, din, dout);
          size_t f = flat3(x, y, z, dout) * 4;
          out[f] = p.p.x;
          out[f + 1] = p.p.y;
          out[f + 2] = p.p.z;
          out[f + 3] = lod;
        }
  }
}

}  // namespace orc

// =================================================================== C ABI
extern "C" {

int orc_num_threads() { return orc::nthreads(); }
// thread count of the later OpenMP regions of the calling thread (the CPU
// baseline's single-thread figure)
void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

#define ORC_DISPATCH(prec, call) \
  do {                           \
    if (prec) {                  \
      using T = double;          \
      call;                      \
    } else {                     \
      using T = float;           \
      call;                      \
    }                            \
  } while (0)

int orc_raymarch_fwd(int prec, const double* input, const int64_t* ishape, const float* M, const float* V,
                     const float* P, const float* F, int nCam, const int64_t* out_dhw, int fm, int bm, int bl,
                     int global, double* output) {
  ORC_DISPATCH(prec, (orc::raymarch_fwd<T>(input, ishape, M, V, P, F, nCam, out_dhw, fm, bm, bl, global, output)));
  return 0;
}
int orc_raymarch_bwd(int prec, int gm, const double* input, const int64_t* ishape, const double* output,
                     const double* out_grad, const float* M, const float* V, const float* P, const float* F, int nCam,
                     const int64_t* out_dhw, int bm, int bl, int global, double* in_grad) {
  ORC_DISPATCH(prec, (orc::raymarch_bwd<T>(gm, input, ishape, output, out_grad, M, V, P, F, nCam, out_dhw, bm, bl,
                                           global, in_grad)));
  return 0;
}
int orc_reduce_blend_fwd(int prec, const double* input, const int64_t* ishape, int bl, int keep, double* output) {
  ORC_DISPATCH(prec, (orc::reduce_blend_fwd<T>(input, ishape, bl, keep, output)));
  return 0;
}
int orc_reduce_blend_bwd(int prec, int gm, const double* input, const int64_t* ishape, const double* output,
                         const double* out_grad, int bl, int keep, double* in_grad) {
  ORC_DISPATCH(prec, (orc::reduce_blend_bwd<T>(gm, input, ishape, output, out_grad, bl, keep, in_grad)));
  return 0;
}
int orc_sample_transform_fwd(int prec, const double* input, const int64_t* ishape, const float* M, const float* V,
                             const float* P, const float* F, int nCam, const int64_t* out_dhw, int fm, int bm, int mm,
                             int levels, double bias, int cm, int global, double* output) {
  ORC_DISPATCH(prec, (orc::sample_transform_fwd<T>(input, ishape, M, V, P, F, nCam, out_dhw, fm, bm, mm, levels, bias,
                                                   cm, global, output)));
  return 0;
}
int orc_sample_transform_bwd(int prec, int gm, const double* input, const int64_t* ishape, const double* out_grad,
                             const float* M, const float* V, const float* P, const float* F, int nCam,
                             const int64_t* out_dhw, int fm, int bm, int mm, int levels, double bias, int cm,
                             int global, double* in_grad) {
  ORC_DISPATCH(prec, (orc::sample_transform_bwd<T>(gm, input, ishape, out_grad, M, V, P, F, nCam, out_dhw, fm, bm, mm,
                                                   levels, bias, cm, global, in_grad)));
  return 0;
}
int orc_sample_lut_fwd(int prec, const double* input, const int64_t* ishape, const double* lut, const int64_t* lshape,
                       int fm, int bm, int mm, int levels, double bias, double offset, int global, int relative,
                       int normalized, double* output) {
  ORC_DISPATCH(prec, (orc::sample_lut_fwd<T>(input, ishape, lut, lshape, fm, bm, mm, levels, bias, offset, global,
                                             relative, normalized, output)));
  return 0;
}
int orc_sample_lut_bwd(int prec, int gm, const double* input, const int64_t* ishape, const double* lut,
                       const int64_t* lshape, const double* out_grad, int fm, int bm, double offset, int global,
                       int relative, int normalized, double* in_grad, double* lut_grad) {
  ORC_DISPATCH(prec, (orc::sample_lut_bwd<T>(gm, input, ishape, lut, lshape, out_grad, fm, bm, offset, global, relative,
                                             normalized, in_grad, lut_grad)));
  return 0;
}
int orc_sample_positions(int prec, const float* M, const float* V, const float* P, const float* F, int nCam,
                         const int64_t* in_dhw, const int64_t* out_dhw, int cm, int lod, double* out) {
  ORC_DISPATCH(prec, (orc::sample_positions<T>(M, V, P, F, nCam, in_dhw, out_dhw, cm, lod, out)));
  return 0;
}
int orc_lod_transform(int prec, const int64_t* in_dhw, const float* MV, const float* P, const float* F, int nCam,
                      const int64_t* out_dhw, int inverse, double* output) {
  ORC_DISPATCH(prec, (orc::lod_transform<T>(in_dhw, MV, P, F, nCam, out_dhw, inverse, output)));
  return 0;
}
int orc_build_mips(int prec, const double* input, const int64_t* dhwc, int levels, double* out_concat) {
  // concatenated levels 1..levels (level 0 omitted), for the mip-pyramid tests
  orc::Dims d{(int)dhwc[2], (int)dhwc[1], (int)dhwc[0]};
  int C = (int)dhwc[3];
  size_t vox = size_t(d.x) * d.y * d.z;
  ORC_DISPATCH(prec, ({
                 std::vector<T> base(vox * C);
                 for (size_t i = 0; i < base.size(); ++i) base[i] = (T)input[i];
                 auto pyr = orc::build_pyramid(base.data(), d, C, levels);
                 size_t o = 0;
                 for (int m = 1; m <= levels; ++m)
                   for (T v : pyr.lv[m]) out_concat[o++] = v;
               }));
  return 0;
}
}
