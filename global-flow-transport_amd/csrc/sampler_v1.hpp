// Shared pieces of the v1 sampler (phitest/render/cuda/src/sampling.hpp):
// mip level pointers, level selection, and the sample position and LoD of
// samplePos3D for the four coordinate modes. Used by the resample ops and the
// batched mipmapped render.
#pragma once
#include "gft_device.hpp"

namespace gft {

__device__ __forceinline__ const float* level_ptr(const float* base, const float* atlas, const MipTable& mt, int b,
                                                  int m) {
  return m == 0 ? base + (size_t)b * mt.base_stride : atlas + (size_t)b * mt.atlas_stride + mt.offset[m];
}
__device__ __forceinline__ float* level_ptr_mut(float* base, float* atlas, const MipTable& mt, int b, int m) {
  return m == 0 ? base + (size_t)b * mt.base_stride : atlas + (size_t)b * mt.atlas_stride + mt.offset[m];
}

// Mip level selection (sampling.hpp:551-571,603-620): returns the number of
// taps (1 or 2) with their levels and weights.
__device__ __forceinline__ int level_taps(const SampleCfg& s, float lod, int* lv, float* w) {
  if (s.mm == MIP_NONE) {
    lv[0] = 0;
    w[0] = 1.f;
    return 1;
  }
  float l = fminf(fmaxf(lod + s.bias, 0.f), (float)s.levels);
  int fl = min((int)(l + (s.mm == MIP_NEAREST ? 0.5f : 0.f)), s.levels);
  if (s.mm == MIP_LINEAR) {
    int cl = min((int)ceilf(l), s.levels);
    if (cl > fl) {
      float fr = l - floorf(l);
      lv[0] = fl;
      w[0] = 1.f - fr;
      lv[1] = cl;
      w[1] = fr;
      return 2;
    }
  }
  lv[0] = fl;
  w[0] = 1.f;
  return 1;
}

// Coordinate maps of samplePos3D (sampling.hpp:59-136) with the v1
// transforms (transformations.hpp:86-145). q is an output index + 0.5.
__device__ __forceinline__ float3 map_pos(const SampleView& v, int cm, float qx, float qy, float qz, float3 oinv,
                                          float3 idims) {
  if (cm == CM_TRANSFORM || cm == CM_LINDEPTH) {
    float nx = qx * oinv.x * 2.f - 1.f, ny = qy * oinv.y * 2.f - 1.f, nz = qz * oinv.z * 2.f - 1.f;
    if (cm == CM_LINDEPTH) {
      float t = qz * oinv.z;
      float z = -(v.nearp + (v.farp - v.nearp) * t);
      nz = -v.p22 - v.p32 / z;
    }
    float4 vs = mat_vec(v.A, nx, ny, nz, 1.f);
    float4 os = mat_vec(v.B, vs.x / vs.w, vs.y / vs.w, vs.z / vs.w, 1.f);
    return make_float3(os.x, os.y, os.z);
  }
  float4 vs = mat_vec(v.A, qx, qy, qz, 1.f);
  float4 ndc = mat_vec(v.B, vs.x, vs.y, vs.z, vs.w);
  float ndx = ndc.x / ndc.w, ndy = ndc.y / ndc.w, ndz = ndc.z / ndc.w;
  float3 r = make_float3((ndx * 0.5f + 0.5f) * idims.x, (ndy * 0.5f + 0.5f) * idims.y, (ndz * 0.5f + 0.5f) * idims.z);
  if (cm == CM_LINDEPTH_REVERSE) {
    float z = -v.p32 / (ndz + v.p22);
    float t = -(z + v.nearp) / (v.farp - v.nearp);
    r.z = t * idims.z;
  }
  return r;
}

__device__ __forceinline__ float len3f(float a, float b, float c) { return sqrtf(a * a + b * b + c * c); }

// Position and LoD of one output cell (forward differences; Jacobian rows for
// the forward modes, columns for the reverse modes).
__device__ __forceinline__ float3 sample_pos(const SampleView& v, int cm, bool lod, float qx, float qy, float qz,
                                             float3 oinv, float3 idims, float& lod_out) {
  float3 p = map_pos(v, cm, qx, qy, qz, oinv, idims);
  lod_out = 0.f;
  if (lod) {
    float3 a = map_pos(v, cm, qx + 1.f, qy, qz, oinv, idims);
    float3 b = map_pos(v, cm, qx, qy + 1.f, qz, oinv, idims);
    float3 c = map_pos(v, cm, qx, qy, qz + 1.f, oinv, idims);
    float3 dx = make_float3(p.x - a.x, p.y - a.y, p.z - a.z);
    float3 dy = make_float3(p.x - b.x, p.y - b.y, p.z - b.z);
    float3 dz = make_float3(p.x - c.x, p.y - c.y, p.z - c.z);
    float fx, fy, fz;
    if (cm == CM_LINDEPTH || cm == CM_TRANSFORM) {
      fx = len3f(dx.x, dy.x, dz.x);
      fy = len3f(dx.y, dy.y, dz.y);
      fz = len3f(dx.z, dy.z, dz.z);
    } else {
      fx = len3f(dx.x, dx.y, dx.z);
      fy = len3f(dy.x, dy.y, dy.z);
      fz = len3f(dz.x, dz.y, dz.z);
    }
    lod_out = fmaxf(0.f, log2f(fmaxf(fx, fmaxf(fy, fz))));
  }
  return p;
}

}  // namespace gft
