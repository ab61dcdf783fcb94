// Corner-quad copies of scalar grids (the layout the fused render and the
// batched mipmapped step read their trilinear taps from).
#pragma once
#include "gft_device.hpp"

namespace gft {

// Corner-quad copy of a scalar grid for BORDER trilinear taps: entry
// (x, y, z) for x in [-1, W-1], y in [-1, H-1], z in [-1, D] holds the float4
// (v(x,y,z), v(x+1,y,z), v(x,y+1,z), v(x+1,y+1,z)) with v = 0 outside the grid
// (the BORDER value of an invalid corner). A trilinear tap is then two 16-byte
// loads (slices z and z+1) instead of eight 4-byte gathers, and needs no
// per-corner validity test.
// The entries are stored in bricks of kQBX x kQBY x kQBZ = 4 x 1 x 2 (x fastest
// inside a brick, bricks x fastest): one brick is one 128-byte cache line, and
// the two slices a tap reads share it whenever z is even. A wave's taps at one
// step lie on a slanted patch of the grid for views that orbit in the xz-plane;
// in an x-fastest layout nearly every lane then touches its own cache line, and
// the forward is bound by lines per load instruction rather than bytes. At c2
// the render forward takes 0.209 ms with rows (8 x 1 x 1 or 8 x 1 x 2/4 bricks)
// and 0.169-0.175 ms with 4 x 1 x 2, 2 x 1 x 4, 4 x 1 x 4, 4 x 1 x 8 or 2 x 2 x 2.
//
// With the density stored in bfloat16 the copy holds bfloat16 corners (8-byte
// entries, uint2: corners 0/1 in x, 2/3 in y, low half first) in 4 x 2 x 2
// bricks, again one 128-byte line per brick: half the bytes per tap.
template <class Q> struct QuadLayout;
template <> struct QuadLayout<float4> {
  static constexpr unsigned Xl = 2, Yl = 0, Zl = 1;  // log2 of the brick dims
};
template <> struct QuadLayout<uint2> {
  static constexpr unsigned Xl = 2, Yl = 1, Zl = 1;
};
template <class Q> struct QuadBrick {
  static constexpr unsigned Xl = QuadLayout<Q>::Xl, Yl = QuadLayout<Q>::Yl, Zl = QuadLayout<Q>::Zl;
  static constexpr unsigned X = 1u << Xl, Y = 1u << Yl, Z = 1u << Zl, N = X * Y * Z;
};
struct QuadDims {
  unsigned qx, qy, qz;     // entries: W+1, H+1, D+2
  unsigned nbx, nby, nbz;  // bricks along each axis
};
template <class Q = float4> __host__ __device__ inline QuadDims make_quad_dims(Dims3 d) {
  using B = QuadBrick<Q>;
  const unsigned qx = (unsigned)d.x + 1u, qy = (unsigned)d.y + 1u, qz = (unsigned)d.z + 2u;
  return QuadDims{qx, qy, qz, (qx + B::X - 1) / B::X, (qy + B::Y - 1) / B::Y, (qz + B::Z - 1) / B::Z};
}
template <class Q = float4> __host__ __device__ inline size_t quad_entries(const QuadDims& q) {
  return (size_t)q.nbx * q.nby * q.nbz * QuadBrick<Q>::N;
}
// index of entry (ix, iy, iz) (0-based: ix = x + 1, iy = y + 1, iz = z + 1)
template <class Q = float4>
__device__ __forceinline__ unsigned quad_index(const QuadDims& q, unsigned ix, unsigned iy, unsigned iz) {
  using B = QuadBrick<Q>;
  const unsigned b = ((iz >> B::Zl) * q.nby + (iy >> B::Yl)) * q.nbx + (ix >> B::Xl);
  const unsigned o = ((((iz & (B::Z - 1)) << B::Yl) | (iy & (B::Y - 1))) << B::Xl) | (ix & (B::X - 1));
  return b * B::N + o;
}

// float -> bfloat16 bits, rounded to nearest even (finite values)
__device__ __forceinline__ unsigned bf16_bits(float f) {
  const unsigned u = __float_as_uint(f);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float4 quad_value(const float4& e) { return e; }
__device__ __forceinline__ float4 quad_value(const uint2& e) {
  return make_float4(__uint_as_float(e.x << 16), __uint_as_float(e.x & 0xffff0000u), __uint_as_float(e.y << 16),
                     __uint_as_float(e.y & 0xffff0000u));
}
__device__ __forceinline__ void quad_store(float4* p, float a, float b, float c, float d) { *p = make_float4(a, b, c, d); }
__device__ __forceinline__ void quad_store(uint2* p, float a, float b, float c, float d) {
  *p = make_uint2(bf16_bits(a) | (bf16_bits(b) << 16), bf16_bits(c) | (bf16_bits(d) << 16));
}

}  // namespace gft
