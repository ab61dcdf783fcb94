// Small utility kernels. Buffers are cleared with a fill kernel rather than
// hipMemsetAsync so that the same code path can be captured in a hipGraph.
#include "gft_device.hpp"

namespace gft {

__global__ __launch_bounds__(256) void k_fill4(float* __restrict__ p, size_t n4, float v) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n4) reinterpret_cast<float4*>(p)[i] = make_float4(v, v, v, v);
}
__global__ __launch_bounds__(256) void k_fill1(float* __restrict__ p, size_t n, float v) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t fill_launch(hipStream_t s, float* p, size_t n, float v) {
  if (n == 0) return hipSuccess;
  KernelTimer t(s, "fill");
  size_t head = 0;
  if ((reinterpret_cast<uintptr_t>(p) & 15) == 0) {
    size_t n4 = n / 4;
    if (n4) {
      hipLaunchKernelGGL(k_fill4, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, s, p, n4, v);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    head = n4 * 4;
  }
  size_t rest = n - head;
  if (rest) hipLaunchKernelGGL(k_fill1, dim3((unsigned)((rest + 255) / 256)), dim3(256), 0, s, p + head, rest, v);
  return hipGetLastError();
}

}  // namespace gft
