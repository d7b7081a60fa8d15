// Plain device fill and copy kernels (the fill replaces memset nodes, which
// graphs must not contain; the copy is the known-size reference kernel for
// the HBM traffic counters).
#include "gt_launch.h"

namespace gt {

__global__ void fill_u32_kernel(uint32_t* __restrict__ d, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = i; j < n; j += stride) d[j] = v;
}

__global__ void copy16_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = i; j < n; j += stride) dst[j] = src[j];
}

hipError_t launch_fill_u32(uint32_t* d, int64_t n, uint32_t value, hipStream_t s) {
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1 << 16);
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d, n, value);
  return hipGetLastError();
}

hipError_t launch_copy16(void* dst, const void* src, int64_t bytes, hipStream_t s) {
  const int64_t n = bytes / 16;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1 << 16);
  hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                     reinterpret_cast<uint4*>(dst), reinterpret_cast<const uint4*>(src), n);
  return hipGetLastError();
}

}  // namespace gt
