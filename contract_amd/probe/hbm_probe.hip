// hbm_probe.hip -- HBM ceilings measured on the running device, for the
// benchmark's roofline report (not part of libcontract.so).
//
// probe_read streams a buffer with 16-byte non-temporal loads (eight in
// flight per thread) and folds it into one word per workgroup, so nothing
// but the loads costs bandwidth; probe_copy is a plain 16-byte copy, the
// same access shape as a propagator-in / propagator-out kernel at 1:1.

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace
{

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int THREADS = 256;
constexpr int UNROLL = 8;

__global__ __launch_bounds__(THREADS) void read_kernel(const u32x4 *__restrict__ src, size_t n16,
                                                       unsigned *__restrict__ sink)
{
  const size_t stride = (size_t)gridDim.x * THREADS;
  size_t i = (size_t)blockIdx.x * THREADS + threadIdx.x;
  u32x4 acc = {0u, 0u, 0u, 0u};
  for (; i + (UNROLL - 1) * stride < n16; i += UNROLL * stride) {
    u32x4 v[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) acc ^= v[u];
  }
  for (; i < n16; i += stride) acc ^= __builtin_nontemporal_load(src + i);
  __shared__ unsigned red[THREADS];
  red[threadIdx.x] = acc.x ^ acc.y ^ acc.z ^ acc.w;
  __syncthreads();
  for (int s = THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] ^= red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) sink[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(THREADS) void copy_kernel(u32x4 *__restrict__ dst, const u32x4 *__restrict__ src,
                                                       size_t n16)
{
  const size_t stride = (size_t)gridDim.x * THREADS;
  for (size_t i = (size_t)blockIdx.x * THREADS + threadIdx.x; i < n16; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

} // namespace

extern "C" {

// Workgroups of the read probe (sink holds one unsigned per workgroup).
int probe_read_blocks(void) { return 16384; }

// Read `bytes` (a multiple of 16) from src; returns a hipError_t.
int probe_read(const void *src, size_t bytes, void *sink, void *stream)
{
  hipLaunchKernelGGL(read_kernel, dim3(probe_read_blocks()), dim3(THREADS), 0, static_cast<hipStream_t>(stream),
                     static_cast<const u32x4 *>(src), bytes / 16, static_cast<unsigned *>(sink));
  return (int)hipGetLastError();
}

// Copy `bytes` (a multiple of 16) from src to dst; returns a hipError_t.
int probe_copy(void *dst, const void *src, size_t bytes, void *stream)
{
  hipLaunchKernelGGL(copy_kernel, dim3(16384), dim3(THREADS), 0, static_cast<hipStream_t>(stream),
                     static_cast<u32x4 *>(dst), static_cast<const u32x4 *>(src), bytes / 16);
  return (int)hipGetLastError();
}

} // extern "C"
