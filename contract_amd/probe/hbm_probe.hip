// hbm_probe.hip -- HBM ceilings measured on the running device, for the
// benchmark's roofline report (not part of libcontract.so).
//
// probe_read streams a buffer with 16-byte non-temporal loads and folds it
// into one word per workgroup, so nothing but the loads costs bandwidth;
// probe_copy is a 16-byte copy, the same access shape as a propagator-in /
// propagator-out kernel at 1:1.  Both take the launch shape (workgroups of
// 256 threads, loads in flight per thread): the rate a shape reaches depends
// on it (e.g. grid strides of 2, 4 or 8 MiB with several loads in flight
// per thread lose up to a fifth of the bandwidth), so the ceiling is the
// best of a few shapes.

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace
{

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int THREADS = 256;
constexpr int MAX_BLOCKS = 16384;

template <int UNROLL>
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

template <int UNROLL>
__global__ __launch_bounds__(THREADS) void copy_kernel(u32x4 *__restrict__ dst, const u32x4 *__restrict__ src,
                                                       size_t n16)
{
  const size_t stride = (size_t)gridDim.x * THREADS;
  size_t i = (size_t)blockIdx.x * THREADS + threadIdx.x;
  for (; i + (UNROLL - 1) * stride < n16; i += UNROLL * stride) {
    u32x4 v[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) __builtin_nontemporal_store(v[u], dst + i + u * stride);
  }
  for (; i < n16; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

bool shape_ok(int blocks, int unroll)
{
  return blocks >= 1 && blocks <= MAX_BLOCKS && (unroll == 1 || unroll == 2 || unroll == 4 || unroll == 8);
}

} // namespace

extern "C" {

// Largest workgroup count of a probe launch (sink holds one unsigned per
// workgroup).
int probe_read_blocks(void) { return MAX_BLOCKS; }

// Read `bytes` (a multiple of 16) from src with `blocks` workgroups and
// `unroll` (1, 2, 4, 8) loads in flight per thread; returns a hipError_t
// (hipErrorInvalidValue for a shape out of range).
int probe_read_shape(const void *src, size_t bytes, void *sink, void *stream, int blocks, int unroll)
{
  if (!shape_ok(blocks, unroll)) return (int)hipErrorInvalidValue;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const u32x4 *s = static_cast<const u32x4 *>(src);
  unsigned *k = static_cast<unsigned *>(sink);
  switch (unroll) {
  case 1: hipLaunchKernelGGL(read_kernel<1>, dim3(blocks), dim3(THREADS), 0, st, s, bytes / 16, k); break;
  case 2: hipLaunchKernelGGL(read_kernel<2>, dim3(blocks), dim3(THREADS), 0, st, s, bytes / 16, k); break;
  case 4: hipLaunchKernelGGL(read_kernel<4>, dim3(blocks), dim3(THREADS), 0, st, s, bytes / 16, k); break;
  default: hipLaunchKernelGGL(read_kernel<8>, dim3(blocks), dim3(THREADS), 0, st, s, bytes / 16, k); break;
  }
  return (int)hipGetLastError();
}

// Copy `bytes` (a multiple of 16) from src to dst, shape as probe_read_shape.
int probe_copy_shape(void *dst, const void *src, size_t bytes, void *stream, int blocks, int unroll)
{
  if (!shape_ok(blocks, unroll)) return (int)hipErrorInvalidValue;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  u32x4 *d = static_cast<u32x4 *>(dst);
  const u32x4 *s = static_cast<const u32x4 *>(src);
  switch (unroll) {
  case 1: hipLaunchKernelGGL(copy_kernel<1>, dim3(blocks), dim3(THREADS), 0, st, d, s, bytes / 16); break;
  case 2: hipLaunchKernelGGL(copy_kernel<2>, dim3(blocks), dim3(THREADS), 0, st, d, s, bytes / 16); break;
  case 4: hipLaunchKernelGGL(copy_kernel<4>, dim3(blocks), dim3(THREADS), 0, st, d, s, bytes / 16); break;
  default: hipLaunchKernelGGL(copy_kernel<8>, dim3(blocks), dim3(THREADS), 0, st, d, s, bytes / 16); break;
  }
  return (int)hipGetLastError();
}

// The earlier fixed shapes: 16384 workgroups, eight loads in flight (read)
// or one (copy).
int probe_read(const void *src, size_t bytes, void *sink, void *stream)
{
  return probe_read_shape(src, bytes, sink, stream, MAX_BLOCKS, 8);
}

int probe_copy(void *dst, const void *src, size_t bytes, void *stream)
{
  return probe_copy_shape(dst, src, bytes, stream, MAX_BLOCKS, 1);
}

} // extern "C"
