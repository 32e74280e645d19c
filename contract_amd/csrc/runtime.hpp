// runtime.hpp -- error handling, device queries and argument validation shared
// by the C-ABI entry points and the extension API.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace contract_amd
{

// The drop-in ABI returns void and terminates the process on a runtime error,
// as the reference does (lib/target/runtime_api_hip.cpp:5-12).
#define CONTRACT_HIP_CHECK(call)                                                                                       \
  do {                                                                                                                 \
    hipError_t _e = (call);                                                                                            \
    if (_e != hipSuccess) {                                                                                            \
      std::fprintf(stderr, "HIP error in %s at line %d: %s\n", __FILE__, __LINE__, hipGetErrorString(_e));            \
      std::exit(_e);                                                                                                   \
    }                                                                                                                  \
  } while (0)

// Validation failures of the ABI print and exit(-1), as the reference's
// launchers do (lib/meson/meson.cu:8-11, lib/baryon/baryon.cu:43-46,61-64).
[[noreturn]] inline void abi_fail(const char *msg, long value)
{
  std::fprintf(stderr, "Error: %s %ld\n", msg, value);
  std::exit(-1);
}

// The reference requires volume % 4 == 0 (4 sites per 64-thread block,
// kernel.cuh:175, meson.cu:8-11) and exits otherwise; the drop-in keeps that
// contract even though these kernels handle ragged tails.
inline void abi_check_volume(size_t volume)
{
  if (volume % 4 != 0) {
    std::fprintf(stderr, "Error: Volume must be a multiple of TILES_PER_BLOCK\n");
    std::exit(-1);
  }
}

inline void abi_check_gamma(int g, const char *name)
{
  if (g < 0 || g > 15) {
    std::fprintf(stderr, "Error: Invalid %s value %d\n", name, g);
    std::exit(-1);
  }
}

int device_cu_count();

// Extension-API error codes (include/contract_ext.h).
enum : int {
  CONTRACT_OK = 0,
  CONTRACT_ERR_ARG = 1,
  CONTRACT_ERR_HIP = 2,
  CONTRACT_ERR_WORKSPACE = 3,
};

void set_last_error(const char *fmt, ...);
// Resolve the current device's scratch of the no-allocation entry point.
int noalloc_setup();
// Stream of the calling thread's drop-in ABI launches (contract_set_stream;
// the null stream unless set).
hipStream_t abi_stream();
// Record the calling thread's kernel-timer events (contract_set_kernel_timer).
hipError_t timer_begin(hipStream_t stream);
hipError_t timer_end(hipStream_t stream);

#define CONTRACT_EXT_HIP(call)                                                                                         \
  do {                                                                                                                 \
    hipError_t _e = (call);                                                                                            \
    if (_e != hipSuccess) {                                                                                            \
      ::contract_amd::set_last_error("HIP error %s at %s:%d", hipGetErrorString(_e), __FILE__, __LINE__);            \
      return ::contract_amd::CONTRACT_ERR_HIP;                                                                        \
    }                                                                                                                  \
  } while (0)

} // namespace contract_amd
