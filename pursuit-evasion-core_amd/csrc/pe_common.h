// Shared host/device helpers of the pe_amd library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "pe_amd.h"

namespace pe {

// Per-thread error text behind pe_last_error().
void set_error(const std::string &message);
void clear_error();

// Returns `code` after recording `message`.
inline int fail(int code, const std::string &message) {
  set_error(message);
  return code;
}

// Checks that a gfx950 device is current; returns 0 or PE_ERR_NODEVICE.
int require_device(int *device);

}  // namespace pe

// Evaluates a HIP call; on error records it and returns PE_ERR_HIP (or
// PE_ERR_MEMORY for allocation failures) from the enclosing function.
#define PE_HIP(call)                                                        \
  do {                                                                      \
    hipError_t pe_status_ = (call);                                         \
    if (pe_status_ != hipSuccess) {                                         \
      (void)hipGetLastError();                                              \
      return ::pe::fail(pe_status_ == hipErrorOutOfMemory ? PE_ERR_MEMORY   \
                                                          : PE_ERR_HIP,     \
                        std::string(#call) + ": " +                         \
                            hipGetErrorString(pe_status_));                 \
    }                                                                       \
  } while (0)

// Checks the launch status of the kernel just enqueued.
#define PE_LAUNCHED(name) PE_HIP(hipGetLastError())

namespace pe {

// Max-reduction of a non-negative double over a wave64.
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Raises *slot to v (v >= 0, stored as the IEEE bits of a double, whose
// unsigned order matches the numeric order for non-negative values).  Reads
// first so that the common case (no increase) issues no atomic.
__device__ inline void raise_max(unsigned long long *slot, double v) {
  if (!(v > 0.0)) return;
  const unsigned long long bits = __double_as_longlong(v);
  if (bits > __atomic_load_n(slot, __ATOMIC_RELAXED)) atomicMax(slot, bits);
}

}  // namespace pe
