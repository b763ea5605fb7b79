// Internal helpers shared by the nbx translation units: status/error plumbing
// and small device utilities. Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "nbx.h"

namespace nbx {

// Records a message in the calling thread's last-error slot and returns status.
int fail(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

inline hipStream_t as_stream(nbx_stream s) { return reinterpret_cast<hipStream_t>(s); }

#define NBX_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::nbx::fail(NBX_ERR_HIP, "HIP error: %s at %s:%d", hipGetErrorString(e_),     \
                         __FILE__, __LINE__);                                              \
  } while (0)

// Checks the launch that was just queued (hipGetLastError, like CUDA_CHECK_KERNEL in
// release builds of the reference, include/nbody/error_handling.hpp:122-129).
#define NBX_LAUNCHED()                                                                     \
  do {                                                                                     \
    hipError_t e_ = hipGetLastError();                                                     \
    if (e_ != hipSuccess)                                                                  \
      return ::nbx::fail(NBX_ERR_HIP, "kernel launch failed: %s at %s:%d",                 \
                         hipGetErrorString(e_), __FILE__, __LINE__);                       \
  } while (0)

#define NBX_TRY(expr)                                                                      \
  do {                                                                                     \
    int rc_ = (expr);                                                                      \
    if (rc_ != NBX_OK) return rc_;                                                         \
  } while (0)

inline int check_particles(const nbx_particles* d, const char* who) {
  if (!d) return fail(NBX_ERR_INVALID, "%s: null particle struct", who);
  if (d->count == 0) return fail(NBX_ERR_INVALID, "%s: particle count must be > 0", who);
  if (d->count > (size_t)INT32_MAX)
    return fail(NBX_ERR_INVALID, "%s: particle count %zu exceeds 2^31-1", who, d->count);
  return NBX_OK;
}

// Device scratch that grows on demand and is reused across calls (one per
// purpose). Growing reallocates; calls that overlap on different streams must
// not share one.
struct Scratch {
  void* ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t need);
  ~Scratch();
};

// IEEE single-rounded add / mul that the compiler may not fuse into an FMA.
// (__fadd_rn / __fmul_rn are plain `+` / `*` in this HIP and get contracted.)
__device__ __forceinline__ float add_rn(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ float sub_rn(float a, float b) {
#pragma clang fp contract(off)
  return a - b;
}
__device__ __forceinline__ float mul_rn(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ float div_rn(float a, float b) {
#pragma clang fp contract(off)
  return a / b;
}

inline unsigned div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// AND of p over a one-dimensional workgroup of W waves (64 * W threads): a
// ballot per wave and, for W > 1, one LDS word per wave between two barriers.
// __syncthreads_and forms a linear thread id from the y and z workitem ids,
// which then occupy two VGPRs for the rest of a kernel that loops over tiles.
template <int W>
__device__ __forceinline__ bool block_all(bool p) {
  const bool wave_all = __ballot(!p) == 0;
  if (W == 1) return wave_all;
  __shared__ int votes[W];
  if ((threadIdx.x & 63) == 0) votes[threadIdx.x >> 6] = wave_all;
  __syncthreads();
  bool all = true;
#pragma unroll
  for (int w = 0; w < W; ++w) all = all && votes[w] != 0;
  __syncthreads();  // the words are rewritten by the next call
  return all;
}

// The self pair contributes exactly zero (dx = 0, finite f) when eps2 keeps
// rsq(eps2)^3 finite for every mass in use; below this the j == i test runs.
constexpr float kUncheckedMinEps2 = 1e-12f;

}  // namespace nbx
