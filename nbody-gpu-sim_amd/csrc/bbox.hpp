// Exact per-axis min/max of a position set, shared by the Barnes-Hut build and
// the spatial-hash build (the reference reuses one kernel for both:
// src/cuda/force_barnes_hut.cu:53-122, declared again at
// src/cuda/force_spatial_hash.cu:184-188). Starting values +-1e30
// (force_barnes_hut.cu:18-19). Min and max are exact, so the result does not
// depend on the reduction order.
#pragma once

#include "common.hpp"

namespace nbx {

// Workgroup of T threads: grid-stride min/max over [0, n), reduced across the
// workgroup; partial[6*blockIdx.x ..] = {min x, y, z, max x, y, z}.
template <int T>
__device__ __forceinline__ void bbox_partial_body(const float* __restrict__ x,
                                                  const float* __restrict__ y,
                                                  const float* __restrict__ z, int n,
                                                  float* partial) {
  float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
  for (int i = blockIdx.x * T + threadIdx.x; i < n; i += gridDim.x * T) {
    const float v[3] = {x[i], y[i], z[i]};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], v[a]);
      hi[a] = fmaxf(hi[a], v[a]);
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    for (int off = 32; off > 0; off >>= 1) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], off, 64));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off, 64));
    }
  }
  __shared__ float sh[6][T / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) {
      sh[a][w] = lo[a];
      sh[3 + a][w] = hi[a];
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    float v = sh[threadIdx.x][0];
    for (int q = 1; q < T / 64; ++q)
      v = threadIdx.x < 3 ? fminf(v, sh[threadIdx.x][q]) : fmaxf(v, sh[threadIdx.x][q]);
    partial[blockIdx.x * 6 + threadIdx.x] = v;
  }
}

// One workgroup of T threads reduces nparts partials. Returns true on thread 0,
// which then holds the box in lo / hi; every other thread returns false.
template <int T>
__device__ __forceinline__ bool bbox_final_body(const float* partial, int nparts, float (&lo)[3],
                                                float (&hi)[3]) {
  for (int a = 0; a < 3; ++a) {
    lo[a] = 1e30f;
    hi[a] = -1e30f;
  }
  for (int b = threadIdx.x; b < nparts; b += T)
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], partial[b * 6 + a]);
      hi[a] = fmaxf(hi[a], partial[b * 6 + 3 + a]);
    }
  for (int a = 0; a < 3; ++a)
    for (int off = 32; off > 0; off >>= 1) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], off, 64));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off, 64));
    }
  __shared__ float sh[6][T / 64];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) {
      sh[a][w] = lo[a];
      sh[3 + a][w] = hi[a];
    }
  __syncthreads();
  if (threadIdx.x != 0) return false;
  for (int a = 0; a < 3; ++a) {
    lo[a] = sh[a][0];
    hi[a] = sh[3 + a][0];
    for (int q = 1; q < T / 64; ++q) {
      lo[a] = fminf(lo[a], sh[a][q]);
      hi[a] = fmaxf(hi[a], sh[3 + a][q]);
    }
  }
  return true;
}

}  // namespace nbx
