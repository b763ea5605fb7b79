// Device core of the direct N^2 force evaluation, shared by the single-system
// kernel (direct.hip) and the batched-ensemble kernel (ensemble.hip) so both
// produce bit-identical accelerations for the same tile shape.
//
// Reference semantics (src/cuda/force_direct.cu:10-85): for each target i,
//   a_i = sum_{j != i} G m_j r_ij * rsqrt(|r_ij|^2 + eps2)^3,   r_ij = x_j - x_i.
//
// Mapping to CDNA4:
//  * j bodies are read with wave-uniform addresses, so the compiler turns them
//    into scalar loads (s_load_dwordx16 of 16 bodies per SoA array). Each body
//    is broadcast to all 64 lanes through SGPR operands: no LDS traffic and no
//    barriers in the inner loop.
//  * i targets live in VGPRs, IPL per lane, so one scalar batch feeds
//    64*IPL interactions.
//  * pair arithmetic is FP32 with the hardware v_rsq_f32; 13 VALU ops per
//    interaction (3 sub, 3 fma, rsq, 3 mul, 3 fma).
//  * each lane sums 64 bodies in FP32, then folds the partial into an FP64
//    accumulator: the rounding error grows with the 64-term block, not with N.
//  * the j range of one i tile is split across WS waves of the workgroup; the
//    WS FP64 partials are added in wave order through LDS, so results are
//    deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace nbx {

constexpr int kFold = 64;   // bodies per FP32 partial before the FP64 fold
constexpr int kBatch = 16;  // bodies per scalar-load batch (s_load_dwordx16)

// Contiguous j chunk of wave w when [0, n) is split over ws waves at multiples of kFold.
__host__ __device__ inline int direct_chunk(int n, int ws) {
  const int folds = (n + kFold - 1) / kFold;
  return ((folds + ws - 1) / ws) * kFold;
}

template <bool CHECKED>
__device__ __forceinline__ void direct_pair(float bx, float by, float bz, float bm, int j, float xi,
                                            float yi, float zi, int i, float eps2, float& fx,
                                            float& fy, float& fz) {
  const float dx = bx - xi;
  const float dy = by - yi;
  const float dz = bz - zi;
  float r2 = __builtin_fmaf(dx, dx, eps2);
  r2 = __builtin_fmaf(dy, dy, r2);
  r2 = __builtin_fmaf(dz, dz, r2);
  const float inv = __builtin_amdgcn_rsqf(r2);
  const float mi = bm * inv;
  const float inv2 = inv * inv;
  float f = mi * inv2;
  if (CHECKED) f = (j == i) ? 0.0f : f;  // reference skips j == i by index (:60)
  fx = __builtin_fmaf(f, dx, fx);
  fy = __builtin_fmaf(f, dy, fy);
  fz = __builtin_fmaf(f, dz, fz);
}

// Accumulates bodies [j0, j1) (wave-uniform bounds) into the lane's IPL targets.
// x/y/z/m point at the system's first body; i indices are system-local.
// CHECKED selects the j == i test, needed only when eps2 is too small for the
// self pair to evaluate to an exact zero.
template <int IPL, bool CHECKED>
__device__ __forceinline__ void direct_accumulate(const float* __restrict__ x,
                                                  const float* __restrict__ y,
                                                  const float* __restrict__ z,
                                                  const float* __restrict__ m, int j0, int j1,
                                                  const float (&xi)[IPL], const float (&yi)[IPL],
                                                  const float (&zi)[IPL], const int (&ii)[IPL],
                                                  float eps2, double (&sx)[IPL],
                                                  double (&sy)[IPL], double (&sz)[IPL]) {
  int j = j0;
  for (; j + kFold <= j1; j += kFold) {
    float fx[IPL], fy[IPL], fz[IPL];
#pragma unroll
    for (int q = 0; q < IPL; ++q) fx[q] = fy[q] = fz[q] = 0.0f;
#pragma unroll 1
    for (int u = 0; u < kFold; u += kBatch) {
      const int jb = j + u;
#pragma unroll
      for (int k = 0; k < kBatch; ++k) {
        const float bx = x[jb + k], by = y[jb + k], bz = z[jb + k], bm = m[jb + k];
#pragma unroll
        for (int q = 0; q < IPL; ++q)
          direct_pair<CHECKED>(bx, by, bz, bm, jb + k, xi[q], yi[q], zi[q], ii[q], eps2, fx[q],
                               fy[q], fz[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
      sx[q] += (double)fx[q];
      sy[q] += (double)fy[q];
      sz[q] += (double)fz[q];
    }
  }
  if (j < j1) {  // tail shorter than one fold
    float fx[IPL], fy[IPL], fz[IPL];
#pragma unroll
    for (int q = 0; q < IPL; ++q) fx[q] = fy[q] = fz[q] = 0.0f;
    for (; j < j1; ++j) {
      const float bx = x[j], by = y[j], bz = z[j], bm = m[j];
#pragma unroll
      for (int q = 0; q < IPL; ++q)
        direct_pair<CHECKED>(bx, by, bz, bm, j, xi[q], yi[q], zi[q], ii[q], eps2, fx[q], fy[q],
                             fz[q]);
    }
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
      sx[q] += (double)fx[q];
      sy[q] += (double)fy[q];
      sz[q] += (double)fz[q];
    }
  }
}

// Epilogue options of the force tile.
struct KickArgs {
  float* vx;
  float* vy;
  float* vz;
  const float* aox;
  const float* aoy;
  const float* aoz;
  float dt_half;  // 0.5f*dt, as integrator.cu:32 forms it
};

// One workgroup: WS waves, tile of 64*IPL targets [ibase, ibase + 64*IPL) of a
// system of n bodies at x..m. Writes acc (and optionally applies the kick
// v += (a_old + a_new)*dt_half of integrator.cu:25-37 with IEEE rounding).
// red: LDS scratch of WS*64*IPL*3 doubles.
template <int IPL, int WS, bool CHECKED, bool KICK>
__device__ __forceinline__ void direct_tile(const float* __restrict__ x,
                                            const float* __restrict__ y,
                                            const float* __restrict__ z,
                                            const float* __restrict__ m, int n, int ibase,
                                            int iend, double G, float eps2, float* ax, float* ay,
                                            float* az, const KickArgs& kick, double* red) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float xi[IPL], yi[IPL], zi[IPL];
  int ii[IPL];
#pragma unroll
  for (int q = 0; q < IPL; ++q) {
    const int i = ibase + q * 64 + lane;
    ii[q] = i;
    const bool ok = i < iend;
    xi[q] = ok ? x[i] : 0.0f;
    yi[q] = ok ? y[i] : 0.0f;
    zi[q] = ok ? z[i] : 0.0f;
  }
  double sx[IPL], sy[IPL], sz[IPL];
#pragma unroll
  for (int q = 0; q < IPL; ++q) sx[q] = sy[q] = sz[q] = 0.0;

  const int chunk = direct_chunk(n, WS);
  const int j0 = min(n, wave * chunk);
  const int j1 = min(n, j0 + chunk);
  direct_accumulate<IPL, CHECKED>(x, y, z, m, j0, j1, xi, yi, zi, ii, eps2, sx, sy, sz);

  constexpr int T = 64 * IPL;
#pragma unroll
  for (int q = 0; q < IPL; ++q) {
    const int s = q * 64 + lane;
    red[(wave * 3 + 0) * T + s] = sx[q];
    red[(wave * 3 + 1) * T + s] = sy[q];
    red[(wave * 3 + 2) * T + s] = sz[q];
  }
  __syncthreads();
  for (int s = threadIdx.x; s < T; s += 64 * WS) {
    const int i = ibase + s;
    if (i >= iend) continue;
    double tx = 0.0, ty = 0.0, tz = 0.0;
#pragma unroll
    for (int w = 0; w < WS; ++w) {
      tx += red[(w * 3 + 0) * T + s];
      ty += red[(w * 3 + 1) * T + s];
      tz += red[(w * 3 + 2) * T + s];
    }
    const float a0 = (float)(G * tx), a1 = (float)(G * ty), a2 = (float)(G * tz);
    ax[i] = a0;
    ay[i] = a1;
    az[i] = a2;
    if (KICK) {
      kick.vx[i] = add_rn(kick.vx[i], mul_rn(add_rn(kick.aox[i], a0), kick.dt_half));
      kick.vy[i] = add_rn(kick.vy[i], mul_rn(add_rn(kick.aoy[i], a1), kick.dt_half));
      kick.vz[i] = add_rn(kick.vz[i], mul_rn(add_rn(kick.aoz[i], a2), kick.dt_half));
    }
  }
  __syncthreads();  // red is reused by the next tile of a persistent workgroup
}

// The self pair contributes exactly zero (dx = 0, finite f) when eps2 keeps
// rsq(eps2)^3 finite for every mass in use; below this the j == i test runs.
constexpr float kUncheckedMinEps2 = 1e-12f;

}  // namespace nbx
