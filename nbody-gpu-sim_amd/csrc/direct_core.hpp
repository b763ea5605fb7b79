// Device core of the direct N^2 force evaluation, shared by the single-system
// kernel (direct.hip) and the batched-ensemble kernel (ensemble.hip) so both
// produce bit-identical accelerations for the same tile shape.
//
// Reference semantics (src/cuda/force_direct.cu:10-85): for each target i,
//   a_i = sum_{j != i} G m_j r_ij * rsqrt(|r_ij|^2 + eps2)^3,   r_ij = x_j - x_i.
//
// Mapping to CDNA4:
//  * j bodies are read with wave-uniform addresses, so the compiler turns them
//    into scalar loads (one s_load_dwordx8 per SoA array for a group of 8
//    bodies). Each body is broadcast to all 64 lanes through SGPR operands: no
//    LDS traffic and no barriers in the inner loop.
//  * i targets live in VGPRs, IPL per lane, so one scalar batch feeds
//    64*IPL interactions.
//  * pair arithmetic is FP32 with the hardware v_rsq_f32; 13 VALU ops per
//    interaction (3 sub, 3 fma, rsq, 3 mul, 3 fma), issued in stages over
//    kChains independent interactions (direct_group).
//  * each lane sums 64 bodies in FP32, then folds the partial into an FP64
//    accumulator: the rounding error grows with the 64-term block, not with N.
//    When the 64 bodies of a fold share one mass (the reference's initialisers
//    default to min_mass == max_mass, include/nbody/types.hpp:157-173), the
//    mass multiplies the fold's FP64 partial once instead of every pair.
//  * the j range of one i tile is split across WS waves of the workgroup; the
//    WS FP64 partials are added in wave order through LDS, so results are
//    deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace nbx {

constexpr int kFold = 64;   // bodies per FP32 partial before the FP64 fold

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

// NB bodies of the j range, held in SGPRs (wave-uniform loads).
template <int NB>
struct BodyGroup {
  float x[NB], y[NB], z[NB], m[NB];
};

template <int NB>
__device__ __forceinline__ void load_group(const float* __restrict__ x,
                                           const float* __restrict__ y,
                                           const float* __restrict__ z,
                                           const float* __restrict__ m, int jb,
                                           BodyGroup<NB>& g) {
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    g.x[b] = x[jb + b];
    g.y[b] = y[jb + b];
    g.z[b] = z[jb + b];
    g.m[b] = m[jb + b];
  }
}

// NB bodies x IPL targets, issued stage by stage: all differences, all r^2,
// all rsq, all force factors, then the accumulations (NB*IPL independent
// interactions in flight per lane). v_rsq_f32 runs at a quarter of the FP32 FMA
// rate and does not overlap with other VALU work; issuing each stage over
// independent chains keeps the transcendental ops back to back (no dependent
// use straight after an rsq, so no hazard wait states) and lets every other
// op issue without waiting on its producer. The loads of the following group
// (nxt_j) are issued after the first stage, once this group's operands have
// arrived: scalar loads return out of order, so a prefetch issued earlier
// would be waited for together with the current group.
// Arithmetic and the per-target accumulation order (body by body) are those
// of direct_pair(), so results are identical to an unstaged loop. UNIFORM
// (every body of the group has the same mass) leaves the mass out of the
// force factor, f = inv^3 (two multiplies instead of three); the caller scales
// the fold's sum by that mass.
template <int IPL, int NB, bool CHECKED, bool UNIFORM>
__device__ __forceinline__ void direct_group(const BodyGroup<NB>& cur, int jb,
                                             const float* __restrict__ x,
                                             const float* __restrict__ y,
                                             const float* __restrict__ z,
                                             const float* __restrict__ m, int nxt_j,
                                             BodyGroup<NB>& nxt, const float (&xi)[IPL],
                                             const float (&yi)[IPL], const float (&zi)[IPL],
                                             const int (&ii)[IPL], float eps2, float (&fx)[IPL],
                                             float (&fy)[IPL], float (&fz)[IPL]) {
  float dx[NB][IPL], dy[NB][IPL], dz[NB][IPL], w[NB][IPL];
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
      dx[b][q] = cur.x[b] - xi[q];
      dy[b][q] = cur.y[b] - yi[q];
      dz[b][q] = cur.z[b] - zi[q];
    }
  __builtin_amdgcn_sched_barrier(0);
  load_group<NB>(x, y, z, m, nxt_j, nxt);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
      float r2 = __builtin_fmaf(dx[b][q], dx[b][q], eps2);
      r2 = __builtin_fmaf(dy[b][q], dy[b][q], r2);
      w[b][q] = __builtin_fmaf(dz[b][q], dz[b][q], r2);
    }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int q = 0; q < IPL; ++q) w[b][q] = __builtin_amdgcn_rsqf(w[b][q]);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
      const float inv = w[b][q];
      float f = UNIFORM ? (inv * inv) * inv : (cur.m[b] * inv) * (inv * inv);
      if (CHECKED) f = (jb + b == ii[q]) ? 0.0f : f;  // reference skips j == i by index (:60)
      w[b][q] = f;
    }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
      fx[q] = __builtin_fmaf(w[b][q], dx[b][q], fx[q]);
      fy[q] = __builtin_fmaf(w[b][q], dy[b][q], fy[q]);
      fz[q] = __builtin_fmaf(w[b][q], dz[b][q], fz[q]);
    }
  __builtin_amdgcn_sched_barrier(0);
}

// One fold of kFold bodies starting at j: the FP32 partial of the fold's
// groups is added to the FP64 sums, scaled by mw (the shared mass) when
// UNIFORM. ga holds the fold's first group on entry and the next fold's
// first group on exit (re-reading the last group when j + kFold == jfull).
template <int IPL, int NB, bool CHECKED, bool UNIFORM>
__device__ __forceinline__ void direct_fold(const float* __restrict__ x,
                                            const float* __restrict__ y,
                                            const float* __restrict__ z,
                                            const float* __restrict__ m, int j, int jfull,
                                            BodyGroup<NB>& ga, BodyGroup<NB>& gb,
                                            const float (&xi)[IPL], const float (&yi)[IPL],
                                            const float (&zi)[IPL], const int (&ii)[IPL],
                                            float eps2, double mw, double (&sx)[IPL],
                                            double (&sy)[IPL], double (&sz)[IPL]) {
  float fx[IPL], fy[IPL], fz[IPL];
#pragma unroll
  for (int q = 0; q < IPL; ++q) fx[q] = fy[q] = fz[q] = 0.0f;
#pragma unroll
  for (int u = 0; u < kFold; u += 2 * NB) {
    const int jb = j + u;
    int jn2 = jb + 2 * NB;
    if (jn2 >= jfull) jn2 = jb + NB;
    direct_group<IPL, NB, CHECKED, UNIFORM>(ga, jb, x, y, z, m, jb + NB, gb, xi, yi, zi, ii, eps2,
                                            fx, fy, fz);
    direct_group<IPL, NB, CHECKED, UNIFORM>(gb, jb + NB, x, y, z, m, jn2, ga, xi, yi, zi, ii,
                                            eps2, fx, fy, fz);
  }
#pragma unroll
  for (int q = 0; q < IPL; ++q) {
    if (UNIFORM) {
      sx[q] += (double)fx[q] * mw;
      sy[q] += (double)fy[q] * mw;
      sz[q] += (double)fz[q] * mw;
    } else {
      sx[q] += (double)fx[q];
      sy[q] += (double)fy[q];
      sz[q] += (double)fz[q];
    }
  }
}

// Accumulates bodies [j0, j1) (wave-uniform bounds) into the lane's IPL targets.
// x/y/z/m point at the system's first body; i indices are system-local.
// CHECKED selects the j == i test, needed only when eps2 is too small for the
// self pair to evaluate to an exact zero.
template <int IPL, int NB, bool CHECKED>
__device__ __forceinline__ void direct_accumulate(const float* __restrict__ x,
                                                  const float* __restrict__ y,
                                                  const float* __restrict__ z,
                                                  const float* __restrict__ m, int j0, int j1,
                                                  const float (&xi)[IPL], const float (&yi)[IPL],
                                                  const float (&zi)[IPL], const int (&ii)[IPL],
                                                  float eps2, double (&sx)[IPL],
                                                  double (&sy)[IPL], double (&sz)[IPL]) {
  static_assert(NB >= 1 && kFold % (2 * NB) == 0, "stage groups must pair up within a fold");
  const int lane = threadIdx.x & 63;
  int j = j0;
  const int jfull = j0 + ((j1 - j0) / kFold) * kFold;  // end of the whole folds
  if (j < jfull) {
    // Groups alternate between two SGPR sets (a, b): group g+1 is loaded
    // while group g computes. The last prefetch re-reads the final group
    // instead of running past jfull.
    BodyGroup<NB> ga, gb;
    load_group<NB>(x, y, z, m, j, ga);
    // When every body of the wave's whole folds shares one mass, the folds
    // take the uniform path (mass applied once per fold). The scan reads the
    // j range once, 64 masses per load, against about 64*IPL*12 VALU ops per
    // body in the folds themselves.
    // Sixteen loads are in flight per step; indices past jfull are clamped
    // onto the last fold (in bounds, already compared).
    const float mref = m[j];
    bool same = true;
    for (int k = j; k < jfull; k += 16 * 64) {
      float mv[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) mv[t] = m[min(k + t * 64, jfull - 64) + lane];
#pragma unroll
      for (int t = 0; t < 16; ++t) same &= (mv[t] == mref);
    }
    const bool uniform = __ballot(!same) == 0ull;
    if (uniform) {
      for (; j < jfull; j += kFold)
        direct_fold<IPL, NB, CHECKED, true>(x, y, z, m, j, jfull, ga, gb, xi, yi, zi, ii, eps2,
                                            (double)mref, sx, sy, sz);
    } else {
      for (; j < jfull; j += kFold)
        direct_fold<IPL, NB, CHECKED, false>(x, y, z, m, j, jfull, ga, gb, xi, yi, zi, ii, eps2,
                                             1.0, sx, sy, sz);
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

// One workgroup: WS waves, stage groups of NB bodies, tile of 64*IPL targets [ibase, ibase + 64*IPL) of a
// system of n bodies at x..m. Writes acc (and optionally applies the kick
// v += (a_old + a_new)*dt_half of integrator.cu:25-37 with IEEE rounding).
// red: LDS scratch of WS*64*IPL*3 doubles.
template <int IPL, int NB, int WS, bool CHECKED, bool KICK>
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
  direct_accumulate<IPL, NB, CHECKED>(x, y, z, m, j0, j1, xi, yi, zi, ii, eps2, sx, sy, sz);

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

// Tile shapes X(IPL, NB, WS) of the direct kernel: IPL targets per lane, NB
// bodies per stage group, WS waves splitting the j range of one tile. The
// single-system and batched kernels instantiate this list in this order, so a
// variant index names the same shape in both.
#ifndef NBX_DIRECT_VARIANTS
#define NBX_DIRECT_VARIANTS(X)                                                            \
  X(1, 4, 4) X(1, 4, 8) X(1, 8, 4) X(1, 8, 8) X(1, 8, 16) X(1, 16, 8) X(2, 4, 8) X(4, 2, 4)
#endif

}  // namespace nbx
