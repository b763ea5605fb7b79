// Device core of the pair-symmetric direct N^2 forces, shared by the
// single-system kernel (direct_sym.hip) and the batched-ensemble kernel
// (ensemble.hip): a system evaluated on its own and inside an ensemble runs the
// same tiles in the same order, so its accelerations are bit-identical.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "direct_core.hpp"

namespace nbx {

constexpr float kPadPos = 1.0e18f;

// Rotation by one lane (lane l receives lane l-1's value) through the LDS
// crossbar: ds_bpermute_b32 moves 64 dwords without a VALU slot. (A DPP
// wave_ror:1 move does the same on the VALU; measured on MI355X, the whole
// kernel runs about 30 % faster with ds_bpermute.) src_addr = ((l + 63) & 63) * 4.
__device__ __forceinline__ float ror1(float v, int src_addr) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute(src_addr, __float_as_int(v)));
}

// One round of a system's tiles: distances d_lo .. d_hi (d_lo = 0 includes the
// diagonal). A whole evaluation is one round (d_lo = 0, d_hi = Dmax) when its
// partial sums fit the scratch, or several consecutive rounds otherwise.
struct SymGeom {
  int n;        // bodies
  int S;        // bodies per super-block
  int K;        // super-blocks
  int Dmax;     // floor(K / 2)
  int d_lo;     // first distance of the round
  int d_hi;     // last distance of the round
  long npad;    // K * S
  float* pi;    // i-side partials [d_hi - d_lo + 1][3][npad], slot d - d_lo
  float* pj;    // j-side partials [d_hi - pj_lo() + 1][3][npad], slot d - pj_lo()
  __host__ __device__ int pj_lo() const { return d_lo > 0 ? d_lo : 1; }
};

// Off-diagonal tiles of row A: distances 1 .. sym_rowlen(A, K).
__host__ __device__ inline int sym_rowlen(int A, int K) {
  if (K & 1) return (K - 1) / 2;
  return A < K / 2 ? K / 2 : K / 2 - 1;
}

// 64 rotation steps of one chunk (xj.. one body per lane) against the wave's
// IPL targets per lane. UNI: all masses equal, so the mass is left out of both
// force factors (the caller scales by it). JSIDE: also accumulate the reaction
// on the chunk's bodies (g*). CHECKED: zero the self pair, which only a
// diagonal chunk holding the targets of slot self_q has, at step 0.
template <int IPL, bool UNI, bool JSIDE, bool CHECKED>
__device__ __forceinline__ void sym_chunk(int src_addr, float& xj, float& yj, float& zj, float& mj,
                                          const float (&xi)[IPL], const float (&yi)[IPL],
                                          const float (&zi)[IPL], const float (&mi)[IPL],
                                          float eps2, int self_q, float (&fx)[IPL],
                                          float (&fy)[IPL], float (&fz)[IPL], float& gx,
                                          float& gy, float& gz) {
#pragma unroll 2
  for (int s = 0; s < 64; ++s) {
    float dx[IPL], dy[IPL], dz[IPL], w[IPL];
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
      dx[q] = xj - xi[q];
      dy[q] = yj - yi[q];
      dz[q] = zj - zi[q];
    }
    __builtin_amdgcn_sched_barrier(0);
    // The chunk's positions are dead from here on in this step: rotate them now,
    // so the LDS round trip overlaps the rest of the step instead of stalling
    // the next step's differences (the mass, used by the force factors below,
    // rotates at the end with the reaction sums).
    xj = ror1(xj, src_addr);
    yj = ror1(yj, src_addr);
    zj = ror1(zj, src_addr);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
      float r2 = __builtin_fmaf(dx[q], dx[q], eps2);
      r2 = __builtin_fmaf(dy[q], dy[q], r2);
      w[q] = __builtin_fmaf(dz[q], dz[q], r2);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < IPL; ++q) w[q] = __builtin_amdgcn_rsqf(w[q]);
    __builtin_amdgcn_sched_barrier(0);
    float wj[IPL];
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
      const float inv = w[q];
      float c = (inv * inv) * inv;
      if (CHECKED) c = (s == 0 && q == self_q) ? 0.0f : c;
      if (UNI) {
        w[q] = c;
        wj[q] = c;
      } else {
        w[q] = mj * c;
        wj[q] = mi[q] * c;
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
      fx[q] = __builtin_fmaf(w[q], dx[q], fx[q]);
      fy[q] = __builtin_fmaf(w[q], dy[q], fy[q]);
      fz[q] = __builtin_fmaf(w[q], dz[q], fz[q]);
      if (JSIDE) {
        gx = __builtin_fmaf(-wj[q], dx[q], gx);
        gy = __builtin_fmaf(-wj[q], dy[q], gy);
        gz = __builtin_fmaf(-wj[q], dz[q], gz);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    if (!UNI) mj = ror1(mj, src_addr);
    if (JSIDE) {
      gx = ror1(gx, src_addr);
      gy = ror1(gy, src_addr);
      gz = ror1(gz, src_addr);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

__device__ __forceinline__ void load_body(const float* __restrict__ x,
                                          const float* __restrict__ y,
                                          const float* __restrict__ z,
                                          const float* __restrict__ m, int j, int n, float& bx,
                                          float& by, float& bz, float& bm) {
  const bool ok = j < n;
  bx = ok ? x[j] : kPadPos;
  by = ok ? y[j] : kPadPos;
  bz = ok ? z[j] : kPadPos;
  bm = ok ? m[j] : 0.0f;
}

// One tile: targets of this wave against every chunk of super-block C.
template <int IPL, int W, bool UNI, bool JSIDE, bool CHECKED>
__device__ __forceinline__ void sym_tile(const float* __restrict__ x,
                                         const float* __restrict__ y,
                                         const float* __restrict__ z,
                                         const float* __restrict__ m, const SymGeom& g, int A,
                                         int C, int d, float mref, float eps2,
                                         const float (&xi)[IPL], const float (&yi)[IPL],
                                         const float (&zi)[IPL], const float (&mi)[IPL],
                                         float (*red)[W][3][64], int tid) {
  constexpr int B = 64 * IPL;
  constexpr int S = W * B;
  constexpr int kChunks = S / 64;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int jbase = C * S;
  const int src_addr = ((lane + 63) & 63) * 4;
  double sx[IPL], sy[IPL], sz[IPL];
#pragma unroll
  for (int q = 0; q < IPL; ++q) sx[q] = sy[q] = sz[q] = 0.0;
  float nx, ny, nz, nm;
  load_body(x, y, z, m, jbase + lane, g.n, nx, ny, nz, nm);
  // eps2 as a VGPR operand: on gfx950 an FP32 VALU instruction that reads an
  // SGPR issues at half the rate of one whose operands are all VGPRs
  // (measured: v_fmac/v_sub 3.8e13 against 5.3-7.2e13 lane-ops/s).
  float eps2v = eps2;
  asm volatile("" : "+v"(eps2v));
  for (int c = 0; c < kChunks; ++c) {
    float xj = nx, yj = ny, zj = nz, mj = nm;
    if (c + 1 < kChunks) load_body(x, y, z, m, jbase + (c + 1) * 64 + lane, g.n, nx, ny, nz, nm);
    float fx[IPL], fy[IPL], fz[IPL];
#pragma unroll
    for (int q = 0; q < IPL; ++q) fx[q] = fy[q] = fz[q] = 0.0f;
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    int self_q = -1;
    if (CHECKED) {
      const int q = c - wave * IPL;
      self_q = (q >= 0 && q < IPL) ? q : -1;
    }
    sym_chunk<IPL, UNI, JSIDE, CHECKED>(src_addr, xj, yj, zj, mj, xi, yi, zi, mi, eps2v, self_q,
                                        fx, fy, fz, gx, gy, gz);
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
      if (UNI) {
        sx[q] += (double)fx[q] * (double)mref;
        sy[q] += (double)fy[q] * (double)mref;
        sz[q] += (double)fz[q] * (double)mref;
      } else {
        sx[q] += (double)fx[q];
        sy[q] += (double)fy[q];
        sz[q] += (double)fz[q];
      }
    }
    if (JSIDE) {
      float(*buf)[3][64] = red[c & 1];
      buf[wave][0][lane] = gx;
      buf[wave][1][lane] = gy;
      buf[wave][2][lane] = gz;
      __syncthreads();
      // 3 components x 64 bodies, whatever the workgroup size
      for (int r = tid; r < 192; r += 64 * W) {
        const int k = r >> 6, l = r & 63;
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < W; ++w) t += (double)buf[w][k][l];
        if (UNI) t *= (double)mref;
        g.pj[((long)(d - g.pj_lo()) * 3 + k) * g.npad + jbase + c * 64 + l] = (float)t;
      }
      // red[c & 1] is rewritten at chunk c + 2, after every thread has
      // passed the barrier of chunk c + 1, so one barrier per chunk suffices.
    }
  }
  const int ibase = A * S + wave * B;
#pragma unroll
  for (int q = 0; q < IPL; ++q) {
    const long i = ibase + q * 64 + lane;
    g.pi[((long)(d - g.d_lo) * 3 + 0) * g.npad + i] = (float)sx[q];
    g.pi[((long)(d - g.d_lo) * 3 + 1) * g.npad + i] = (float)sy[q];
    g.pi[((long)(d - g.d_lo) * 3 + 2) * g.npad + i] = (float)sz[q];
  }
}

template <int IPL, int W, bool UNI, bool CHECKED>
__device__ __forceinline__ void sym_dispatch(const float* __restrict__ x,
                                             const float* __restrict__ y,
                                             const float* __restrict__ z,
                                             const float* __restrict__ m, const SymGeom& g, int A,
                                             int C, int d, float mref, float eps2,
                                             const float (&xi)[IPL], const float (&yi)[IPL],
                                             const float (&zi)[IPL], const float (&mi)[IPL],
                                             float (*red)[W][3][64], int tid) {
  if (d == 0)
    sym_tile<IPL, W, UNI, false, CHECKED>(x, y, z, m, g, A, C, d, mref, eps2, xi, yi, zi, mi, red,
                                          tid);
  else
    sym_tile<IPL, W, UNI, true, false>(x, y, z, m, g, A, C, d, mref, eps2, xi, yi, zi, mi, red,
                                       tid);
}

// Tile t of a round: its diagonal tiles first (t < K, round starting at d = 0),
// then row-major over (A, d), d = pj_lo() .. d_hi. Returns false for a tile
// owned by the other row (even K, d = K/2).
__device__ __forceinline__ bool sym_tile_of(const SymGeom& g, int t, int& A, int& d) {
  if (g.d_lo == 0) {
    if (t < g.K) {
      A = t;
      d = 0;
      return true;
    }
    t -= g.K;
  }
  const int nd = g.d_hi - g.pj_lo() + 1;
  A = t / nd;
  d = t % nd + g.pj_lo();
  return d <= sym_rowlen(A, g.K);
}

// One tile (A, d) of the system at x..m, by one workgroup of W waves.
template <int IPL, int W, bool CHECKED>
__device__ __forceinline__ void sym_run_tile(const float* __restrict__ x,
                                             const float* __restrict__ y,
                                             const float* __restrict__ z,
                                             const float* __restrict__ m, const SymGeom& g, int A,
                                             int d, float eps2, float (*red)[W][3][64]) {
  constexpr int B = 64 * IPL;
  constexpr int S = W * B;
  // The thread index through an empty asm: a kernel that runs many tiles
  // (the batched ensemble) then derives the lane values per tile instead of
  // holding them in VGPRs across its whole item loop.
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int C = (A + d) % g.K;
  float xi[IPL], yi[IPL], zi[IPL], mi[IPL];
#pragma unroll
  for (int q = 0; q < IPL; ++q)
    load_body(x, y, z, m, A * S + wave * B + q * 64 + lane, g.n, xi[q], yi[q], zi[q], mi[q]);
  // Do all real bodies of A and C share one mass?
  const float mref = m[A * S];
  bool same = true;
  for (int k = tid; k < S; k += 64 * W) {
    const int ia = A * S + k, jc = C * S + k;
    if (ia < g.n) same = same && (m[ia] == mref);
    if (jc < g.n) same = same && (m[jc] == mref);
  }
  if (block_all<W>(same))
    sym_dispatch<IPL, W, true, CHECKED>(x, y, z, m, g, A, C, d, mref, eps2, xi, yi, zi, mi, red,
                                        tid);
  else
    sym_dispatch<IPL, W, false, CHECKED>(x, y, z, m, g, A, C, d, mref, eps2, xi, yi, zi, mi, red,
                                         tid);
}

// Body b's partials of one round added to t in the canonical order: for d
// ascending, the i-side partial of distance d, then the j-side one. The order
// does not depend on how the distances are split into rounds, and t is carried
// between rounds in FP64 (exact), so every round size gives the same bits.
// b is system-local.
__device__ __forceinline__ void sym_round_add(const SymGeom& g, int b, double (&t)[3]) {
  const int C = b / g.S;
  int d = g.d_lo;
  if (d == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] += (double)g.pi[((long)k) * g.npad + b];
    d = 1;
  }
  // Both sides exist for every row at d <= Dmax - 1 (sym_rowlen >= floor(K/2) - 1).
  // Four distances' loads (24 per lane) are issued before their adds: the pass
  // is bound by HBM latency otherwise.
  const int dsafe = min(g.d_hi, g.Dmax - 1);
  for (; d + 3 <= dsafe; d += 4) {
    float vi[4][3], vj[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        vi[u][k] = g.pi[((long)(d + u - g.d_lo) * 3 + k) * g.npad + b];
        vj[u][k] = g.pj[((long)(d + u - g.pj_lo()) * 3 + k) * g.npad + b];
      }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        t[k] += (double)vi[u][k];
        t[k] += (double)vj[u][k];
      }
  }
  for (; d <= g.d_hi; ++d) {
    const bool own = d <= sym_rowlen(C, g.K);
    const bool other = d <= sym_rowlen((C - d + g.K) % g.K, g.K);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (own) t[k] += (double)g.pi[((long)(d - g.d_lo) * 3 + k) * g.npad + b];
      if (other) t[k] += (double)g.pj[((long)(d - g.pj_lo()) * 3 + k) * g.npad + b];
    }
  }
}

// One round's reduction for body b: t starts at zero (first round) or from the
// FP64 carry acc64[3][npad]; after the last round a_b = G * t is written with
// the optional kick (as in direct_tile()), otherwise t goes back to the carry.
// b is system-local; a*, kick and acc64 point at the system's first body.
template <bool KICK>
__device__ __forceinline__ void sym_reduce_body(const SymGeom& g, int b, double G, float* ax,
                                                float* ay, float* az, const KickArgs& kick,
                                                double* acc64, bool last) {
  double t[3] = {0.0, 0.0, 0.0};
  if (g.d_lo > 0)
#pragma unroll
    for (int k = 0; k < 3; ++k) t[k] = acc64[(long)k * g.npad + b];
  sym_round_add(g, b, t);
  if (!last) {
#pragma unroll
    for (int k = 0; k < 3; ++k) acc64[(long)k * g.npad + b] = t[k];
    return;
  }
  const float a0 = (float)(G * t[0]), a1 = (float)(G * t[1]), a2 = (float)(G * t[2]);
  ax[b] = a0;
  ay[b] = a1;
  az[b] = a2;
  if (KICK) {
    kick.vx[b] = add_rn(kick.vx[b], mul_rn(add_rn(kick.aox[b], a0), kick.dt_half));
    kick.vy[b] = add_rn(kick.vy[b], mul_rn(add_rn(kick.aoy[b], a1), kick.dt_half));
    kick.vz[b] = add_rn(kick.vz[b], mul_rn(add_rn(kick.aoz[b], a2), kick.dt_half));
  }
}

// Geometry of a whole system of n bodies in super-blocks of S as one round;
// partial sums at pi.
__host__ __device__ inline SymGeom sym_geom(int n, int S, float* pi) {
  SymGeom g{};
  g.n = n;
  g.S = S;
  g.K = (n + S - 1) / S;
  g.Dmax = g.K / 2;
  g.d_lo = 0;
  g.d_hi = g.Dmax;
  g.npad = (long)g.K * S;
  g.pi = pi;
  g.pj = pi + (size_t)(g.Dmax + 1) * 3 * g.npad;
  return g;
}

// The round of distances d_lo .. d_hi of g's system, partial sums at pi.
__host__ __device__ inline SymGeom sym_round(SymGeom g, int d_lo, int d_hi, float* pi) {
  g.d_lo = d_lo;
  g.d_hi = d_hi;
  g.pi = pi;
  g.pj = pi + (size_t)(d_hi - d_lo + 1) * 3 * g.npad;
  return g;
}

// Floats of partial sums for n bodies in super-blocks of S, as one round.
__host__ __device__ inline size_t sym_partial_floats(size_t n, int S) {
  const size_t K = (n + S - 1) / S;
  return (2 * (K / 2) + 1) * 3 * K * (size_t)S;
}

// Tiles of a round (including the skipped ones of even K).
__host__ __device__ inline int sym_tiles(const SymGeom& g) {
  const int off = g.d_hi >= g.pj_lo() ? g.K * (g.d_hi - g.pj_lo() + 1) : 0;
  return (g.d_lo == 0 ? g.K : 0) + off;
}

}  // namespace nbx
