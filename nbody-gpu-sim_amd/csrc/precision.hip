// Precision form of the direct N^2 path (an extension; the reference has no
// counterpart — its direct kernel is force_direct.cu:10-85 in plain FP32).
//
//  * positions are double-single: x = hi + lo with two FP32 arrays, so the
//    drift x += v dt + a dt^2/2 does not lose the low bits of small
//    increments over long runs;
//  * pair forces stay FP32 (v_rsq_f32); r_ij = (x_j.hi - x_i.hi) + (x_j.lo - x_i.lo);
//  * each target's sum is Kahan-compensated in FP32 over its whole j chunk, and
//    the compensated sums are combined (across waves, chunks, tiles and rounds)
//    as double-single FP32 pairs (ds_add: TwoSum keeps every addition's
//    rounding error);
//  * FP64 appears only in the hi/lo renormalisation of the drift and the
//    energy/momentum reductions.
#include <algorithm>

#include "common.hpp"
#include "direct_core.hpp"
#include "direct_sym.hpp"
#include "direct_sym_core.hpp"
#include "verlet.hpp"

namespace nbx {

namespace {

constexpr int kDsBatch = 8;  // 7 scalar loads per body: keep a batch within the SGPR budget

// hi + lo += x in FP32, the addition's rounding error kept in lo (Knuth's TwoSum).
__device__ __forceinline__ void ds_add(float& hi, float& lo, float x) {
  const float s = add_rn(hi, x);
  const float bp = sub_rn(s, hi);
  lo = add_rn(lo, add_rn(sub_rn(hi, sub_rn(s, bp)), sub_rn(x, bp)));
  hi = s;
}

// (hi + lo) * k as one FP32 value: the product's rounding error of hi * k is
// recovered with an FMA and added back with lo * k.
__device__ __forceinline__ float ds_scale(float hi, float lo, float k) {
  const float p = mul_rn(hi, k);
  return add_rn(p, add_rn(__builtin_fmaf(hi, k, -p), mul_rn(lo, k)));
}

__device__ __forceinline__ double dadd_rn(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ double dmul_rn(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}

template <bool CHECKED>
__device__ __forceinline__ void ds_pair(float bxh, float byh, float bzh, float bxl, float byl,
                                        float bzl, float bm, int j, float xh, float yh, float zh,
                                        float xl, float yl, float zl, int i, float eps2,
                                        float (&s)[3], float (&c)[3]) {
  const float dx = (bxh - xh) + (bxl - xl);
  const float dy = (byh - yh) + (byl - yl);
  const float dz = (bzh - zh) + (bzl - zl);
  float r2 = __builtin_fmaf(dx, dx, eps2);
  r2 = __builtin_fmaf(dy, dy, r2);
  r2 = __builtin_fmaf(dz, dz, r2);
  const float inv = __builtin_amdgcn_rsqf(r2);
  float f = (bm * inv) * (inv * inv);
  if (CHECKED) f = (j == i) ? 0.0f : f;
  const float d[3] = {dx, dy, dz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {  // Kahan: s + (f d) with the lost low part kept in c
    const float y = __builtin_fmaf(f, d[a], -c[a]);
    const float t = s[a] + y;
    c[a] = (t - s[a]) - y;
    s[a] = t;
  }
}

template <int IPL, int WS, bool CHECKED, bool KICK>
__global__ __launch_bounds__(64 * WS) void ds_forces_kernel(
    const float* __restrict__ xh, const float* __restrict__ yh, const float* __restrict__ zh,
    const float* __restrict__ xl, const float* __restrict__ yl, const float* __restrict__ zl,
    const float* __restrict__ m, int n, float G, float eps2, float* ax, float* ay, float* az,
    KickArgs kick) {
  __shared__ float2 red[WS * 3 * 64 * IPL];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ibase = (int)blockIdx.x * 64 * IPL;
  float pxh[IPL], pyh[IPL], pzh[IPL], pxl[IPL], pyl[IPL], pzl[IPL];
  int ii[IPL];
#pragma unroll
  for (int q = 0; q < IPL; ++q) {
    const int i = ibase + q * 64 + lane;
    ii[q] = i;
    const int ic = i < n ? i : 0;
    pxh[q] = xh[ic];
    pyh[q] = yh[ic];
    pzh[q] = zh[ic];
    pxl[q] = xl[ic];
    pyl[q] = yl[ic];
    pzl[q] = zl[ic];
  }
  float s[IPL][3], c[IPL][3];
#pragma unroll
  for (int q = 0; q < IPL; ++q)
    for (int a = 0; a < 3; ++a) s[q][a] = c[q][a] = 0.0f;
  const int chunk = direct_chunk(n, WS);
  const int j0 = min(n, wave * chunk);
  const int j1 = min(n, j0 + chunk);
  int j = j0;
  for (; j + kDsBatch <= j1; j += kDsBatch) {
#pragma unroll
    for (int k = 0; k < kDsBatch; ++k) {
      const int jj = j + k;
      const float bxh = xh[jj], byh = yh[jj], bzh = zh[jj], bxl = xl[jj], byl = yl[jj],
                  bzl = zl[jj], bm = m[jj];
#pragma unroll
      for (int q = 0; q < IPL; ++q)
        ds_pair<CHECKED>(bxh, byh, bzh, bxl, byl, bzl, bm, jj, pxh[q], pyh[q], pzh[q], pxl[q],
                         pyl[q], pzl[q], ii[q], eps2, s[q], c[q]);
    }
  }
  for (; j < j1; ++j) {
    const float bxh = xh[j], byh = yh[j], bzh = zh[j], bxl = xl[j], byl = yl[j], bzl = zl[j],
                bm = m[j];
#pragma unroll
    for (int q = 0; q < IPL; ++q)
      ds_pair<CHECKED>(bxh, byh, bzh, bxl, byl, bzl, bm, j, pxh[q], pyh[q], pzh[q], pxl[q],
                       pyl[q], pzl[q], ii[q], eps2, s[q], c[q]);
  }
  constexpr int T = 64 * IPL;
#pragma unroll
  for (int q = 0; q < IPL; ++q) {
    const int sl = q * 64 + lane;
    for (int a = 0; a < 3; ++a) red[(wave * 3 + a) * T + sl] = make_float2(s[q][a], -c[q][a]);
  }
  __syncthreads();
  for (int sl = threadIdx.x; sl < T; sl += 64 * WS) {
    const int i = ibase + sl;
    if (i >= n) continue;
    float hi[3] = {0.f, 0.f, 0.f}, lo[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int w = 0; w < WS; ++w)
      for (int a = 0; a < 3; ++a) {
        const float2 r = red[(w * 3 + a) * T + sl];
        ds_add(hi[a], lo[a], r.x);
        lo[a] = add_rn(lo[a], r.y);
      }
    const float a0 = ds_scale(hi[0], lo[0], G), a1 = ds_scale(hi[1], lo[1], G),
                a2 = ds_scale(hi[2], lo[2], G);
    ax[i] = a0;
    ay[i] = a1;
    az[i] = a2;
    if (KICK) {
      kick.vx[i] = add_rn(kick.vx[i], mul_rn(add_rn(kick.aox[i], a0), kick.dt_half));
      kick.vy[i] = add_rn(kick.vy[i], mul_rn(add_rn(kick.aoy[i], a1), kick.dt_half));
      kick.vz[i] = add_rn(kick.vz[i], mul_rn(add_rn(kick.aoz[i], a2), kick.dt_half));
    }
  }
}

// ---- pair-symmetric form (the default above 2^13 bodies) -------------------
// The tiles, partial sums and reduction of the single-system symmetric kernel
// (direct_sym_core.hpp) with double-single positions and Kahan-compensated
// sums on both sides of every pair: the targets' sums per chunk, and the
// reactions, whose sums and compensations rotate with their bodies.

__device__ __forceinline__ void load_ds_body(const float* __restrict__ xh,
                                             const float* __restrict__ yh,
                                             const float* __restrict__ zh,
                                             const float* __restrict__ xl,
                                             const float* __restrict__ yl,
                                             const float* __restrict__ zl,
                                             const float* __restrict__ m, int j, int n,
                                             float (&b)[7]) {
  const bool ok = j < n;
  b[0] = ok ? xh[j] : kPadPos;
  b[1] = ok ? yh[j] : kPadPos;
  b[2] = ok ? zh[j] : kPadPos;
  b[3] = ok ? xl[j] : 0.0f;
  b[4] = ok ? yl[j] : 0.0f;
  b[5] = ok ? zl[j] : 0.0f;
  b[6] = ok ? m[j] : 0.0f;
}

// s + v with the lost low part kept in c (Kahan).
__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
  const float y = v - c;
  const float t = s + y;
  c = (t - s) - y;
  s = t;
}

template <int IPL, bool UNI, bool JSIDE, bool CHECKED>
__device__ __forceinline__ void ds_sym_chunk(int src_addr, float (&b)[7],
                                             const float (&ih)[IPL][3], const float (&il)[IPL][3],
                                             const float (&mi)[IPL], float eps2, int self_q,
                                             float (&s)[IPL][3], float (&c)[IPL][3],
                                             float (&g)[3], float (&cg)[3]) {
#pragma unroll 2
  for (int st = 0; st < 64; ++st) {
    float d[IPL][3], w[IPL], wj[IPL];
#pragma unroll
    for (int q = 0; q < IPL; ++q)
#pragma unroll
      for (int a = 0; a < 3; ++a) d[q][a] = (b[a] - ih[q][a]) + (b[3 + a] - il[q][a]);
    __builtin_amdgcn_sched_barrier(0);
    // the chunk's positions are dead for the rest of the step: rotate them now,
    // so the step's arithmetic hides the LDS round trip
#pragma unroll
    for (int k = 0; k < 6; ++k) b[k] = ror1(b[k], src_addr);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
      float r2 = __builtin_fmaf(d[q][0], d[q][0], eps2);
      r2 = __builtin_fmaf(d[q][1], d[q][1], r2);
      w[q] = __builtin_fmaf(d[q][2], d[q][2], r2);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < IPL; ++q) w[q] = __builtin_amdgcn_rsqf(w[q]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < IPL; ++q) {
      const float inv = w[q];
      float f = (inv * inv) * inv;
      if (CHECKED) f = (st == 0 && q == self_q) ? 0.0f : f;
      w[q] = UNI ? f : b[6] * f;
      wj[q] = UNI ? f : mi[q] * f;
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < IPL; ++q)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        kahan_add(s[q][a], c[q][a], w[q] * d[q][a]);
        if (JSIDE) kahan_add(g[a], cg[a], -(wj[q] * d[q][a]));
      }
    __builtin_amdgcn_sched_barrier(0);
    if (!UNI) b[6] = ror1(b[6], src_addr);
    if (JSIDE)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        g[a] = ror1(g[a], src_addr);
        cg[a] = ror1(cg[a], src_addr);
      }
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int IPL, int W, bool UNI, bool JSIDE, bool CHECKED>
__device__ __forceinline__ void ds_sym_tile(const float* __restrict__ xh,
                                            const float* __restrict__ yh,
                                            const float* __restrict__ zh,
                                            const float* __restrict__ xl,
                                            const float* __restrict__ yl,
                                            const float* __restrict__ zl,
                                            const float* __restrict__ m, const SymGeom& g, int A,
                                            int C, int dd, float mref, float eps2,
                                            const float (&ih)[IPL][3], const float (&il)[IPL][3],
                                            const float (&mi)[IPL], float2 (*red)[W][3][64]) {
  constexpr int B = 64 * IPL;
  constexpr int S = W * B;
  constexpr int kChunks = S / 64;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int jbase = C * S;
  const int src_addr = ((lane + 63) & 63) * 4;
  float sh[IPL][3], sl[IPL][3];  // the targets' sums over chunks, double-single
#pragma unroll
  for (int q = 0; q < IPL; ++q)
#pragma unroll
    for (int a = 0; a < 3; ++a) sh[q][a] = sl[q][a] = 0.0f;
  float nb[7];
  load_ds_body(xh, yh, zh, xl, yl, zl, m, jbase + lane, g.n, nb);
  float eps2v = eps2;  // a VGPR operand (an SGPR operand halves the FP32 issue rate)
  asm volatile("" : "+v"(eps2v));
  for (int ch = 0; ch < kChunks; ++ch) {
    float b[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) b[k] = nb[k];
    if (ch + 1 < kChunks) load_ds_body(xh, yh, zh, xl, yl, zl, m, jbase + (ch + 1) * 64 + lane, g.n, nb);
    float s[IPL][3], c[IPL][3], gs[3] = {0.f, 0.f, 0.f}, gc[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < IPL; ++q)
#pragma unroll
      for (int a = 0; a < 3; ++a) s[q][a] = c[q][a] = 0.0f;
    int self_q = -1;
    if (CHECKED) {
      const int q = ch - wave * IPL;
      self_q = (q >= 0 && q < IPL) ? q : -1;
    }
    ds_sym_chunk<IPL, UNI, JSIDE, CHECKED>(src_addr, b, ih, il, mi, eps2v, self_q, s, c, gs, gc);
#pragma unroll
    for (int q = 0; q < IPL; ++q)
#pragma unroll
      for (int a = 0; a < 3; ++a) {  // the chunk's Kahan sum is s - c
        ds_add(sh[q][a], sl[q][a], s[q][a]);
        sl[q][a] = sub_rn(sl[q][a], c[q][a]);
      }
    if (JSIDE) {
      float2(*buf)[3][64] = red[ch & 1];
#pragma unroll
      for (int a = 0; a < 3; ++a) buf[wave][a][lane] = make_float2(gs[a], -gc[a]);
      __syncthreads();
      for (int r = threadIdx.x; r < 192; r += 64 * W) {
        const int k = r >> 6, l = r & 63;
        float th = 0.0f, tl = 0.0f;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          ds_add(th, tl, buf[w][k][l].x);
          tl = add_rn(tl, buf[w][k][l].y);
        }
        g.pj[((long)(dd - g.pj_lo()) * 3 + k) * g.npad + jbase + ch * 64 + l] =
            UNI ? ds_scale(th, tl, mref) : add_rn(th, tl);
      }
    }
  }
  const int ibase = A * S + wave * B;
#pragma unroll
  for (int q = 0; q < IPL; ++q)
#pragma unroll
    for (int a = 0; a < 3; ++a)
      g.pi[((long)(dd - g.d_lo) * 3 + a) * g.npad + ibase + q * 64 + lane] =
          UNI ? ds_scale(sh[q][a], sl[q][a], mref) : add_rn(sh[q][a], sl[q][a]);
}

// At least 4 waves per SIMD: unbounded, the unchecked form takes 129 VGPRs (3
// waves); held to 128 it spills 2 outside its inner loop and runs 2.7 % faster.
template <int IPL, int W, bool CHECKED>
__global__ __launch_bounds__(64 * W, 4) void ds_sym_kernel(
    const float* __restrict__ xh, const float* __restrict__ yh, const float* __restrict__ zh,
    const float* __restrict__ xl, const float* __restrict__ yl, const float* __restrict__ zl,
    const float* __restrict__ m, SymGeom g, float eps2) {
  constexpr int B = 64 * IPL;
  constexpr int S = W * B;
  __shared__ float2 red[2][W][3][64];
  int A, dd;
  if (!sym_tile_of(g, (int)blockIdx.x, A, dd)) return;  // whole workgroup: the other row's tile
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int C = (A + dd) % g.K;
  float ih[IPL][3], il[IPL][3], mi[IPL];
#pragma unroll
  for (int q = 0; q < IPL; ++q) {
    float b[7];
    load_ds_body(xh, yh, zh, xl, yl, zl, m, A * S + wave * B + q * 64 + lane, g.n, b);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      ih[q][a] = b[a];
      il[q][a] = b[3 + a];
    }
    mi[q] = b[6];
  }
  const float mref = m[A * S];
  bool same = true;
  for (int k = threadIdx.x; k < S; k += 64 * W) {
    const int ia = A * S + k, jc = C * S + k;
    if (ia < g.n) same = same && (m[ia] == mref);
    if (jc < g.n) same = same && (m[jc] == mref);
  }
  const bool uni = block_all<W>(same);
#define NBX_DS_TILE(U, J, CK) \
  ds_sym_tile<IPL, W, U, J, CK>(xh, yh, zh, xl, yl, zl, m, g, A, C, dd, mref, eps2, ih, il, mi, red)
  if (uni) {
    if (dd == 0) NBX_DS_TILE(true, false, CHECKED);
    else NBX_DS_TILE(true, true, false);
  } else {
    if (dd == 0) NBX_DS_TILE(false, false, CHECKED);
    else NBX_DS_TILE(false, true, false);
  }
#undef NBX_DS_TILE
}

// Body b's partials of one round added to (hi, lo) in the canonical order of
// sym_round_add (for d ascending, the i-side partial of distance d, then the
// j-side one), as double-single FP32 sums; the carry between rounds holds the
// pair exactly, so every round size gives the same bits.
__device__ __forceinline__ void ds_round_add(const SymGeom& g, int b, float (&hi)[3],
                                             float (&lo)[3]) {
  const int C = b / g.S;
  int d = g.d_lo;
  if (d == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) ds_add(hi[k], lo[k], g.pi[(long)k * g.npad + b]);
    d = 1;
  }
  // both sides exist for every row at d <= Dmax - 1; four distances' loads are
  // issued before their adds, which keep the canonical order
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
        ds_add(hi[k], lo[k], vi[u][k]);
        ds_add(hi[k], lo[k], vj[u][k]);
      }
  }
  for (; d <= g.d_hi; ++d) {
    const bool own = d <= sym_rowlen(C, g.K);
    const bool other = d <= sym_rowlen((C - d + g.K) % g.K, g.K);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (own) ds_add(hi[k], lo[k], g.pi[((long)(d - g.d_lo) * 3 + k) * g.npad + b]);
      if (other) ds_add(hi[k], lo[k], g.pj[((long)(d - g.pj_lo()) * 3 + k) * g.npad + b]);
    }
  }
}

// One round's reduction of the double-single partial sums: (hi, lo) starts at
// zero (first round) or from the carry (3 float pairs per padded body); after
// the last round a = G (hi + lo) and the optional kick, as sym_reduce_body.
template <bool KICK>
__global__ __launch_bounds__(256) void ds_sym_reduce(SymGeom g, float G, float* ax, float* ay,
                                                     float* az, KickArgs kick, float2* carry,
                                                     int last) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.n) return;
  float hi[3] = {0.f, 0.f, 0.f}, lo[3] = {0.f, 0.f, 0.f};
  if (g.d_lo > 0)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const float2 c = carry[(long)k * g.npad + b];
      hi[k] = c.x;
      lo[k] = c.y;
    }
  ds_round_add(g, b, hi, lo);
  if (!last) {
#pragma unroll
    for (int k = 0; k < 3; ++k) carry[(long)k * g.npad + b] = make_float2(hi[k], lo[k]);
    return;
  }
  const float a0 = ds_scale(hi[0], lo[0], G), a1 = ds_scale(hi[1], lo[1], G),
              a2 = ds_scale(hi[2], lo[2], G);
  ax[b] = a0;
  ay[b] = a1;
  az[b] = a2;
  if (KICK) {
    kick.vx[b] = add_rn(kick.vx[b], mul_rn(add_rn(kick.aox[b], a0), kick.dt_half));
    kick.vy[b] = add_rn(kick.vy[b], mul_rn(add_rn(kick.aoy[b], a1), kick.dt_half));
    kick.vz[b] = add_rn(kick.vz[b], mul_rn(add_rn(kick.aoz[b], a2), kick.dt_half));
  }
}

// acc_old <- acc; (hi, lo) += v dt + a dt^2/2 evaluated and renormalised in FP64.
__global__ void ds_drift_kernel(float* __restrict__ xh, float* __restrict__ yh,
                                float* __restrict__ zh, float* __restrict__ xl,
                                float* __restrict__ yl, float* __restrict__ zl,
                                const float* __restrict__ vx, const float* __restrict__ vy,
                                const float* __restrict__ vz, const float* __restrict__ ax,
                                const float* __restrict__ ay, const float* __restrict__ az,
                                float* __restrict__ aox, float* __restrict__ aoy,
                                float* __restrict__ aoz, int n, double dt, double dt2_half) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float* hi[3] = {xh, yh, zh};
  float* lo[3] = {xl, yl, zl};
  const float* v[3] = {vx, vy, vz};
  const float* a[3] = {ax, ay, az};
  float* ao[3] = {aox, aoy, aoz};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float ak = a[k][i];
    ao[k][i] = ak;
    const double inc = dadd_rn(dmul_rn((double)v[k][i], dt), dmul_rn((double)ak, dt2_half));
    const double x = dadd_rn(dadd_rn((double)hi[k][i], (double)lo[k][i]), inc);
    const float h = (float)x;
    hi[k][i] = h;
    lo[k][i] = (float)dadd_rn(x, -(double)h);
  }
}

constexpr int kIpl = 2, kWs = 8;           // one-sided shape
constexpr int kSymIpl = 4, kSymW = 2;      // pair-symmetric shape, super-blocks of 512
constexpr int kSymS = 64 * kSymIpl * kSymW;

// -1: symmetric above 2^13 bodies when the call brings a workspace (0.237 against
// 0.293 ms at 12,288 bodies, 0.240 against 0.389 at 16,384; even at 8,192), 0: one-sided,
// 1: symmetric (needs a workspace).
int g_ds_kernel = -1;

// The symmetric tiles in rounds of distances, as direct_sym_launch.
int ds_sym_launch(nbx_ds_particles* d, float G, float eps2, const KickArgs* kick,
                  const SymWork& w, hipStream_t s) {
  nbx_particles& p = d->p;
  const int n = (int)p.count;
  const int R = direct_sym_round_len((size_t)n, kSymS, w.part_bytes);
  const SymGeom g = sym_geom(n, kSymS, w.part);
  if (R < 0 || !w.part)
    return fail(NBX_ERR_CAPACITY, "pair-symmetric scratch of %zu bytes too small for %d bodies",
                w.part_bytes, n);
  if (R < g.Dmax && (!w.carry || w.carry_bytes < direct_sym_carry_bytes((size_t)n, kSymS)))
    return fail(NBX_ERR_CAPACITY, "pair-symmetric evaluation of %d bodies in rounds needs a "
                "%zu-byte carry", n, direct_sym_carry_bytes((size_t)n, kSymS));
  KickArgs k{};
  if (kick) k = *kick;
  const unsigned rb = div_up((size_t)n, 256);
  for (int lo = 0;;) {
    const int hi = std::min(g.Dmax, lo == 0 ? R : lo + R - 1);
    const int last = hi == g.Dmax ? 1 : 0;
    const SymGeom gr = sym_round(g, lo, hi, w.part);
    const unsigned tiles = (unsigned)sym_tiles(gr);
    if (eps2 < kUncheckedMinEps2)
      hipLaunchKernelGGL((ds_sym_kernel<kSymIpl, kSymW, true>), dim3(tiles), dim3(64 * kSymW), 0,
                         s, p.pos_x, p.pos_y, p.pos_z, d->pos_lo_x, d->pos_lo_y, d->pos_lo_z,
                         p.mass, gr, eps2);
    else
      hipLaunchKernelGGL((ds_sym_kernel<kSymIpl, kSymW, false>), dim3(tiles), dim3(64 * kSymW),
                         0, s, p.pos_x, p.pos_y, p.pos_z, d->pos_lo_x, d->pos_lo_y, d->pos_lo_z,
                         p.mass, gr, eps2);
    NBX_LAUNCHED();
    float2* carry = reinterpret_cast<float2*>(w.carry);  // 8 bytes per component either way
    if (kick)
      hipLaunchKernelGGL(ds_sym_reduce<true>, dim3(rb), dim3(256), 0, s, gr, G, p.acc_x, p.acc_y,
                         p.acc_z, k, carry, last);
    else
      hipLaunchKernelGGL(ds_sym_reduce<false>, dim3(rb), dim3(256), 0, s, gr, G, p.acc_x, p.acc_y,
                         p.acc_z, k, carry, last);
    NBX_LAUNCHED();
    if (last) break;
    lo = hi + 1;
  }
  return NBX_OK;
}

int ds_forces_launch(nbx_direct* op, nbx_ds_particles* d, float G, float eps2,
                     const KickArgs* kick, hipStream_t s) {
  nbx_particles& p = d->p;
  if (!(eps2 >= 0.0f) || !(eps2 < 3.0e38f))
    return fail(NBX_ERR_INVALID, "softening eps2 must be finite and >= 0");
  const int n = (int)p.count;
  if (g_ds_kernel == 1 && !op)
    return fail(NBX_ERR_INVALID, "the pair-symmetric double-single kernel needs a workspace");
  const bool sym = op && (g_ds_kernel == 1 || (g_ds_kernel == -1 && n > (1 << 13)));
  if (sym) return ds_sym_launch(d, G, eps2, kick, op->work, s);
  dim3 grid(div_up(n, 64 * kIpl)), block(64 * kWs);
  KickArgs k{};
  if (kick) k = *kick;
  const bool checked = eps2 < kUncheckedMinEps2;
#define NBX_DS_LAUNCH(C, K)                                                                    \
  hipLaunchKernelGGL((ds_forces_kernel<kIpl, kWs, C, K>), grid, block, 0, s, p.pos_x, p.pos_y, \
                     p.pos_z, d->pos_lo_x, d->pos_lo_y, d->pos_lo_z, p.mass, n, G,            \
                     eps2, p.acc_x, p.acc_y, p.acc_z, k)
  if (checked) {
    if (kick) NBX_DS_LAUNCH(true, true);
    else NBX_DS_LAUNCH(true, false);
  } else {
    if (kick) NBX_DS_LAUNCH(false, true);
    else NBX_DS_LAUNCH(false, false);
  }
#undef NBX_DS_LAUNCH
  NBX_LAUNCHED();
  return NBX_OK;
}

int ds_drift_launch(nbx_ds_particles* d, float dt, hipStream_t s) {
  nbx_particles& p = d->p;
  const int n = (int)p.count;
  const double ddt = (double)dt;
  hipLaunchKernelGGL(ds_drift_kernel, dim3(div_up(n, 256)), dim3(256), 0, s, p.pos_x, p.pos_y,
                     p.pos_z, d->pos_lo_x, d->pos_lo_y, d->pos_lo_z, p.vel_x, p.vel_y, p.vel_z,
                     p.acc_x, p.acc_y, p.acc_z, p.acc_old_x, p.acc_old_y, p.acc_old_z, n, ddt,
                     0.5 * ddt * ddt);
  NBX_LAUNCHED();
  return NBX_OK;
}

int check_ds(const nbx_ds_particles* d, const char* who) {
  if (!d) return fail(NBX_ERR_INVALID, "%s: null struct", who);
  NBX_TRY(check_particles(&d->p, who));
  if (!d->pos_lo_x || !d->pos_lo_y || !d->pos_lo_z)
    return fail(NBX_ERR_INVALID, "%s: missing low-part position arrays", who);
  return NBX_OK;
}

int check_ds_ws(const nbx_direct* op, const nbx_ds_particles* d, const char* who) {
  if (!op) return NBX_OK;
  if (d->p.count > op->max_particles)
    return fail(NBX_ERR_CAPACITY, "%s: %zu bodies exceed the workspace's %zu", who, d->p.count,
                op->max_particles);
  int dev = 0;
  NBX_HIP(hipGetDevice(&dev));
  if (dev != op->device)
    return fail(NBX_ERR_STATE, "%s: workspace of device %d used on device %d", who, op->device,
                dev);
  return NBX_OK;
}

}  // namespace

}  // namespace nbx

using namespace nbx;

extern "C" {

int nbx_ds_alloc(nbx_ds_particles* d, size_t count) {
  if (!d) return fail(NBX_ERR_INVALID, "nbx_ds_alloc: null struct");
  d->pos_lo_x = d->pos_lo_y = d->pos_lo_z = nullptr;
  NBX_TRY(nbx_particles_alloc(&d->p, count));
  float** lo[3] = {&d->pos_lo_x, &d->pos_lo_y, &d->pos_lo_z};
  for (float** f : lo) {
    if (hipMalloc(reinterpret_cast<void**>(f), std::max<size_t>(count, 1) * sizeof(float)) !=
            hipSuccess ||
        hipMemset(*f, 0, std::max<size_t>(count, 1) * sizeof(float)) != hipSuccess) {
      (void)hipGetLastError();
      nbx_ds_free(d);
      return fail(NBX_ERR_NOMEM, "double-single position arrays for %zu particles", count);
    }
  }
  return NBX_OK;
}

int nbx_ds_free(nbx_ds_particles* d) {
  if (!d) return NBX_OK;
  float** lo[3] = {&d->pos_lo_x, &d->pos_lo_y, &d->pos_lo_z};
  for (float** f : lo) {
    if (*f) (void)hipFree(*f);
    *f = nullptr;
  }
  return nbx_particles_free(&d->p);
}

int nbx_ds_forces(nbx_direct* op, nbx_ds_particles* d, float G, float eps2, nbx_stream s) {
  NBX_TRY(check_ds(d, "nbx_ds_forces"));
  NBX_TRY(check_ds_ws(op, d, "nbx_ds_forces"));
  return ds_forces_launch(op, d, G, eps2, nullptr, as_stream(s));
}

int nbx_ds_set_kernel(int kind) {
  if (kind < -1 || kind > 1) return fail(NBX_ERR_INVALID, "nbx_ds_set_kernel: kind %d", kind);
  g_ds_kernel = kind;
  return NBX_OK;
}

int nbx_ds_drift(nbx_ds_particles* d, float dt, nbx_stream s) {
  NBX_TRY(check_ds(d, "nbx_ds_drift"));
  return ds_drift_launch(d, dt, as_stream(s));
}

int nbx_ds_step(nbx_direct* op, nbx_ds_particles* d, float G, float eps2, float dt,
                nbx_stream s) {
  NBX_TRY(check_ds(d, "nbx_ds_step"));
  NBX_TRY(check_ds_ws(op, d, "nbx_ds_step"));
  hipStream_t st = as_stream(s);
  NBX_TRY(ds_drift_launch(d, dt, st));
  nbx_particles& p = d->p;
  KickArgs k{p.vel_x, p.vel_y, p.vel_z, p.acc_old_x, p.acc_old_y, p.acc_old_z, 0.5f * dt};
  return ds_forces_launch(op, d, G, eps2, &k, st);
}

int nbx_ds_energy(const nbx_ds_particles* d, float G, float eps2, double* ke, double* pe,
                  double momentum[3], nbx_stream s) {
  NBX_TRY(check_ds(d, "nbx_ds_energy"));
  return energy_launch(&d->p, d->pos_lo_x, d->pos_lo_y, d->pos_lo_z, G, eps2, ke, pe, momentum,
                       as_stream(s));
}

}  // extern "C"
