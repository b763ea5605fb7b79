// Spatial-hash (uniform cell grid) short-range forces: the reference's
// SpatialHashGrid and SpatialHashCalculator (src/cuda/force_spatial_hash.cu,
// include/nbody/spatial_hash_grid.hpp), on the GPU.
//
// Build (SpatialHashGrid::build, :241-309):
//   bbox    exact per-axis min/max (the Barnes-Hut reduction, which the
//           reference reuses, :184-188), read back once (24 bytes). The host
//           pads it by 0.001f (:232-238), sets dims = (int)ceilf((max-min)/cell)+1
//           in FP32 (:250-252) and refuses grids above 1e8 cells (:254-260).
//   assign  cell of each particle: (int)floorf((p - min) / cell) per axis,
//           clamped to the grid, linear cx + cy*dx + cz*dx*dy (:36-49).
//   sort    stable radix sort of (key, particle index), key = cell << 12 | the
//           Morton index of the particle's sub-cell in a 16 x 16 x 16 grid of
//           the cell (above 2^20 cells: cell << 6 | a 4 x 4 x 4 one; above 2^26:
//           the cell alone). Within a cell the order is therefore spatially
//           compact runs, then index. The reference's atomic scatter (:72-81)
//           leaves that order to the scheduler; this is one of its outcomes, and
//           it makes every later sum deterministic.
//   ranges  cell_start / cell_end of every cell, the exclusive scan of the
//           per-cell counts (:287-300), from one merge of the sorted keys with
//           the cell numbers: each thread takes a fixed slice of the merged
//           sequence, so a sparse grid (many empty cells) and a dense one cost
//           the same per element, and no count array or memset is needed.
//
// Forces (spatialHashForceKernel, :84-158). For target i in cell c, each j != i
// in the 27 cells around c (clamped at the grid edge) with r^2 < cutoff^2
// adds G m_j r_ij (r^2 + eps^2)^{-3/2}; cells in dz, dy, dx order, bodies in
// sorted order. The three dx cells of one (dy, dz) row are adjacent in the
// linear numbering, so they are one contiguous range of the sorted bodies.
//   * bodies are gathered once per call into sorted order as float4
//     (x, y, z, G*m), G*m being the reference's first product (:144), with
//     the bounding box of every aligned group of 4.
//   * a wave takes 64 consecutive sorted targets. When they all lie in one
//     (cy, cz) row of cells (the common case once cells hold dozens of
//     bodies), the wave walks the union of the lanes' ranges with
//     wave-uniform addresses, skipping the groups whose box is provably
//     beyond the cutoff from the box of its targets (walk_culled): the bodies
//     of two groups at a time arrive through scalar loads and reach every lane
//     as SGPR operands, and a lane drops the bodies outside its own three cells
//     with a mask.
//     Otherwise every lane walks its own ranges with vector loads.
//   * r^2 is fma(rz, rz, fma(rx, rx, ry*ry)) on both sides (here and in the
//     oracle), so the set of pairs kept by the cutoff test is exactly the
//     oracle's; the kept pairs use v_rsq_f32 and FMA accumulation
//     (tolerance-level differences).
//   * both walks add a target's nonzero terms in the same order with the same
//     arithmetic (skipped or masked bodies add exact zeros), so a target's
//     result does not depend on which one ran.
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "bbox.hpp"
#include "common.hpp"
#include "radix.hpp"

namespace nbx {
namespace {

constexpr int kThreads = 256;
constexpr int kMergeSpan = 32;             // merged elements per thread in grid_ranges
constexpr long long kMaxCells = 100000000;  // :257-260
constexpr int kNB = 8;                      // bodies per batch
constexpr int kGroup = 4;                   // bodies per culling box
constexpr int kPer = kNB / kGroup;          // culling groups per batch
// Order within a cell: the Morton index of a 16 x 16 x 16 sub-cell grid (12 key
// bits) up to 2^20 cells, of a 4 x 4 x 4 one (6 bits) up to 2^26, none beyond.
constexpr long long kSub16MaxCells = 1LL << 20;
constexpr long long kSub4MaxCells = 1LL << 26;

struct GridGeom {
  float lo[3];  // padded lower corner
  float cell;
  int dims[3];
  int total;
  int shift;  // sort key = cell << shift | sub-cell (12: 16^3 sub-cells, 6: 4^3, 0: cell only)
};

// :36-49 and :103-108 (the force kernel repeats the arithmetic and the clamp).
__device__ __forceinline__ int grid_axis(float p, float lo, float cell, int dim) {
  const int c = (int)floorf(div_rn(sub_rn(p, lo), cell));
  return max(0, min(c, dim - 1));
}

// Cell and sub-cell coordinate on one axis: u = (p - lo) / cell as above,
// sub = floor(ns u) - ns c clamped to 0..ns-1, ns = 16 or 4 (ns u is exact).
__device__ __forceinline__ int grid_axis_sub(float p, float lo, float cell, int dim, int ns,
                                             int& sub) {
  const float u = div_rn(sub_rn(p, lo), cell);
  const int c = max(0, min((int)floorf(u), dim - 1));
  sub = max(0, min((int)floorf(u * (float)ns) - ns * c, ns - 1));
  return c;
}

__device__ __forceinline__ uint32_t spread4(int v) {  // bits 3..0 -> 9, 6, 3, 0
  return (uint32_t)((v & 1) | ((v & 2) << 2) | ((v & 4) << 4) | ((v & 8) << 6));
}

__global__ __launch_bounds__(kThreads) void grid_bbox_partial(const float* __restrict__ x,
                                                              const float* __restrict__ y,
                                                              const float* __restrict__ z, int n,
                                                              float* partial) {
  bbox_partial_body<kThreads>(x, y, z, n, partial);
}

__global__ __launch_bounds__(kThreads) void grid_bbox_final(const float* partial, int nparts,
                                                            float* box) {
  float lo[3], hi[3];
  if (!bbox_final_body<kThreads>(partial, nparts, lo, hi)) return;
  for (int a = 0; a < 3; ++a) {
    box[a] = lo[a];
    box[3 + a] = hi[a];
  }
}

// Cell of each particle (:36-49) and its sort key: the cell, then (shift 12 or 6)
// the Morton index of its sub-cell (16 or 4 per axis), so that a cell's bodies
// are ordered in compact spatial groups. The reference's scatter leaves the order
// within a cell open (:72-81); this order is one of its outcomes.
__global__ __launch_bounds__(kThreads) void grid_assign(const float* __restrict__ x,
                                                        const float* __restrict__ y,
                                                        const float* __restrict__ z, int n,
                                                        GridGeom g, int32_t* __restrict__ pcell,
                                                        uint32_t* __restrict__ keys,
                                                        int32_t* __restrict__ idx,
                                                        int32_t* __restrict__ n_dev) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i == 0) *n_dev = n;
  if (i >= n) return;
  const int ns = g.shift == 12 ? 16 : 4;
  int sx, sy, sz;
  const int cx = grid_axis_sub(x[i], g.lo[0], g.cell, g.dims[0], ns, sx);
  const int cy = grid_axis_sub(y[i], g.lo[1], g.cell, g.dims[1], ns, sy);
  const int cz = grid_axis_sub(z[i], g.lo[2], g.cell, g.dims[2], ns, sz);
  const int c = cx + cy * g.dims[0] + cz * g.dims[0] * g.dims[1];
  const uint32_t sub = (spread4(sx) << 2) | (spread4(sy) << 1) | spread4(sz);
  pcell[i] = c;
  keys[i] = g.shift ? ((uint32_t)c << g.shift) | sub : (uint32_t)c;
  idx[i] = i;
}

// Merge of the cell numbers 0..total (total is a sentinel) with the cells of the
// n sorted keys (key >> shift), cells placed before keys of equal cell. Cell c
// then sits after exactly start[c] = #{keys of cell < c} keys, and key k sits at
// merged position k + cell_k + 1 (strictly increasing in k). A thread finds its first key by
// binary search on that position, then walks kMergeSpan merged elements,
// writing start[c] and end[c-1] (= start[c]) for every cell it passes.
__global__ __launch_bounds__(kThreads) void grid_ranges(const uint32_t* __restrict__ keys, int n,
                                                        int total, int shift,
                                                        int32_t* __restrict__ start,
                                                        int32_t* __restrict__ end) {
  const long long m = (long long)total + 1 + n;
  const long long q0 = ((long long)blockIdx.x * kThreads + threadIdx.x) * kMergeSpan;
  if (q0 >= m) return;
  int lo = 0, hi = n;  // first key whose merged position is >= q0
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((long long)mid + (keys[mid] >> shift) + 1 < q0)
      lo = mid + 1;
    else
      hi = mid;
  }
  int j = lo;
  long long c = q0 - j;  // cells before q0
  const long long q1 = q0 + kMergeSpan < m ? q0 + kMergeSpan : m;
  for (long long q = q0; q < q1; ++q) {
    if (j < n && (long long)j + (keys[j] >> shift) + 1 == q) {
      ++j;
      continue;
    }
    if (c < total) start[c] = j;
    if (c >= 1) end[c - 1] = j;
    ++c;
  }
}

// Bodies in sorted order with G*m (:144 forms G * mass[j] first), padded with
// zeros to a multiple of kNB, then the pad group (kGroup zero-mass bodies at
// 1e18, beyond any cutoff), and the bounding box of every aligned group of
// kGroup bodies (box[2g] = min, box[2g+1] = max over the group's real bodies).
// The grid covers the padded length; the groups' lanes sit in one wave.
// ublk[block] = 1 when every body of the block has the G m of sorted body 0.
__global__ __launch_bounds__(kThreads) void grid_gather(const int32_t* __restrict__ idx,
                                                        const float* __restrict__ x,
                                                        const float* __restrict__ y,
                                                        const float* __restrict__ z,
                                                        const float* __restrict__ m, int n,
                                                        int n_pad, float G,
                                                        float4* __restrict__ body,
                                                        float4* __restrict__ box,
                                                        int32_t* __restrict__ ublk) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  float4 b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (k < n) {
    const int i = idx[k];
    b = make_float4(x[i], y[i], z[i], mul_rn(G, m[i]));
    lo[0] = hi[0] = b.x;
    lo[1] = hi[1] = b.y;
    lo[2] = hi[2] = b.z;
  }
#pragma unroll
  for (int off = 1; off < kGroup; off <<= 1)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], off, 64));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off, 64));
    }
  const float gm0 = mul_rn(G, m[idx[0]]);
  const bool same = __syncthreads_and(k >= n || b.w == gm0);
  if (threadIdx.x == 0) ublk[blockIdx.x] = same ? 1 : 0;
  if (k < kGroup) body[n_pad + k] = make_float4(1.0e18f, 1.0e18f, 1.0e18f, 0.0f);
  if (k >= n_pad) return;
  body[k] = b;
  if ((k & (kGroup - 1)) == 0) {
    box[2 * (k / kGroup)] = make_float4(lo[0], lo[1], lo[2], 0.0f);
    box[2 * (k / kGroup) + 1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
  }
}

// uni = 1 when every block of the gather found one G m (all bodies share it).
__global__ __launch_bounds__(kThreads) void grid_uniform(const int32_t* __restrict__ ublk, int nb,
                                                         int32_t* __restrict__ uni) {
  int ok = 1;
  for (int i = threadIdx.x; i < nb; i += kThreads) ok &= ublk[i];
  ok = __syncthreads_and(ok);
  if (threadIdx.x == 0) *uni = ok;
}

// r^2 of one pair as fma(rz, rz, fma(rx, rx, ry*ry)): the fused form an LLVM-based
// compiler gives the reference's dx*dx + dy*dy + dz*dz (:137) under its default
// contraction. The oracle evaluates the same form, so the pairs kept by
// r^2 < cutoff^2 are exactly the same on both sides.
__device__ __forceinline__ float grid_d2(float rx, float ry, float rz) {
  return __builtin_fmaf(rz, rz, __builtin_fmaf(rx, rx, mul_rn(ry, ry)));
}

// v_cndmask with the keep mask in SGPRs (a plain select may be turned into a
// branch around the force factor).
__device__ __forceinline__ float keep_or_zero(uint64_t keep, float v) {
  float r;
  asm("v_cndmask_b32_e64 %0, 0, %1, %2" : "=v"(r) : "v"(v), "s"(keep));
  return r;
}

// The bodies of kPer groups of kGroup: j[0] .. j[0]+kGroup-1, then j[1] ...
__device__ __forceinline__ void load_groups(const float4* __restrict__ body,
                                            const int (&j)[kPer], float4 (&b)[kNB]) {
#pragma unroll
  for (int g = 0; g < kPer; ++g)
#pragma unroll
    for (int q = 0; q < kGroup; ++q) b[g * kGroup + q] = body[j[g] + q];
}

// kNB bodies (wave-uniform, in SGPRs: the groups at j[0 .. kPer-1]) against
// each lane's target, issued stage by stage over independent pairs: differences,
// then (once this batch's operands are in) the loads of the next groups (nj),
// r^2 and the masks, the rsq back to back, the force factors, the
// accumulations. Per pair (:133-149): keep = r^2 < cutoff2 [and j != self] [and
// j in the lane's own range], f = (G m_j) ((inv inv) inv), inv = rsq(r^2 + eps2),
// a += f r. UNI: every body has the same G m, which the caller applies to the
// finished sum instead (f = (inv inv) inv).
template <bool SELF, bool MEMBER, bool UNI>
__device__ __forceinline__ void grid_batch(const float4 (&b)[kNB], const int (&jg)[kPer],
                                           const float4* __restrict__ body,
                                           const int (&nj)[kPer],
                                           float4 (&nxt)[kNB], int own_lo, int own_hi, int self,
                                           float xi, float yi, float zi, float cutoff2,
                                           float eps2, float& fx, float& fy, float& fz) {
  float rx[kNB], ry[kNB], rz[kNB], w[kNB];
#pragma unroll
  for (int q = 0; q < kNB; ++q) {
    rx[q] = sub_rn(b[q].x, xi);
    ry[q] = sub_rn(b[q].y, yi);
    rz[q] = sub_rn(b[q].z, zi);
  }
  __builtin_amdgcn_sched_barrier(0);
  load_groups(body, nj, nxt);
  __builtin_amdgcn_sched_barrier(0);
  // A dropped pair enters v_rsq as +inf, whose result +0 makes its term an
  // exact +0 (the same value the per-lane walk's mask gives): no lane mask
  // has to live across the batch, which keeps the SGPRs to the bodies.
#pragma unroll
  for (int q = 0; q < kNB; ++q) {
    const int j = jg[q / kGroup] + q % kGroup;
    const float d2 = grid_d2(rx[q], ry[q], rz[q]);
    bool k = d2 < cutoff2;
    if (SELF) k = k && j != self;
    if (MEMBER) k = k && j >= own_lo && j < own_hi;
    w[q] = k ? add_rn(d2, eps2) : __builtin_inff();
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int q = 0; q < kNB; ++q) w[q] = __builtin_amdgcn_rsqf(w[q]);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int q = 0; q < kNB; ++q) {
    const float inv = w[q];
    w[q] = UNI ? mul_rn(mul_rn(inv, inv), inv) : mul_rn(b[q].w, mul_rn(mul_rn(inv, inv), inv));
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int q = 0; q < kNB; ++q) {
    fx = __builtin_fmaf(w[q], rx[q], fx);
    fy = __builtin_fmaf(w[q], ry[q], fy);
    fz = __builtin_fmaf(w[q], rz[q], fz);
  }
}

// One pair with the same arithmetic (tails of a walk, per-lane walks).
template <bool UNI>
__device__ __forceinline__ void grid_pair(float4 b, bool keep, float xi, float yi, float zi,
                                          float cutoff2, float eps2, float& fx, float& fy,
                                          float& fz) {
  const float rx = sub_rn(b.x, xi), ry = sub_rn(b.y, yi), rz = sub_rn(b.z, zi);
  const float d2 = grid_d2(rx, ry, rz);
  const uint64_t k = __builtin_amdgcn_ballot_w64(keep & (d2 < cutoff2));
  const float inv = __builtin_amdgcn_rsqf(add_rn(d2, eps2));
  const float c = mul_rn(mul_rn(inv, inv), inv);
  const float f = keep_or_zero(k, UNI ? c : mul_rn(b.w, c));
  fx = __builtin_fmaf(f, rx, fx);
  fy = __builtin_fmaf(f, ry, fy);
  fz = __builtin_fmaf(f, rz, fz);
}

// Walk of [ulo, uhi) with wave-uniform indices (scalar loads, SGPR operands) over
// the aligned groups of kGroup bodies that can hold a kept pair. Each lane tests
// one group's box against the box of the wave's targets; a group whose per-axis
// gaps (sx, sy, sz) give fma(sz, sz, fma(sx, sx, sy*sy)) >= cutoff2 is skipped.
// No pair of a skipped group can be kept: every pair difference is at least the
// gap on each axis, the gaps are rounded from the same coordinates by the same
// monotone operations, and r^2 is the same monotone expression. The groups left
// are evaluated in ascending order, two per batch, with the next two's loads in
// flight; a skipped group's pairs would only have added exact zeros, so the sums
// are those of a walk over every body.
// SELF: the range may hold a lane's own body. multi: the wave's targets lie in
// more than one cell, so each lane keeps only its own [own_lo, own_hi); edge
// groups keep [ulo, uhi) the same way.
template <bool SELF, bool UNI>
__device__ __forceinline__ void walk_culled(const float4* __restrict__ body,
                                            const float4* __restrict__ box, int ulo, int uhi,
                                            int pad, bool multi, int own_lo, int own_hi,
                                            int self, const float (&tlo)[3],
                                            const float (&thi)[3], float xi, float yi, float zi,
                                            float cutoff2, float eps2, float& fx, float& fy,
                                            float& fz) {
  if (ulo >= uhi) return;
  const int lane = (int)(threadIdx.x & 63u);
  const int g_first = ulo / kGroup, g_last = (uhi - 1) / kGroup;
  const int mlo = multi ? own_lo : ulo, mhi = multi ? own_hi : uhi;
  for (int gt = g_first; gt <= g_last; gt += 64) {
    const int gl = gt + lane;
    bool pass = false;
    if (gl <= g_last) {
      const float4 lo = box[2 * gl], hi = box[2 * gl + 1];
      const float sx = fmaxf(fmaxf(sub_rn(lo.x, thi[0]), sub_rn(tlo[0], hi.x)), 0.0f);
      const float sy = fmaxf(fmaxf(sub_rn(lo.y, thi[1]), sub_rn(tlo[1], hi.y)), 0.0f);
      const float sz = fmaxf(fmaxf(sub_rn(lo.z, thi[2]), sub_rn(tlo[2], hi.z)), 0.0f);
      pass = grid_d2(sx, sy, sz) < cutoff2;
    }
    uint64_t todo = __builtin_amdgcn_ballot_w64(pass);
    // Groups left, kPer per batch in ascending order; a missing group is the pad
    // group (zero mass, far beyond any cutoff: exact zeros). Once the groups run
    // out the walk takes pad as its next batch and stops there.
    auto next = [&](int (&j)[kPer]) {
#pragma unroll
      for (int g = 0; g < kPer; ++g) {
        j[g] = todo ? (gt + (int)__builtin_ctzll(todo)) * kGroup : pad;
        todo &= todo - 1;
      }
    };
    if (!todo) continue;
    int ja[kPer], jb[kPer];
    next(ja);
    float4 ba[kNB], bb[kNB];
    load_groups(body, ja, ba);
    for (;;) {  // two batches per turn, so the buffers alternate without copies
      next(jb);
      if (multi || ja[0] < ulo || ja[kPer - 1] + kGroup > uhi)
        grid_batch<SELF, true, UNI>(ba, ja, body, jb, bb, mlo, mhi, self, xi, yi, zi, cutoff2,
                                    eps2, fx, fy, fz);
      else
        grid_batch<SELF, false, UNI>(ba, ja, body, jb, bb, 0, 0, self, xi, yi, zi, cutoff2, eps2,
                                     fx, fy, fz);
      if (jb[0] == pad) break;
      next(ja);
      if (multi || jb[0] < ulo || jb[kPer - 1] + kGroup > uhi)
        grid_batch<SELF, true, UNI>(bb, jb, body, ja, ba, mlo, mhi, self, xi, yi, zi, cutoff2,
                                    eps2, fx, fy, fz);
      else
        grid_batch<SELF, false, UNI>(bb, jb, body, ja, ba, 0, 0, self, xi, yi, zi, cutoff2, eps2,
                                     fx, fy, fz);
      if (ja[0] == pad) break;
    }
  }
}

// Walk of a lane's own [lo, hi) with per-lane (vector) loads.
template <bool UNI>
__device__ __forceinline__ void walk_lane(const float4* __restrict__ body, int lo, int hi,
                                          int self, float xi, float yi, float zi, float cutoff2,
                                          float eps2, float& fx, float& fy, float& fz) {
  for (int j = lo; j < hi; ++j)
    grid_pair<UNI>(body[j], j != self, xi, yi, zi, cutoff2, eps2, fx, fy, fz);
}

__device__ __forceinline__ int wave_min(int v) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_max(int v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ float wave_fmin(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float wave_fmax(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// Workgroups are dealt to the 8 XCDs round-robin; renumber them so that each
// XCD takes one contiguous run of sorted targets, whose neighbour cells then
// stay in that XCD's L2.
__device__ __forceinline__ int xcd_block(int i, int nb) {
  const int per = nb / 8, rem = nb % 8;
  const int xcd = i % 8, slot = i / 8;
  return xcd < rem ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

// mode 0: uniform walk where a wave's targets share a row; 1: per-lane walks only.
// UNI: every body has the same G m (that of sorted body 0), applied to each
// target's finished sum.
template <bool UNI>
__device__ __forceinline__ void grid_forces_body(const float4* __restrict__ body,
                                                 const float4* __restrict__ box,
                                                 const int32_t* __restrict__ idx,
                                                 const int32_t* __restrict__ cstart,
                                                 const int32_t* __restrict__ cend, GridGeom g,
                                                 float cutoff2, float eps2, int n, int n_pad,
                                                 int mode, float* __restrict__ ax,
                                                 float* __restrict__ ay, float* __restrict__ az) {
  const int b = xcd_block((int)blockIdx.x, (int)gridDim.x);
  const int k0 = b * kThreads + (int)(threadIdx.x & ~63u);
  if (k0 >= n) return;  // the whole wave is past the end
  const int k = k0 + (int)(threadIdx.x & 63u);
  const bool active = k < n;
  const int kc = active ? k : n - 1;  // idle lanes mirror the wave's last target
  const float4 me = body[kc];
  const int cx = grid_axis(me.x, g.lo[0], g.cell, g.dims[0]);
  const int cy = grid_axis(me.y, g.lo[1], g.cell, g.dims[1]);
  const int cz = grid_axis(me.z, g.lo[2], g.cell, g.dims[2]);
  const int lo_x = max(cx - 1, 0), hi_x = min(cx + 1, g.dims[0] - 1);
  const int dxy = g.dims[0] * g.dims[1];
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  const int row = cy + cz * g.dims[1];
  const bool one_row = mode == 0 && __all(row == __builtin_amdgcn_readfirstlane(row));
  const bool self_test = eps2 < kUncheckedMinEps2;
  if (one_row) {
    const int cy0 = __builtin_amdgcn_readfirstlane(cy), cz0 = __builtin_amdgcn_readfirstlane(cz);
    const int xmin = wave_min(cx), xmax = wave_max(cx);
    const int ulo_x = max(xmin - 1, 0), uhi_x = min(xmax + 1, g.dims[0] - 1);
    const bool multi = xmin != xmax;
    float tlo[3] = {wave_fmin(me.x), wave_fmin(me.y), wave_fmin(me.z)};
    float thi[3] = {wave_fmax(me.x), wave_fmax(me.y), wave_fmax(me.z)};
    // operands of vector instructions only, kept in VGPRs (an SGPR operand
    // halves an FP32 instruction's issue rate, and the SGPRs hold the bodies)
    float cut2v = cutoff2, eps2v = eps2;
    asm volatile("" : "+v"(tlo[0]), "+v"(tlo[1]), "+v"(tlo[2]), "+v"(thi[0]), "+v"(thi[1]),
                 "+v"(thi[2]), "+v"(cut2v), "+v"(eps2v));
    for (int dz = -1; dz <= 1; ++dz) {
      const int nz = cz0 + dz;
      if (nz < 0 || nz >= g.dims[2]) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        const int ny = cy0 + dy;
        if (ny < 0 || ny >= g.dims[1]) continue;
        const int base = ny * g.dims[0] + nz * dxy;
        const int ulo = cstart[base + ulo_x], uhi = cend[base + uhi_x];
        const int olo = multi ? cstart[base + lo_x] : 0, ohi = multi ? cend[base + hi_x] : 0;
        // the self test is needed only where a lane's own body lies in the range
        // (the centre row, unless positions moved since the build), and only for
        // eps2 below kUncheckedMinEps2: above it the self pair adds an exact zero
        if (self_test && __any((k >= ulo) & (k < uhi)))
          walk_culled<true, UNI>(body, box, ulo, uhi, n_pad, multi, olo, ohi, k, tlo, thi, me.x,
                                 me.y, me.z, cut2v, eps2v, fx, fy, fz);
        else
          walk_culled<false, UNI>(body, box, ulo, uhi, n_pad, multi, olo, ohi, k, tlo, thi,
                                  me.x, me.y, me.z, cut2v, eps2v, fx, fy, fz);
      }
    }
  } else {
    for (int dz = -1; dz <= 1; ++dz) {
      const int nz = cz + dz;
      for (int dy = -1; dy <= 1; ++dy) {
        const int ny = cy + dy;
        if (nz < 0 || nz >= g.dims[2] || ny < 0 || ny >= g.dims[1]) continue;
        const int base = ny * g.dims[0] + nz * dxy;
        walk_lane<UNI>(body, cstart[base + lo_x], cend[base + hi_x], k, me.x, me.y, me.z,
                       cutoff2, eps2, fx, fy, fz);
      }
    }
  }
  if (UNI) {
    const float gm = body[0].w;
    fx = mul_rn(gm, fx);
    fy = mul_rn(gm, fy);
    fz = mul_rn(gm, fz);
  }
  if (active) {
    const int i = idx[k];
    ax[i] = fx;
    ay[i] = fy;
    az[i] = fz;
  }
}

__global__ __launch_bounds__(kThreads) void grid_forces(const float4* __restrict__ body,
                                                        const float4* __restrict__ box,
                                                        const int32_t* __restrict__ idx,
                                                        const int32_t* __restrict__ cstart,
                                                        const int32_t* __restrict__ cend,
                                                        GridGeom g, float cutoff2, float eps2,
                                                        int n, int n_pad, int mode,
                                                        const int32_t* __restrict__ uni,
                                                        float* __restrict__ ax,
                                                        float* __restrict__ ay,
                                                        float* __restrict__ az) {
  if (*uni)
    grid_forces_body<true>(body, box, idx, cstart, cend, g, cutoff2, eps2, n, n_pad, mode, ax,
                           ay, az);
  else
    grid_forces_body<false>(body, box, idx, cstart, cend, g, cutoff2, eps2, n, n_pad, mode, ax,
                            ay, az);
}

int bits_for(int total) {  // key bits covering cell numbers 0 .. total-1
  int b = 0;
  while (b < 31 && (1LL << b) < (long long)total) ++b;
  return b;
}

}  // namespace
}  // namespace nbx

using nbx::GridGeom;

struct nbx_grid {
  size_t max_particles = 0;
  float cell = 0.0f;
  int bbox_blocks = 0;
  std::vector<void*> allocs;
  float* bbox_part = nullptr;
  float* box_dev = nullptr;
  float* box_host = nullptr;  // pinned, 6 floats
  int32_t* pcell = nullptr;
  uint32_t* keys = nullptr;
  int32_t* idx = nullptr;
  int32_t* n_dev = nullptr;
  float4* body = nullptr;  // sorted bodies, padded to a multiple of kNB
  float4* group_box = nullptr;  // per group of kGroup sorted bodies: min, max
  int32_t* ublk = nullptr;       // per gather block: its bodies share sorted body 0's G m
  int32_t* uni = nullptr;        // every body shares it
  int32_t* cstart = nullptr;
  int32_t* cend = nullptr;
  size_t cell_cap = 0;
  nbx::RadixSort sort;
  GridGeom geom{};
  float box[6] = {0, 0, 0, 0, 0, 0};  // padded bbox of the last build
  int n = 0;                          // particles of the last build (0: not built)
  int mode = 0;
};

namespace nbx {
namespace {

template <typename T>
int galloc(nbx_grid* g, T** p, size_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(NBX_ERR_NOMEM, "spatial hash buffers: allocation of %zu bytes failed",
                sizeof(T) * count);
  }
  g->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return NBX_OK;
}

void grid_release(nbx_grid* g) {
  for (void* p : g->allocs) (void)hipFree(p);
  g->allocs.clear();
  if (g->cstart) (void)hipFree(g->cstart);
  if (g->cend) (void)hipFree(g->cend);
  g->cstart = g->cend = nullptr;
  if (g->box_host) (void)hipHostFree(g->box_host);
  g->box_host = nullptr;
  g->sort.release();
}

int grid_blocks(size_t n, int cap) {
  return (int)std::max<size_t>(1, std::min<size_t>((n + kThreads - 1) / kThreads, (size_t)cap));
}

}  // namespace
}  // namespace nbx

using namespace nbx;

extern "C" {

int nbx_grid_create(size_t max_particles, float cell_size, nbx_grid** out) {
  if (!out) return fail(NBX_ERR_INVALID, "nbx_grid_create: null output");
  *out = nullptr;
  if (max_particles == 0) return fail(NBX_ERR_INVALID, "spatial hash grid needs max_particles > 0");
  if (max_particles > (size_t)INT32_MAX - 1024)
    return fail(NBX_ERR_INVALID, "spatial hash grid supports at most 2^31-1025 particles");
  if (!(cell_size > 0.0f) || std::isinf(cell_size))
    return fail(NBX_ERR_INVALID, "Spatial hash cell size must be positive and finite");
  nbx_grid* g = new (std::nothrow) nbx_grid();
  if (!g) return fail(NBX_ERR_NOMEM, "nbx_grid_create: host allocation");
  const size_t P = max_particles;
  g->max_particles = P;
  g->cell = cell_size;
  g->bbox_blocks = grid_blocks(P, 1024);
  int rc = NBX_OK;
  auto A = [&](auto** p, size_t c) {
    if (rc == NBX_OK) rc = galloc(g, p, c);
  };
  A(&g->bbox_part, (size_t)g->bbox_blocks * 6);
  A(&g->box_dev, 6);
  A(&g->pcell, P);
  A(&g->keys, P);
  A(&g->idx, P);
  A(&g->n_dev, 1);
  const size_t P8 = (P + kNB - 1) / kNB * kNB;
  A(&g->body, P8 + kGroup);  // + the pad group
  A(&g->group_box, 2 * (P8 / kGroup));
  A(&g->ublk, (size_t)grid_blocks(P8, INT32_MAX));
  A(&g->uni, 1);
  if (rc == NBX_OK && hipHostMalloc((void**)&g->box_host, 6 * sizeof(float)) != hipSuccess) {
    (void)hipGetLastError();
    g->box_host = nullptr;
    rc = fail(NBX_ERR_NOMEM, "spatial hash: pinned host allocation failed");
  }
  if (rc == NBX_OK) rc = g->sort.init((int)P);
  if (rc != NBX_OK) {
    grid_release(g);
    delete g;
    return rc;
  }
  *out = g;
  return NBX_OK;
}

int nbx_grid_destroy(nbx_grid* g) {
  if (!g) return NBX_OK;
  grid_release(g);
  delete g;
  return NBX_OK;
}

int nbx_grid_build(nbx_grid* g, const nbx_particles* d, nbx_stream stream) {
  if (!g) return fail(NBX_ERR_INVALID, "nbx_grid_build: null grid");
  NBX_TRY(check_particles(d, "nbx_grid_build"));
  if (d->count > g->max_particles)
    return fail(NBX_ERR_CAPACITY, "spatial hash grid created for %zu particles, got %zu",
                g->max_particles, d->count);
  hipStream_t s = as_stream(stream);
  const int n = (int)d->count;
  g->n = 0;
  // bbox (:190-239): device reduction, one 24-byte read-back
  const int bb = std::min(g->bbox_blocks, grid_blocks(n, 1024));
  hipLaunchKernelGGL(grid_bbox_partial, dim3(bb), dim3(kThreads), 0, s, d->pos_x, d->pos_y,
                     d->pos_z, n, g->bbox_part);
  hipLaunchKernelGGL(grid_bbox_final, dim3(1), dim3(kThreads), 0, s, g->bbox_part, bb,
                     g->box_dev);
  NBX_LAUNCHED();
  NBX_HIP(hipMemcpyAsync(g->box_host, g->box_dev, 6 * sizeof(float), hipMemcpyDeviceToHost, s));
  NBX_HIP(hipStreamSynchronize(s));
  float lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = g->box_host[a] - 0.001f;  // :232-238, FP32 on the host
    hi[a] = g->box_host[3 + a] + 0.001f;
  }
  // dims (:249-260)
  GridGeom geom{};
  long long total = 1;
  bool too_large = false;
  for (int a = 0; a < 3; ++a) {
    const float r = (hi[a] - lo[a]) / g->cell;
    if (!(r <= 1e8f)) {  // also catches non-finite positions
      too_large = true;
      break;
    }
    geom.dims[a] = (int)std::ceil(r) + 1;
    total *= geom.dims[a];
    if (total > kMaxCells) too_large = true;
  }
  if (too_large)
    return fail(NBX_ERR_CAPACITY, "Spatial hash grid too large: reduce cell_size or bounding box");
  for (int a = 0; a < 3; ++a) {
    geom.lo[a] = lo[a];
    g->box[a] = lo[a];
    g->box[3 + a] = hi[a];
  }
  geom.cell = g->cell;
  geom.total = (int)total;
  geom.shift = total <= kSub16MaxCells ? 12 : (total <= kSub4MaxCells ? 6 : 0);
  if ((size_t)total > g->cell_cap) {  // grow-only cell arrays
    if (g->cstart) (void)hipFree(g->cstart);
    if (g->cend) (void)hipFree(g->cend);
    g->cstart = g->cend = nullptr;
    g->cell_cap = 0;
    if (hipMalloc((void**)&g->cstart, sizeof(int32_t) * total) != hipSuccess ||
        hipMalloc((void**)&g->cend, sizeof(int32_t) * total) != hipSuccess) {
      (void)hipGetLastError();
      return fail(NBX_ERR_NOMEM, "spatial hash: allocation of %lld cells failed", total);
    }
    g->cell_cap = (size_t)total;
  }
  g->geom = geom;
  // assign (:281-285), stable sort by cell, ranges (:287-300)
  hipLaunchKernelGGL(grid_assign, dim3(grid_blocks(n, INT32_MAX)), dim3(kThreads), 0, s, d->pos_x,
                     d->pos_y, d->pos_z, n, geom, g->pcell, g->keys, g->idx, g->n_dev);
  NBX_LAUNCHED();
  NBX_TRY(g->sort.sort(g->keys, g->idx, g->n_dev, bits_for(geom.total) + geom.shift, s));
  const long long merged = total + 1 + n;
  const long long threads = (merged + kMergeSpan - 1) / kMergeSpan;
  hipLaunchKernelGGL(grid_ranges, dim3((unsigned)((threads + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, s, g->keys, n, geom.total, geom.shift, g->cstart,
                     g->cend);
  NBX_LAUNCHED();
  g->n = n;
  return NBX_OK;
}

int nbx_grid_forces(nbx_grid* g, nbx_particles* d, float cutoff, float G, float eps,
                    nbx_stream stream) {
  if (!g) return fail(NBX_ERR_INVALID, "nbx_grid_forces: null grid");
  NBX_TRY(check_particles(d, "nbx_grid_forces"));
  if (g->n == 0) return fail(NBX_ERR_STATE, "spatial hash forces before a successful build");
  if ((int)d->count != g->n)
    return fail(NBX_ERR_STATE, "spatial hash grid was built for %d particles, forces asked for %zu",
                g->n, d->count);
  hipStream_t s = as_stream(stream);
  const int n = g->n;
  const float cutoff2 = cutoff * cutoff;  // :320
  const float eps2 = eps * eps;
  const int n_pad = (n + kNB - 1) / kNB * kNB;
  const int gblocks = grid_blocks(n_pad, INT32_MAX);
  hipLaunchKernelGGL(grid_gather, dim3(gblocks), dim3(kThreads), 0, s, g->idx, d->pos_x, d->pos_y,
                     d->pos_z, d->mass, n, n_pad, G, g->body, g->group_box, g->ublk);
  hipLaunchKernelGGL(grid_uniform, dim3(1), dim3(kThreads), 0, s, g->ublk, gblocks, g->uni);
  hipLaunchKernelGGL(grid_forces, dim3(grid_blocks(n, INT32_MAX)), dim3(kThreads), 0, s, g->body,
                     g->group_box, g->idx, g->cstart, g->cend, g->geom, cutoff2, eps2, n, n_pad,
                     g->mode, g->uni, d->acc_x, d->acc_y, d->acc_z);
  NBX_LAUNCHED();
  return NBX_OK;
}

int nbx_grid_info(nbx_grid* g, int32_t dims[3], float* cell_size, int32_t* total_cells,
                  float bbox[6]) {
  if (!g) return fail(NBX_ERR_INVALID, "nbx_grid_info: null grid");
  const bool built = g->geom.total > 0;
  for (int a = 0; a < 3; ++a)
    if (dims) dims[a] = built ? g->geom.dims[a] : 0;
  if (cell_size) *cell_size = g->cell;
  if (total_cells) *total_cells = built ? g->geom.total : 0;
  if (bbox)
    for (int a = 0; a < 6; ++a) bbox[a] = g->box[a];
  return NBX_OK;
}

int nbx_grid_copy_cells(nbx_grid* g, int32_t* cell_start, int32_t* cell_end,
                        int32_t* particle_cells, int32_t* sorted_indices) {
  if (!g) return fail(NBX_ERR_INVALID, "nbx_grid_copy_cells: null grid");
  NBX_HIP(hipDeviceSynchronize());
  const size_t total = g->geom.total > 0 ? (size_t)g->geom.total : 0;
  const size_t n = (size_t)g->n, P = g->max_particles;
  if (cell_start && total)
    NBX_HIP(hipMemcpy(cell_start, g->cstart, total * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (cell_end && total)
    NBX_HIP(hipMemcpy(cell_end, g->cend, total * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (particle_cells) {
    if (n) NBX_HIP(hipMemcpy(particle_cells, g->pcell, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::fill(particle_cells + n, particle_cells + P, -1);
  }
  if (sorted_indices) {
    if (n) NBX_HIP(hipMemcpy(sorted_indices, g->idx, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    std::fill(sorted_indices + n, sorted_indices + P, -1);
  }
  return NBX_OK;
}

int nbx_grid_set_walk(nbx_grid* g, int mode) {
  if (!g) return fail(NBX_ERR_INVALID, "nbx_grid_set_walk: null grid");
  if (mode != 0 && mode != 1) return fail(NBX_ERR_INVALID, "nbx_grid_set_walk: mode must be 0 or 1");
  g->mode = mode;
  return NBX_OK;
}

}  // extern "C"
