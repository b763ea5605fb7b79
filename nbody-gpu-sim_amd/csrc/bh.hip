// Barnes-Hut pipeline on the GPU, bit-exact with the reference's host build
// (src/cuda/force_barnes_hut.cu). Every phase runs on the device and every
// count the build needs lives in device memory, so a build issues no host
// round trip.
//
//  1 bbox       exact per-axis min/max (:78-122, :266-306), then the Morton
//               reciprocals (:313-320) and the root cell (:418-422) on device
//  2 morton     30-bit codes of the per-axis normalised positions (:125-139)
//  3 sort       stable LSD radix sort of (code, index)        (:324-328)
//  4 paths      each particle's octant digit at depths 0..31, walking the
//               float centre chain c +- h/2 exactly as allocateLeaf does (:375-379)
//  5 tree       the PR-octree from the particles' order by octant path: a
//               second sort (by path), adjacent common prefixes, one scan. The
//               cells come out numbered in the traversal's preorder, with
//               boxes, counts and subtree ends (see section 5)
//  6 com        bottom-up, children in octant order, IEEE float ops (:512-538);
//               depth-32 cells holding several particles merge them in
//               insertion order (:355-365, :490-507)
//  numbering    on demand (nbx_bh_copy_nodes), the reference's node indices.
//               The reference inserts particles one by one in sorted order; its
//               numbering follows from two ranks per cell: t1 = first inserted
//               particle, t2 = second. A non-root cell X with parent P is
//               created at insertion k(X) = max(t2(P), t1(X)); within one
//               insertion the displaced particle's chain comes first by depth,
//               then the new particle's leaf (t1 == k). Sorting the cells by
//               (k, t1 == k, depth) reproduces the reference indices exactly.
//  7 traverse   theta-criterion traversal with the reference's acceptance test,
//               self-exclusion rule and DFS order (:142-249), one shared stack
//               per wave (lanes carry a 64-bit "still needs this subtree" mask),
//               so node data is read with wave-uniform scalar loads.
//
// When the reference's node budget could stop a split (node_count + 33 >=
// 2*max_particles, :449-453) the tree depends on insertion order in ways the
// parallel build does not model; then a single-thread kernel replays the
// reference's insertion loop exactly. Both builders write the same 76-byte
// node array.
#include <algorithm>
#include <climits>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"
#include "bbox.hpp"
#include "radix.hpp"

namespace nbx {

namespace {

constexpr int kMaxDepth = 32;  // force_barnes_hut.cu:24
constexpr int kThreads = 256;

// 76-byte node, byte layout of nbody::OctreeNode (barnes_hut_tree.hpp:9-30).
struct RefNode {
  float cx, cy, cz, half;
  float comx, comy, comz, mass;
  int32_t child[8];
  int32_t pidx;
  uint8_t leaf;
  uint8_t pad[3];
  int32_t count;
};
static_assert(sizeof(RefNode) == NBX_OCTREE_NODE_BYTES, "OctreeNode layout");

// Device-resident bookkeeping of one build.
struct BhState {
  float bmin[3], bmax[3];
  float inv[3];
  float root_c[3];
  float root_h;
  int32_t n;
  int32_t node_count;  // parallel builder: cells
  int32_t deepest;      // deepest leaf
  int32_t multi_count;  // depth-32 leaves holding several particles
  int32_t run_count;    // runs of positions sharing the sorted path digits
  int32_t use_seq;
  int32_t final_count;  // node_count_ of the reference
  int32_t max_depth;    // max_depth_ of the reference
  int32_t budget_merges;
  int32_t depth_merges;
  int32_t long_runs;  // runs longer than kRefineMax (ordered by path_refine_long)
  int32_t sort_n;    // cells to number (0 when the sequential builder runs)
  int32_t truncated; // sequential builder: bodies left out when the node array ran out
};

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {  // :34-40
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__device__ __forceinline__ uint32_t quantise(float t) {  // :45-47
  const float c = fminf(fmaxf(t, 0.0f), 1.0f);
  return (uint32_t)mul_rn(c, 1023.0f);
}

__device__ __forceinline__ int octant_of(float px, float py, float pz, float cx, float cy,
                                         float cz) {  // :339-348
  return (px >= cx ? 1 : 0) | (py >= cy ? 2 : 0) | (pz >= cz ? 4 : 0);
}

// Sort key of ten digits stored 3 bits each from bit 0: digit i becomes 7 - digit
// at bits 3 (9 - i), so the first digit is the most significant.
__device__ __forceinline__ uint32_t order_key10(uint32_t w) {
  uint32_t k = 0;
#pragma unroll
  for (int i = 0; i < 10; ++i) k |= ((~w >> (3 * i)) & 7u) << (3 * (9 - i));
  return k;
}

// Exclusive block scan of one int per thread (kThreads threads); returns the
// exclusive prefix and writes the block total.
__device__ int block_exclusive_scan(int v, int* total) {
  __shared__ int wsum[kThreads / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int incl = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int before = 0, tot = 0;
  for (int q = 0; q < kThreads / 64; ++q) {
    if (q < w) before += wsum[q];
    tot += wsum[q];
  }
  __syncthreads();
  *total = tot;
  return before + incl - v;
}

// ---------------------------------------------------------------- 1: bbox

__global__ __launch_bounds__(kThreads) void bbox_partial(const float* __restrict__ x,
                                                         const float* __restrict__ y,
                                                         const float* __restrict__ z, int n,
                                                         float* partial) {
  bbox_partial_body<kThreads>(x, y, z, n, partial);
}

__global__ __launch_bounds__(kThreads) void bbox_final(const float* partial, int nparts, int n,
                                                       BhState* st) {
  float lo[3], hi[3];
  if (!bbox_final_body<kThreads>(partial, nparts, lo, hi)) return;
  float ext[3];
  for (int a = 0; a < 3; ++a) {
    st->bmin[a] = lo[a];
    st->bmax[a] = hi[a];
    const float size = add_rn(sub_rn(hi[a], lo[a]), 0.001f);  // :313-315
    st->inv[a] = div_rn(1.0f, size);                          // :319-320
    st->root_c[a] = mul_rn(add_rn(lo[a], hi[a]), 0.5f);       // :418-419
    ext[a] = sub_rn(hi[a], lo[a]);
  }
  float big = ext[0];  // std::max({...}) :420-421
  if (ext[1] > big) big = ext[1];
  if (ext[2] > big) big = ext[2];
  st->root_h = add_rn(mul_rn(big, 0.5f), 0.001f);  // :422
  st->n = n;
  st->deepest = 0;
  st->multi_count = 0;
  st->run_count = 0;
  st->long_runs = 0;
  st->truncated = 0;
}

// ---------------------------------------------------------------- 2: morton

// Also packs each particle as (x, y, z, m), so that the paths pass gathers one
// 16-byte record per particle through the sort permutation instead of four.
// Bits 0..20 of v to bits 0, 3, ..., 60.
__device__ __forceinline__ uint64_t spread3_64(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

// Skilling's Hilbert index (AIP Conf. Proc. 707, 381 (2004)) of a cell of the
// 2^B x 2^B x 2^B grid: the transpose form of the three B-bit coordinates, then
// its bits interleaved, the first axis most significant (3B bits).
template <int B>
__device__ __forceinline__ uint64_t hilbert_index(uint32_t x, uint32_t y, uint32_t z) {
  uint32_t X[3] = {x, y, z};
  for (uint32_t q = 1u << (B - 1); q > 1; q >>= 1) {
    const uint32_t p = q - 1;
    for (int i = 0; i < 3; ++i) {
      if (X[i] & q) {
        X[0] ^= p;
      } else {
        const uint32_t t = (X[0] ^ X[i]) & p;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  X[1] ^= X[0];
  X[2] ^= X[1];
  uint32_t t = 0;
  for (uint32_t q = 1u << (B - 1); q > 1; q >>= 1)
    if (X[2] & q) t ^= q - 1;
  return (spread3_64(X[0] ^ t) << 2) | (spread3_64(X[1] ^ t) << 1) | spread3_64(X[2] ^ t);
}

// The traversal's task key: the Hilbert index of the particle's cell of the
// 2048^3 grid over the per-axis normalised box (33 bits), its top 32 bits.
// (Task order only: forces do not depend on it.)
__device__ __forceinline__ uint32_t task_hilbert_key(float nx, float ny, float nz) {
  constexpr int B = 11;
  const float top = (float)((1u << B) - 1);
  const uint32_t hx = (uint32_t)(fminf(fmaxf(nx, 0.0f), 1.0f) * top);
  const uint32_t hy = (uint32_t)(fminf(fmaxf(ny, 0.0f), 1.0f) * top);
  const uint32_t hz = (uint32_t)(fminf(fmaxf(nz, 0.0f), 1.0f) * top);
  return (uint32_t)(hilbert_index<B>(hx, hy, hz) >> (3 * B - 32));
}

__global__ __launch_bounds__(kThreads) void morton_kernel(const float* __restrict__ x,
                                                          const float* __restrict__ y,
                                                          const float* __restrict__ z,
                                                          const float* __restrict__ m, int n,
                                                          const BhState* __restrict__ st,
                                                          uint32_t* codes, int32_t* idx,
                                                          float4* packed, uint32_t* hkey) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float px = x[i], py = y[i], pz = z[i];
  packed[i] = make_float4(px, py, pz, m[i]);
  const float nx = mul_rn(sub_rn(px, st->bmin[0]), st->inv[0]);
  const float ny = mul_rn(sub_rn(py, st->bmin[1]), st->inv[1]);
  const float nz = mul_rn(sub_rn(pz, st->bmin[2]), st->inv[2]);
  const uint32_t qx = quantise(nx), qy = quantise(ny), qz = quantise(nz);
  codes[i] = (expand_bits(qx) << 2) | (expand_bits(qy) << 1) | expand_bits(qz);
  idx[i] = i;
  hkey[i] = task_hilbert_key(nx, ny, nz);  // the traversal's task order (task_keys)
}

// ---------------------------------------------------------------- 4: paths

// Also copies each particle to its rank as (x, y, z, m), and writes the keys of
// the path sort: digits 10..19 (sorted first) and digits 0..9 (see section 5).
__global__ __launch_bounds__(kThreads) void path_kernel(const float4* __restrict__ packed,
                                                        const int32_t* __restrict__ sidx, int n,
                                                        const BhState* __restrict__ st,
                                                        uint4* path, float4* sbody,
                                                        uint32_t* key, uint32_t* key_hi,
                                                        int32_t* rank) {
  const int r = blockIdx.x * kThreads + threadIdx.x;
  if (r >= n) return;
  const float4 b = packed[sidx[r]];
  const float px = b.x, py = b.y, pz = b.z;
  sbody[r] = b;
  float cx = st->root_c[0], cy = st->root_c[1], cz = st->root_c[2], h = st->root_h;
  uint32_t w[4] = {0, 0, 0, 0};
  for (int d = 0; d < kMaxDepth; ++d) {
    const int o = octant_of(px, py, pz, cx, cy, cz);
    w[d / 10] |= (uint32_t)o << (3 * (d % 10));
    const float hs = mul_rn(h, 0.5f);  // :375
    cx = add_rn(cx, (o & 1) ? hs : -hs);
    cy = add_rn(cy, (o & 2) ? hs : -hs);
    cz = add_rn(cz, (o & 4) ? hs : -hs);
    h = hs;
  }
  path[r] = make_uint4(w[0], w[1], w[2], w[3]);
  key[r] = order_key10(w[1]);
  key_hi[r] = order_key10(w[0]);
  rank[r] = r;
}

// Traversal layout: cells in the order the reference's stack walk visits them
// (:160-241: children pushed 0..7, so popped 7..0), i.e. depth-first preorder
// with children 7 down to 0.
// Record p (64 bytes, so that one scalar address reaches all of it):
//   rec[4p]   = (centre of mass, G * mass; written by trav_prep per force call)
//   rec[4p+1] = (cell centre, half size; negative for a leaf)
//   rec[4p+2] = (position after the cell's subtree as int bits, the opening-test
//                size (2h)^2 = 4 h h rounded as the traversal's test (:218-224)
//                rounds it, S_open, S_acc; the last two written by trav_prep)
//   rec[4p+3] = (mass, 0, 0, 0)
// A leaf's (2h)^2 is written as -1: its thresholds are negative for every
// theta >= 0, so the theta test never opens it, and the walk needs the leaf flag
// only where a lane's own particle lies in the cell. A node of zero mass, which
// the reference skips with its whole subtree (:172-173), gets its centre of mass
// and cell centre at kFarPos, a zero half size and the same -1: no particle lies
// in it, it is never opened, and its force term is an exact zero (G * 0 times a
// finite inverse cube).
constexpr float kFarPos = 1e18f;
constexpr float kNeverOpen = -1.0f;

// The per-call fields of a record for one (theta, G): G * mass and the opening
// thresholds S_open/S_acc (their bounds are stated at trav_prep). The layout
// passes write them for the (theta, G) of the last force call, so that a force
// call with the same values needs no trav_prep pass.
struct TravParams {
  float inv_theta2, G;
  int valid;
};
__device__ __forceinline__ void trav_fields(float size2, float mass, float inv_theta2, float G,
                                            float* gm, float* s_open, float* s_acc) {
  const float sc = mul_rn(size2, inv_theta2);
  *gm = mul_rn(G, mass);
  // Leaves and zero-mass cells (size2 < 0) are never opened by the test, so
  // their S_open holds -mass (<= 0, still never above d2): the walk's common
  // step compares a lane's own mass with a leaf's from it.
  *s_open = size2 < 0.0f ? -mass : mul_rn(sc, 1.0f - 0x1p-18f);
  *s_acc = mul_rn(sc, 1.0f + 0x1p-18f);
}

__device__ __forceinline__ void write_trav(float4* rec, int p, int size, float4 com, float4 box,
                                           bool leaf, TravParams tp) {
  const bool empty = com.w == 0.0f;
  const float size2 = (leaf || empty) ? kNeverOpen : mul_rn(mul_rn(4.0f, box.w), box.w);
  float gm = com.w, s_open = 0.0f, s_acc = 0.0f;
  if (tp.valid) trav_fields(size2, com.w, tp.inv_theta2, tp.G, &gm, &s_open, &s_acc);
  if (empty) {
    rec[4 * p] = make_float4(kFarPos, kFarPos, kFarPos, gm);
    rec[4 * p + 1] = make_float4(kFarPos, kFarPos, kFarPos, leaf ? -0.0f : 0.0f);
  } else {
    rec[4 * p] = make_float4(com.x, com.y, com.z, gm);
    rec[4 * p + 1] = make_float4(box.x, box.y, box.z, leaf ? -box.w - 0.0f : box.w);
  }
  rec[4 * p + 2] = make_float4(__int_as_float(p + size), size2, s_open, s_acc);
  rec[4 * p + 3] = make_float4(com.w, 0.0f, 0.0f, 0.0f);
}

// ---------------------------------------------------------------- 5: tree

// Cells of the PR-octree from the particles' order by octant path.
//
// Sort the particles by their paths, comparing octants as 7 - o (ascending
// order then lists the children of a cell 7 down to 0, the order in which the
// reference's stack walk visits them, :160-241), ties by insertion rank. Every
// cell is then a run of consecutive positions. With lcp[j] the number of
// leading digits position j shares with position j - 1 (lcp[0] = lcp[n] = 0):
//   - position j is the first particle of exactly the cells at depths
//     lcp[j] + 1 .. L[j], L[j] = min(32, max(lcp[j], lcp[j + 1]) + 1), the
//     deepest being its leaf (a depth-32 leaf holds every particle with the
//     same 32 digits); the root, depth 0, is cell 0;
//   - numbering these cells position by position, parent before child, gives
//     the depth-first preorder of the tree with children 7..0:
//     pre(j, d) = first_pre[j] + d - lcp[j] - 1, first_pre = 1 + the exclusive
//     scan of L[j] - lcp[j];
//   - a cell whose run ends before position e has its subtree end (the
//     traversal's skip) at first_pre[e], and holds e - j particles.
// So the cells, their boxes and the whole traversal layout come from one sort,
// one scan and one pass per position. The centres of mass follow bottom-up:
// each workgroup does the cells inside its block of positions level by level,
// then one workgroup does the few cells that straddle block boundaries.
//
// The path sort is a stable LSD radix sort over digits 0..19 (two 30-bit
// keys). Particles whose first 20 digits agree, rare outside exact or
// near-coincident duplicates and dense clusters, form runs that one workgroup
// per run orders by the remaining 12 digits and rank: in LDS up to kRefineMax
// positions (path_refine), by a stable radix sort through global scratch
// beyond that (path_refine_long).

// Cells ("tmp nodes"), indexed by preorder position.
struct TmpNodes {
  int32_t* t1;      // first inserted rank in the cell
  int32_t* t2;      // second inserted rank (INT_MAX if none)
  int32_t* k;       // insertion step that creates the cell (written by its parent)
  int32_t* count;   // particles in the cell
  int32_t* pidx;    // leaf: particle index of its first particle; else -1
  uint32_t* info;   // depth | octant << 8 | leaf << 11
  int32_t* child;   // [8] per internal cell, -1 where absent
  float4* box;      // centre, half size
  float4* com;      // centre of mass, total mass
  int32_t* skip;    // preorder position after the cell's subtree
  int32_t* sowner;  // internal cell straddling a block of positions: its block; else -1
  int32_t* pre;     // sequential builder: scratch
  int32_t* size;    // sequential builder: scratch
};

// Per-position arrays of the path-order build.
struct PathOrder {
  uint32_t* key;      // sort keys (in place)
  uint32_t* key_hi;   // digits 0..9 key of every rank
  int32_t* rank;      // ranks in path order (in place)
  uint4* spath;       // paths in path order
  uint8_t* lcp;       // n + 1 entries
  int32_t* first_pre; // n + 1 entries
  int32_t* bsum;      // per block of kPosBlock positions: cells, then their offsets
  int32_t* lvl;       // straddling cells by (hierarchy level, unit, depth), or -1
  int32_t* runs;      // starts of runs sharing 20 digits
  int2* lrun;         // runs longer than kRefineMax: (start, length)
  unsigned long long* lkey[2];  // path_refine_long: keys, ping-pong, by position
  int32_t* lval[2];             // path_refine_long: ranks, ping-pong, by position
  int8_t* lmin1;      // min lcp per chunk of 64 positions
  int8_t* lmin2;      // min lcp per block of kPosBlock positions
  int8_t* lmin3;      // min lcp per kSuperBlocks blocks
  int32_t* bstat;     // per block: deepest leaf, depth-32 multi leaves
  int32_t* ppart;     // particle index at every position: the traversal's order
};

constexpr int kPosBlock = 1024;  // positions per workgroup of the counting passes (16 chunks)
constexpr int kPosRounds = kPosBlock / kThreads;
// positions per block of the bottom-up pass: small enough that four workgroups'
// cells fit in a CU's LDS at once
constexpr int kLocalBlock = 512;
constexpr int kLocalRounds = kLocalBlock / kThreads;
constexpr int kSuperBlocks = 64;
constexpr int kLocalCap = 1024;  // cells of one block held in LDS by cell_local

// Cells that straddle blocks are reduced by a hierarchy of groups: level 2 takes
// those whose run stays inside one group of 64 blocks, level 3 inside one group
// of 64 level-2 groups, level 4 the rest. A cell's unit at level L is its owner
// block >> 6 (L - 2); at one depth a unit holds at most one such cell (the one
// containing the unit's last position). lvl[] files them by (L, unit, depth).
constexpr int kFan = 64;
constexpr int kMaxLevel = 4;
__host__ __device__ inline int lvl_units(int nb, int L) {
  int u = nb;
  for (int l = 2; l < L; ++l) u = (u + kFan - 1) / kFan;
  return u;
}
__host__ __device__ inline int lvl_base(int nb, int L) {
  int off = 0;
  for (int l = 2; l < L; ++l) off += lvl_units(nb, l) * (kMaxDepth + 1);
  return off;
}
// the level that reduces the root (it contains every position)
__host__ __device__ inline int root_level(int nb) { return nb <= kFan ? 2 : (nb <= kFan * kFan ? 3 : 4); }
constexpr uint32_t kLeafBit = 1u << 11;  // in TmpNodes::info
constexpr int kRefineMax = 1024;
constexpr int kSortDigits = 20;

// Leading octant digits shared by two paths (0..32).
__device__ __forceinline__ int path_lcp(uint4 a, uint4 b) {
  uint32_t x = a.x ^ b.x;
  if (x) return __builtin_ctz(x) / 3;
  x = a.y ^ b.y;
  if (x) return 10 + __builtin_ctz(x) / 3;
  x = a.z ^ b.z;
  if (x) return 20 + __builtin_ctz(x) / 3;
  x = a.w ^ b.w;
  if (x) return 30 + __builtin_ctz(x) / 3;
  return kMaxDepth;
}

__device__ __forceinline__ int path_digit(uint4 p, int d) {
  const uint32_t w = d < 10 ? p.x : (d < 20 ? p.y : (d < 30 ? p.z : p.w));
  return (int)((w >> (3 * (d % 10))) & 7u);
}

// Cells started by position j (see above); lcp_next = lcp[j + 1] (0 past the end).
__device__ __forceinline__ int leaf_depth(int l0, int l1) { return min(kMaxDepth, max(l0, l1) + 1); }

__global__ __launch_bounds__(kThreads) void path_gather_key(const int32_t* __restrict__ n_dev,
                                                            PathOrder O) {
  const int n = *n_dev;
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < n) O.key[i] = O.key_hi[O.rank[i]];
}

// lcp of every position with its predecessor, the paths in path order, and
// the starts of runs that share their first kSortDigits digits.
__global__ __launch_bounds__(kThreads) void path_lcp_kernel(BhState* st,
                                                            const uint4* __restrict__ path,
                                                            PathOrder O) {
  const int n = st->n;
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= n) {  // lcp[n] = 0, and zeros to the end of its chunk of 64
    if (j < ((n + 64) & ~63)) O.lcp[j] = 0;
    return;
  }
  const uint4 a = path[O.rank[j]];
  O.spath[j] = a;
  int l = 0;
  if (j > 0) {
    const uint4 b = path[O.rank[j - 1]];
    l = path_lcp(a, b);
    if (l >= kSortDigits && (j == 1 || path_lcp(b, path[O.rank[j - 2]]) < kSortDigits))
      O.runs[atomicAdd(&st->run_count, 1)] = j - 1;
  }
  O.lcp[j] = (uint8_t)l;
}

// Orders each run of positions sharing kSortDigits digits by its remaining
// digits, then rank (the order a full stable path sort gives), and redoes the
// run's lcp values. Bitonic sort in LDS of (digits 20..31 key << 25 | rank).
__global__ __launch_bounds__(kThreads) void path_refine(BhState* st,
                                                        const uint4* __restrict__ path,
                                                        PathOrder O) {
  __shared__ unsigned long long key[kRefineMax];
  __shared__ int sh_end;
  const int nr = st->run_count;
  const int n = st->n;
  for (int g = blockIdx.x; g < nr; g += gridDim.x) {
    const int s = O.runs[g];
    if (threadIdx.x == 0) sh_end = INT_MAX;
    __syncthreads();
    for (int base = s + 1;; base += kThreads) {
      const int x = base + (int)threadIdx.x;
      if (x >= n || O.lcp[x] < kSortDigits) atomicMin(&sh_end, min(x, n));
      __syncthreads();
      if (sh_end != INT_MAX) break;
      __syncthreads();
    }
    const int m = sh_end - s;
    if (m > kRefineMax) {  // left to path_refine_long
      if (threadIdx.x == 0) O.lrun[atomicAdd(&st->long_runs, 1)] = make_int2(s, m);
      __syncthreads();
      continue;
    }
    int p2 = 1;
    while (p2 < m) p2 <<= 1;
    for (int i = threadIdx.x; i < p2; i += kThreads) {
      unsigned long long k = ~0ull;
      if (i < m) {
        const int r = O.rank[s + i];
        const uint4 p = path[r];
        const uint32_t lo = ((~p.w & 7u) << 3) | ((~p.w >> 3) & 7u);
        k = ((((unsigned long long)order_key10(p.z) << 6) | lo) << 25) | (unsigned)r;
      }
      key[i] = k;
    }
    __syncthreads();
    for (int size = 2; size <= p2; size <<= 1)
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int i = threadIdx.x; i < p2; i += kThreads) {
          const int q = i ^ stride;
          if (q > i) {
            const bool up = (i & size) == 0;
            const unsigned long long u = key[i], v = key[q];
            if ((u > v) == up) {
              key[i] = v;
              key[q] = u;
            }
          }
        }
        __syncthreads();
      }
    for (int i = threadIdx.x; i < m; i += kThreads) {
      const int r = (int)(key[i] & 0x1FFFFFFull);
      O.rank[s + i] = r;
      O.spath[s + i] = path[r];
    }
    __syncthreads();
    for (int i = 1 + (int)threadIdx.x; i < m; i += kThreads)
      O.lcp[s + i] = (uint8_t)path_lcp(O.spath[s + i - 1], O.spath[s + i]);
    __syncthreads();
  }
}

// Digits 20..31 of a path as one 36-bit key, octants compared as 7 - o.
__device__ __forceinline__ unsigned long long low_digits_key(uint4 p) {
  const uint32_t lo = ((~p.w & 7u) << 3) | ((~p.w >> 3) & 7u);
  return ((unsigned long long)order_key10(p.z) << 6) | lo;
}

// Runs longer than kRefineMax, one workgroup per run: a stable LSD radix sort of
// the run by its remaining 12 digits (36 bits) in nine 4-bit counting passes
// through global scratch, each pass ranking kLongThreads positions at a time in
// LDS (a lane's rank among the earlier lanes of its digit from ballots, the
// waves' counts in LDS). The path sort left every run in rank order, so
// stability keeps equal paths in rank order: the order a full stable sort of
// the 32-digit paths gives. Then the run's ranks, paths and lcp values are
// rewritten as path_refine does. Costs one idle launch when no run is long.
constexpr int kLongThreads = 1024;
constexpr int kLongWaves = kLongThreads / 64;
__global__ __launch_bounds__(kLongThreads) void path_refine_long(BhState* st,
                                                                 const uint4* __restrict__ path,
                                                                 PathOrder O) {
  __shared__ int hist[16];
  __shared__ int base[16];
  __shared__ int wcnt[kLongWaves][16];
  const int nl = st->long_runs;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t below_me = (1ull << lane) - 1;
  for (int g = blockIdx.x; g < nl; g += gridDim.x) {
    const int s = O.lrun[g].x, m = O.lrun[g].y;
    unsigned long long* key[2] = {O.lkey[0] + s, O.lkey[1] + s};
    int32_t* val[2] = {O.lval[0] + s, O.lval[1] + s};
    for (int i = threadIdx.x; i < m; i += kLongThreads) {
      const int r = O.rank[s + i];
      key[0][i] = low_digits_key(path[r]);
      val[0][i] = r;
    }
    __syncthreads();
    for (int pass = 0; pass < 9; ++pass) {
      const int shift = 4 * pass;
      const unsigned long long* sk = key[pass & 1];
      const int32_t* sv = val[pass & 1];
      unsigned long long* dk = key[(pass + 1) & 1];
      int32_t* dv = val[(pass + 1) & 1];
      if (threadIdx.x < 16) hist[threadIdx.x] = 0;
      __syncthreads();
      for (int i = threadIdx.x; i < m; i += kLongThreads) atomicAdd(&hist[(sk[i] >> shift) & 15], 1);
      __syncthreads();
      if (threadIdx.x == 0) {
        int acc = 0;
        for (int b = 0; b < 16; ++b) {
          base[b] = acc;
          acc += hist[b];
        }
      }
      __syncthreads();
      for (int r0 = 0; r0 < m; r0 += kLongThreads) {
        const int i = r0 + (int)threadIdx.x;
        const bool valid = i < m;
        const unsigned long long k = valid ? sk[i] : 0ull;
        const int v = valid ? sv[i] : 0;
        const int d = (int)((k >> shift) & 15u);
        if (threadIdx.x < kLongWaves * 16) (&wcnt[0][0])[threadIdx.x] = 0;
        __syncthreads();
        uint64_t match = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const uint64_t set = __ballot((d >> b) & 1);
          match &= ((d >> b) & 1) ? set : ~set;
        }
        const int below = __popcll(match & below_me);
        if (valid && below == 0) wcnt[wave][d] = __popcll(match);
        __syncthreads();
        if (valid) {
          int off = base[d] + below;
          for (int w = 0; w < wave; ++w) off += wcnt[w][d];
          dk[off] = k;
          dv[off] = v;
        }
        __syncthreads();
        if (threadIdx.x < 16) {
          int t = 0;
          for (int w = 0; w < kLongWaves; ++w) t += wcnt[w][threadIdx.x];
          base[threadIdx.x] += t;
        }
        __syncthreads();
      }
    }
    // nine passes: the sorted run is in key[1] / val[1]
    for (int i = threadIdx.x; i < m; i += kLongThreads) {
      const int r = val[1][i];
      O.rank[s + i] = r;
      O.spath[s + i] = path[r];
    }
    __syncthreads();
    for (int i = 1 + (int)threadIdx.x; i < m; i += kLongThreads)
      O.lcp[s + i] = (uint8_t)path_lcp(O.spath[s + i - 1], O.spath[s + i]);
    __syncthreads();
  }
}

// Cells per block of positions; the deepest leaf; the depth-32 leaves that
// hold several particles; and the minima of lcp per chunk of 64 positions and
// per block, two levels of the index next_lower() searches.
__global__ __launch_bounds__(kThreads) void cell_count(BhState* st, PathOrder O) {
  __shared__ int sh_sum[kThreads / 64], sh_min[kThreads / 64];
  const int n = st->n;
  const int a = blockIdx.x * kPosBlock;
  int sum = 0, deep = 0, multi = 0, bmin = 127;
  for (int i = 0; i < kPosRounds; ++i) {
    const int j = a + i * kThreads + (int)threadIdx.x;
    int lv = 0;
    if (j < n) {
      const int l0 = O.lcp[j], l1 = O.lcp[j + 1];
      const int L = leaf_depth(l0, l1);
      sum += L - l0;
      deep = max(deep, L);
      multi += (l0 < kMaxDepth && l1 == kMaxDepth) ? 1 : 0;
      lv = l0;
    }
    int cm = lv;  // positions at and past n count as 0: every search stops at n
    for (int off = 32; off > 0; off >>= 1) cm = min(cm, __shfl_xor(cm, off, 64));
    if ((threadIdx.x & 63) == 0) O.lmin1[(a + i * kThreads + (int)threadIdx.x) >> 6] = (int8_t)cm;
    bmin = min(bmin, cm);
  }
  for (int off = 32; off > 0; off >>= 1) {
    sum += __shfl_xor(sum, off, 64);
    deep = max(deep, __shfl_xor(deep, off, 64));
    multi += __shfl_xor(multi, off, 64);
  }
  __shared__ int sh_deep[kThreads / 64], sh_multi[kThreads / 64];
  if ((threadIdx.x & 63) == 0) {
    sh_sum[threadIdx.x >> 6] = sum;
    sh_min[threadIdx.x >> 6] = bmin;
    sh_deep[threadIdx.x >> 6] = deep;
    sh_multi[threadIdx.x >> 6] = multi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0, m = 127, dp = 0, mu = 0;
    for (int w = 0; w < kThreads / 64; ++w) {
      t += sh_sum[w];
      m = min(m, sh_min[w]);
      dp = max(dp, sh_deep[w]);
      mu += sh_multi[w];
    }
    O.bsum[blockIdx.x] = t;
    O.lmin2[blockIdx.x] = (int8_t)m;
    O.bstat[3 * blockIdx.x] = dp;
    O.bstat[3 * blockIdx.x + 1] = mu;
  }
}

// One workgroup: block offsets, the cell count, the builder choice, the root,
// and the third level of the lcp index (minimum per kSuperBlocks blocks).
__global__ __launch_bounds__(1024) void cell_offsets(BhState* st, PathOrder O, TmpNodes T,
                                                     int max_nodes) {
  __shared__ int wsum[16];
  __shared__ int sh_carry, sh_deep, sh_multi;
  const int n = st->n;
  const int nb = n / kPosBlock + 1;  // blocks of cell_count (through position n)
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (threadIdx.x == 0) {
    sh_carry = 0;
    sh_deep = 0;
    sh_multi = 0;
  }
  __syncthreads();
  {
    int dp = 0, mu = 0;
    for (int k = threadIdx.x; k < nb; k += 1024) {
      dp = max(dp, O.bstat[3 * k]);
      mu += O.bstat[3 * k + 1];
    }
    if (dp) atomicMax(&sh_deep, dp);
    if (mu) atomicAdd(&sh_multi, mu);
  }
  for (int s = threadIdx.x; s * kSuperBlocks < nb; s += 1024) {
    int m = 127;
    for (int k = s * kSuperBlocks; k < min(nb, (s + 1) * kSuperBlocks); ++k) m = min(m, (int)O.lmin2[k]);
    O.lmin3[s] = (int8_t)m;
  }
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + (int)threadIdx.x;
    const int v = i < nb ? O.bsum[i] : 0;
    int incl = v;
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int before = sh_carry, tot = 0;
    for (int q = 0; q < 16; ++q) {
      if (q < w) before += wsum[q];
      tot += wsum[q];
    }
    if (i < nb) O.bsum[i] = before + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) sh_carry += tot;
    __syncthreads();
  }
  {
    const int nbl = (n + kLocalBlock - 1) / kLocalBlock;
    const int total = lvl_base(nbl, kMaxLevel + 1);
    for (int i = threadIdx.x; i < total; i += 1024) O.lvl[i] = -1;
    __syncthreads();
    if (threadIdx.x == 0) O.lvl[lvl_base(nbl, root_level(nbl))] = 0;  // unit 0, depth 0
    T.sowner[0] = root_level(nbl);  // owner block 0
  }
  if (threadIdx.x != 0) return;
  const int cells = 1 + sh_carry;
  const int seq = (cells + kMaxDepth + 1 >= max_nodes) ? 1 : 0;  // :449
  st->node_count = cells;
  st->final_count = cells;
  st->use_seq = seq;
  st->sort_n = seq ? 0 : cells;
  st->deepest = sh_deep;
  st->multi_count = sh_multi;
  st->max_depth = sh_multi > 0 ? kMaxDepth : sh_deep - 1;
  st->budget_merges = 0;
  st->depth_merges = 0;
  O.first_pre[n] = cells;
  if (seq) return;
  T.box[0] = make_float4(st->root_c[0], st->root_c[1], st->root_c[2], st->root_h);
  T.info[0] = 0;
  T.count[0] = n;
  T.pidx[0] = -1;
  T.skip[0] = cells;
}

// first_pre[j] for every position.
__global__ __launch_bounds__(kThreads) void cell_first_pre(const BhState* __restrict__ st,
                                                           PathOrder O) {
  if (st->use_seq) return;
  const int n = st->n;
  const int a = blockIdx.x * kPosBlock;
  int base = 1 + O.bsum[blockIdx.x];
  for (int i = 0; i < kPosRounds; ++i) {
    const int j = a + i * kThreads + (int)threadIdx.x;
    int c = 0;
    if (j < n) {
      const int l0 = O.lcp[j];
      c = leaf_depth(l0, O.lcp[j + 1]) - l0;
    }
    int tot;
    const int off = block_exclusive_scan(c, &tot);
    if (j < n) O.first_pre[j] = base + off;
    base += tot;
  }
}

// Offset of the first of the 16 bytes at p, from byte `from` on, whose value
// (signed when sgn) is below d; 16 if none. p is 16-byte aligned.
__device__ __forceinline__ int first_below16(const void* p, int from, int d, bool sgn) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int b = 0; b < 16; ++b) {
    const uint32_t u = (w[b >> 2] >> (8 * (b & 3))) & 0xFFu;
    const int val = sgn ? (int)(int8_t)u : (int)u;
    if (b >= from && val < d) return b;
  }
  return 16;
}

// First of the 64 bytes at p (64-aligned), from `from` on, below d; 64 if none.
__device__ __forceinline__ int first_below64(const void* p, int from, int d, bool sgn) {
  const uint8_t* q = static_cast<const uint8_t*>(p);
  int r[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) r[k] = first_below16(q + 16 * k, from - 16 * k, d, sgn);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (r[k] < 16) return 16 * k + r[k];
  return 64;
}

// First position x >= x0 with lcp[x] < d (d >= 1, x0 <= n): the end of the run
// of a depth-d cell. lcp[n] = 0, so the answer is at most n. The index levels:
// lcp per position (chunks of 64), lmin1 per chunk (16 per block), lmin2 per
// block (kSuperBlocks per super block), lmin3 per super block.
__device__ int next_lower(const PathOrder& O, int x0, int d, int n, int nsuper) {
  int c = x0 >> 6;
  int r = first_below64(O.lcp + (c << 6), x0 & 63, d, false);
  if (r < 64) return (c << 6) + r;
  int k = c >> 4;
  r = first_below16(O.lmin1 + (k << 4), (c & 15) + 1, d, true);
  if (r >= 16) {
    int s = k / kSuperBlocks;
    r = first_below64(O.lmin2 + s * kSuperBlocks, k % kSuperBlocks + 1, d, true);
    if (r >= 64) {
      for (++s; s < nsuper; s += 64) {  // super blocks, 64 per look (never past the end)
        const int f = first_below64(O.lmin3 + (s & ~63), s & 63, d, true);
        if (f < 64) {
          s = (s & ~63) + f;
          break;
        }
        s = (s & ~63);
      }
      if (s >= nsuper) return n;  // not reached: lcp[n] = 0
      r = first_below64(O.lmin2 + s * kSuperBlocks, 0, d, true);
    }
    k = s * kSuperBlocks + r;
    r = first_below16(O.lmin1 + (k << 4), 0, d, true);
  }
  c = (k << 4) + r;
  return (c << 6) + first_below64(O.lcp + (c << 6), 0, d, false);
}

// Every cell a position starts: box (the reference's centre chain, :375-379),
// depth, octant, particle count, subtree end; leaves completely (their
// particle, or the depth-32 merge of :355-365 and :490-507 in insertion order)
// with their traversal record.
__global__ __launch_bounds__(kThreads) void cell_emit(const BhState* __restrict__ st, PathOrder O,
                                                      TmpNodes T, const int32_t* __restrict__ sidx,
                                                      const float4* __restrict__ sbody,
                                                      float4* rec, TravParams tp) {
  if (st->use_seq) return;
  const int n = st->n;
  const int j = blockIdx.x * kThreads + threadIdx.x;
  if (j >= n) return;
  const int r = O.rank[j];
  const int pid = sidx[r];
  O.ppart[j] = pid;
  const int l0 = O.lcp[j], l1 = O.lcp[j + 1];
  const int L = leaf_depth(l0, l1);
  if (L <= l0) return;  // not the first particle of any cell (depth-32 duplicate)
  const int pre0 = O.first_pre[j] - l0 - 1;  // pre(j, d) = pre0 + d
  const uint4 pj = O.spath[j];
  float4 q = sbody[r];  // the leaf's particle: load issued before the chain
  float cx = st->root_c[0], cy = st->root_c[1], cz = st->root_c[2], h = st->root_h;
  for (int d = 1; d <= L; ++d) {
    const int o = path_digit(pj, d - 1);
    const float hs = mul_rn(h, 0.5f);  // :375
    cx = add_rn(cx, (o & 1) ? hs : -hs);
    cy = add_rn(cy, (o & 2) ? hs : -hs);
    cz = add_rn(cz, (o & 4) ? hs : -hs);
    h = hs;
    if (d <= l0) continue;
    const int p = pre0 + d;
    T.box[p] = make_float4(cx, cy, cz, h);
    T.info[p] = (uint32_t)d | ((uint32_t)o << 8) | (d == L ? kLeafBit : 0u);
    if (d < L) T.pidx[p] = -1;
  }
  {  // the leaf (j, L): its particle, or the depth-32 merge in insertion order
    const int p = pre0 + L;
    int e = j + 1;
    int t2 = INT_MAX;
    if (l1 == kMaxDepth) {  // depth-32 cell with several particles, ranks ascending
      t2 = O.rank[j + 1];
      while (e < n && O.lcp[e] == kMaxDepth) {
        const float4 u = sbody[O.rank[e]];
        const float tot = add_rn(q.w, u.w);
        q.x = div_rn(add_rn(mul_rn(q.x, q.w), mul_rn(u.x, u.w)), tot);
        q.y = div_rn(add_rn(mul_rn(q.y, q.w), mul_rn(u.y, u.w)), tot);
        q.z = div_rn(add_rn(mul_rn(q.z, q.w), mul_rn(u.z, u.w)), tot);
        q.w = tot;
        ++e;
      }
    }
    T.com[p] = q;
    T.pidx[p] = pid;
    T.t1[p] = r;
    T.t2[p] = t2;
    T.count[p] = e - j;
    T.skip[p] = p + 1;
    T.sowner[p] = -1;  // leaves are done here, never by cell_group
    write_trav(rec, p, 1, q, make_float4(cx, cy, cz, h), true, tp);
  }
  // internal cells: their first position, for cell_ends
  for (int d = l0 + 1; d < L; ++d) T.skip[pre0 + d] = j;
}

// End of every internal cell's run, one thread per cell so that the searches
// run side by side: the run of (j, d) ends at the next position with fewer
// than d shared digits. Gives the subtree end (skip) and the particle count.
__global__ __launch_bounds__(kThreads) void cell_ends(const BhState* __restrict__ st, PathOrder O,
                                                      TmpNodes T) {
  if (st->use_seq) return;
  const int n = st->n, nn = st->final_count;
  const int nsuper = (n / kPosBlock + 1 + kSuperBlocks - 1) / kSuperBlocks;
  for (int p = 1 + blockIdx.x * kThreads + threadIdx.x; p < nn; p += gridDim.x * kThreads) {
    const uint32_t u = T.info[p];
    if (u & kLeafBit) continue;
    const int j = T.skip[p];  // the first position, written by cell_emit
    const int e = next_lower(O, j + 1, (int)(u & 0xFFu), n, nsuper);
    T.count[p] = e - j;
    T.skip[p] = O.first_pre[e];
    const int kb = j / kLocalBlock;
    if (kb == (e - 1) / kLocalBlock) {
      T.sowner[p] = -1;  // inside its block: cell_local reduces it
      continue;
    }
    int L = 2;
    while (L < kMaxLevel &&
           (long long)e > (((long long)(kb >> (6 * (L - 1))) + 1) << (6 * (L - 1))) * kLocalBlock)
      ++L;
    const int nbl = (n + kLocalBlock - 1) / kLocalBlock;
    T.sowner[p] = kb * 8 + L;
    O.lvl[lvl_base(nbl, L) + (kb >> (6 * (L - 2))) * (kMaxDepth + 1) + (int)(u & 0xFFu)] = p;
  }
}

// Children of every internal cell: first child p + 1, each next child at the
// previous one's skip (children 7..0), stored by octant.
__global__ __launch_bounds__(kThreads) void cell_link(const BhState* __restrict__ st, TmpNodes T) {
  if (st->use_seq) return;
  const int nn = st->final_count;
  for (int p = blockIdx.x * kThreads + threadIdx.x; p < nn; p += gridDim.x * kThreads) {
    if (T.info[p] & kLeafBit) continue;
    int c[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    const int end = T.skip[p];
    // at most 8 children, and skips only grow: the bounds keep the walk finite
    for (int q = p + 1, it = 0; q < end && it < 8; ++it) {
      c[(T.info[q] >> 8) & 7u] = q;
      const int next = T.skip[q];
      q = next > q ? next : end;
    }
    int4* out = reinterpret_cast<int4*>(T.child + (size_t)p * 8);
    out[0] = make_int4(c[0], c[1], c[2], c[3]);
    out[1] = make_int4(c[4], c[5], c[6], c[7]);
  }
}

constexpr uint32_t kSpanBit = 1u << 12;  // cell_local only: the cell straddles the block's end

// Children of local cell i (index within the block's cells) in visiting order
// (octants descending): i + 1, then each previous child's skip, while the depth
// is d + 1. Returns their number.
__device__ __forceinline__ int local_children(int i, int d, int ncell, int base,
                                              const uint32_t* info, const int* skip, int ch[8]) {
  int nc = 0;
  for (int q = i + 1; nc < 8 && q < ncell && (int)(info[q] & 0xFFu) == d + 1;) {
    ch[nc++] = q;
    const int next = skip[q] - base;
    q = next > q ? next : ncell;
  }
  return nc;
}

// Local cell ci at depth d from its children (visiting order, so summed in
// reverse: octants ascending, :514-538).
__device__ __forceinline__ void local_reduce(int ci, int d, int ncell, int base,
                                             const uint32_t* info, const int* skip, float4* com,
                                             int* t1, int* t2) {
  int ch[8];
  const int nc = local_children(ci, d, ncell, base, info, skip, ch);
  float tm = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
  int r1 = INT_MAX, r2 = INT_MAX;
  for (int q = nc - 1; q >= 0; --q) {
    const float4 u = com[ch[q]];
    tm = add_rn(tm, u.w);
    sx = add_rn(sx, mul_rn(u.x, u.w));
    sy = add_rn(sy, mul_rn(u.y, u.w));
    sz = add_rn(sz, mul_rn(u.z, u.w));
    const int ra = t1[ch[q]], rb = t2[ch[q]];
    if (ra < r1) {
      r2 = min(r1, rb);
      r1 = ra;
    } else if (ra < r2) {
      r2 = ra;
    }
  }
  float4 res = make_float4(0.f, 0.f, 0.f, tm);
  if (tm > 0.0f) {
    res.x = div_rn(sx, tm);
    res.y = div_rn(sy, tm);
    res.z = div_rn(sz, tm);
  }
  com[ci] = res;
  t1[ci] = r1;
  t2[ci] = r2;
}

// Bottom-up over the internal cells whose runs lie inside this block of
// positions, deepest level first, with the block's cells in LDS (in global
// memory when there are more than kLocalCap). The cells that contain the
// block's last position and the next one (at most one per depth; the root is
// always one) are left to cell_group.
__global__ __launch_bounds__(kThreads) void cell_local(BhState* st, PathOrder O, TmpNodes T,
                                                       float4* rec, TravParams tp) {
  __shared__ float4 s_com[kLocalCap];
  __shared__ int s_t1[kLocalCap], s_t2[kLocalCap], s_skip[kLocalCap];
  __shared__ uint32_t s_info[kLocalCap];
  __shared__ int8_t smin[kLocalBlock + 1];  // smin[x - a] = min(lcp[x..b]), -1 at n
  __shared__ int seg[kThreads / 64];  // per-wave minima
  __shared__ uint16_t lvl_list[kLocalCap];  // local internal cells grouped by depth
  __shared__ int lvl_cnt[kMaxDepth + 1], lvl_start[kMaxDepth + 1], lvl_fill[kMaxDepth + 1];
  __shared__ int sh_top;
  __shared__ int span[kMaxDepth + 1];  // straddling cells of this block, to unmark (fallback path)
  if (st->use_seq) return;
  const int n = st->n;
  const int a = blockIdx.x * kLocalBlock;
  const int b = min(a + kLocalBlock, n);  // the block's positions are [a, b)
  const int P0 = O.first_pre[a];
  const int ncell = O.first_pre[b] - P0;
  const bool fits = ncell <= kLocalCap;
  float4* com = fits ? s_com : T.com + P0;
  int* t1 = fits ? s_t1 : T.t1 + P0;
  int* t2 = fits ? s_t2 : T.t2 + P0;
  int* skip = fits ? s_skip : T.skip + P0;
  uint32_t* info = fits ? s_info : T.info + P0;
  if (threadIdx.x <= kMaxDepth) span[threadIdx.x] = -1;
  if (threadIdx.x == 0) sh_top = 0;
  // the boxes of this thread's cells, loaded now and used at the end (written by
  // cell_emit; a load in the final loop would wait a full memory latency per cell)
  constexpr int kBoxRounds = kLocalCap / kThreads;
  float4 box_r[kBoxRounds];
  if (fits) {
#pragma unroll
    for (int k = 0; k < kBoxRounds; ++k) {
      const int i = threadIdx.x + k * kThreads;
      box_r[k] = i < ncell ? T.box[P0 + i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    for (int i = threadIdx.x; i < ncell; i += kThreads) {
      const int p = P0 + i;
      const uint32_t u = T.info[p];
      s_info[i] = u;
      s_skip[i] = T.skip[p];
      if (u & kLeafBit) {
        s_com[i] = T.com[p];
        s_t1[i] = T.t1[p];
        s_t2[i] = T.t2[p];
      }
    }
  }
  // suffix minima of lcp over (a, b], with lcp[b] = -1 when b == n
  {
    const int t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    constexpr int kSeg = kLocalBlock / kThreads;
    int v[kSeg];
    int m = 127;
    for (int i = kSeg - 1; i >= 0; --i) {
      const int x = a + 1 + t * kSeg + i;
      int l = 127;
      if (x <= b) l = x < n ? (int)O.lcp[x] : -1;
      m = min(m, l);
      v[i] = m;
    }
    // minimum over this lane's segment and the later lanes' of the wave
    int sfx = m;
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_down(sfx, off, 64);
      if (lane + off < 64) sfx = min(sfx, o);
    }
    if (lane == 0) seg[w] = sfx;
    int after = __shfl_down(sfx, 1, 64);
    if (lane == 63) after = 127;
    __syncthreads();
    for (int q = w + 1; q < kThreads / 64; ++q) after = min(after, seg[q]);
    for (int i = 0; i < kSeg; ++i) smin[1 + t * kSeg + i] = (int8_t)min(v[i], after);
  }
  __syncthreads();
  int l0v[kLocalRounds], Lv[kLocalRounds], pv[kLocalRounds], mv[kLocalRounds];
  int top = 0;
  for (int i = 0; i < kLocalRounds; ++i) {
    const int j = a + i * kThreads + (int)threadIdx.x;
    l0v[i] = 0;
    Lv[i] = 0;
    pv[i] = 0;
    mv[i] = -1;
    if (j >= b) continue;
    const int l0 = O.lcp[j];
    const int L = leaf_depth(l0, O.lcp[j + 1]);
    l0v[i] = l0;
    Lv[i] = L;
    pv[i] = O.first_pre[j] - l0 - 1 - P0;  // local index of cell (j, d) = pv + d
    mv[i] = smin[j + 1 - a];
    top = max(top, L - 1);
    for (int d = l0 + 1; d < L && d <= mv[i]; ++d) {  // internal cells containing position b
      span[d] = P0 + pv[i] + d;
      info[pv[i] + d] |= kSpanBit;
    }
  }
  if (top) atomicMax(&sh_top, top);
  if (threadIdx.x <= kMaxDepth) {
    lvl_cnt[threadIdx.x] = 0;
    lvl_fill[threadIdx.x] = 0;
  }
  __syncthreads();
  if (fits) {
    // the local internal cells by depth, then one parallel round per level
    for (int i = threadIdx.x; i < ncell; i += kThreads) {
      const uint32_t u = s_info[i];
      if (!(u & (kLeafBit | kSpanBit))) atomicAdd(&lvl_cnt[u & 0xFFu], 1);
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of the depth counts (kMaxDepth < 64)
      const int lane = threadIdx.x;
      const int c = lane <= kMaxDepth ? lvl_cnt[lane] : 0;
      int incl = c;
      for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
      }
      if (lane <= kMaxDepth) lvl_start[lane] = incl - c;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < ncell; i += kThreads) {
      const uint32_t u = s_info[i];
      if (!(u & (kLeafBit | kSpanBit))) {
        const int d = (int)(u & 0xFFu);
        lvl_list[lvl_start[d] + atomicAdd(&lvl_fill[d], 1)] = (uint16_t)i;
      }
    }
    __syncthreads();
    for (int d = sh_top; d >= 1; --d) {
      if (lvl_cnt[d] == 0) continue;  // the same for every thread: no barrier needed
      for (int q = threadIdx.x; q < lvl_cnt[d]; q += kThreads)
        local_reduce(lvl_list[lvl_start[d] + q], d, ncell, P0, s_info, s_skip, s_com, s_t1, s_t2);
      __syncthreads();
    }
  } else {
    for (int d = sh_top; d >= 1; --d) {
      for (int i = 0; i < kLocalRounds; ++i)
        if (l0v[i] < d && d < Lv[i] && mv[i] < d)
          local_reduce(pv[i] + d, d, ncell, P0, info, skip, com, t1, t2);
      __syncthreads();
    }
  }
  // results of the block's internal cells (not the straddling ones) to global
  // memory: centre of mass, ranks, child table, the children's creation steps,
  // the traversal record
  auto finish = [&](int i, float4 box) {
    const uint32_t u = info[i];
    if (u & (kLeafBit | kSpanBit)) return;
    const int p = P0 + i;
    const float4 res = com[i];
    const int r2 = t2[i];
    if (fits) {
      T.com[p] = res;
      T.t1[p] = t1[i];
      T.t2[p] = r2;
    }
    int ch[8];
    const int nc = local_children(i, (int)(u & 0xFFu), ncell, P0, info, skip, ch);
    for (int q = 0; q < nc; ++q) T.k[P0 + ch[q]] = max(r2, t1[ch[q]]);
    write_trav(rec, p, skip[i] - p, res, box, false, tp);
  };
  if (fits) {
#pragma unroll
    for (int k = 0; k < kBoxRounds; ++k) {
      const int i = threadIdx.x + k * kThreads;
      if (i < ncell) finish(i, box_r[k]);
    }
  } else {
    for (int i = threadIdx.x; i < ncell; i += kThreads) finish(i, T.box[P0 + i]);
  }
  if (!fits) {  // the span marks went to the global info array: clear them
    __syncthreads();
    for (int d = 1; d <= kMaxDepth; ++d)
      if (threadIdx.x == 0 && span[d] >= 0) T.info[span[d]] &= ~kSpanBit;
  }
}

// Level L of the hierarchy: one workgroup per group of kFan units reduces the
// group's straddling cells, deepest first. The group's cells and their child
// tables are loaded once into LDS; each depth's results stay in LDS, where a
// child reduced at the same level is found by its unit (T.sowner); every other
// child was finished before (cell_emit, cell_local, or a lower level).
__global__ __launch_bounds__(kThreads) void cell_group(const BhState* __restrict__ st, PathOrder O,
                                                       TmpNodes T, float4* rec, TravParams tp,
                                                       int L) {
  __shared__ int cells[kMaxDepth + 1][kFan];
  __shared__ int4 tab[kMaxDepth + 1][kFan][2];
  __shared__ float4 r_com[2][kFan];
  __shared__ int r_t1[2][kFan], r_t2[2][kFan];
  __shared__ int sh_dmax;
  if (st->use_seq) return;
  const int nb = (st->n + kLocalBlock - 1) / kLocalBlock;
  const int units = lvl_units(nb, L);
  const int u0 = blockIdx.x * kFan;
  const int nu = min(kFan, units - u0);
  const int* lv = O.lvl + lvl_base(nb, L) + (size_t)u0 * (kMaxDepth + 1);
  if (threadIdx.x == 0) sh_dmax = -1;
  __syncthreads();
  int dmax = -1;
  for (int idx = threadIdx.x; idx < kFan * (kMaxDepth + 1); idx += kThreads) {
    const int u = idx / (kMaxDepth + 1), d = idx % (kMaxDepth + 1);
    const int p = u < nu ? lv[idx] : -1;
    cells[d][u] = p;
    if (p >= 0) {
      const int4* cp = reinterpret_cast<const int4*>(T.child + (size_t)p * 8);
      tab[d][u][0] = cp[0];
      tab[d][u][1] = cp[1];
      dmax = max(dmax, d);
    }
  }
  if (dmax >= 0) atomicMax(&sh_dmax, dmax);
  __syncthreads();
  const int shift = 6 * (L - 2);  // owner block -> unit at this level
  for (int d = sh_dmax; d >= 0; --d) {
    const int cur = d & 1;
    const int u = threadIdx.x;
    const int p = u < kFan ? cells[d][u] : -1;
    if (p >= 0) {
      const int4 c0 = tab[d][u][0], c1 = tab[d][u][1];
      const int c[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
      float4 q[8];
      int a[8], b[8], own[8];
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        q[o] = make_float4(0.f, 0.f, 0.f, 0.f);
        a[o] = b[o] = INT_MAX;
        own[o] = -1;
        if (c[o] >= 0) {
          own[o] = T.sowner[c[o]];
          q[o] = T.com[c[o]];
          a[o] = T.t1[c[o]];
          b[o] = T.t2[c[o]];
        }
      }
      const float4 box = T.box[p];
      const int skip = T.skip[p];
      float tm = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
      int r1 = INT_MAX, r2 = INT_MAX;
#pragma unroll
      for (int o = 0; o < 8; ++o) {  // octants ascending (:514-538)
        if (c[o] < 0) continue;
        if (own[o] >= 0 && (own[o] & 7) == L) {  // reduced one depth below, here
          const int cu = ((own[o] >> 3) >> shift) - u0;
          q[o] = r_com[cur ^ 1][cu];
          a[o] = r_t1[cur ^ 1][cu];
          b[o] = r_t2[cur ^ 1][cu];
        }
        tm = add_rn(tm, q[o].w);
        sx = add_rn(sx, mul_rn(q[o].x, q[o].w));
        sy = add_rn(sy, mul_rn(q[o].y, q[o].w));
        sz = add_rn(sz, mul_rn(q[o].z, q[o].w));
        if (a[o] < r1) {
          r2 = min(r1, b[o]);
          r1 = a[o];
        } else if (a[o] < r2) {
          r2 = a[o];
        }
      }
      float4 com = make_float4(0.f, 0.f, 0.f, tm);
      if (tm > 0.0f) {
        com.x = div_rn(sx, tm);
        com.y = div_rn(sy, tm);
        com.z = div_rn(sz, tm);
      }
      r_com[cur][u] = com;
      r_t1[cur][u] = r1;
      r_t2[cur][u] = r2;
      T.com[p] = com;
      T.t1[p] = r1;
      T.t2[p] = r2;
#pragma unroll
      for (int o = 0; o < 8; ++o)
        if (c[o] >= 0) T.k[c[o]] = p == 0 ? a[o] : max(r2, a[o]);  // root's children: own t1
      write_trav(rec, p, skip - p, com, box, false, tp);
    }
    __syncthreads();
  }
}

// Reference node-numbering key of every cell (root first).
__global__ __launch_bounds__(kThreads) void node_keys(const BhState* __restrict__ st, TmpNodes T,
                                                      uint32_t* keys, int32_t* vals) {
  const int nn = st->sort_n;
  for (int t = blockIdx.x * kThreads + threadIdx.x; t < nn; t += gridDim.x * kThreads) {
    uint32_t key = 0;
    if (t > 0) {
      const uint32_t depth = T.info[t] & 0xFFu;
      const uint32_t flag = T.t1[t] == T.k[t] ? 1u : 0u;
      key = ((uint32_t)(T.k[t] + 1) << 7) | (flag << 6) | depth;
    }
    keys[t] = key;
    vals[t] = t;
  }
}

__global__ __launch_bounds__(kThreads) void invert_perm(const BhState* __restrict__ st,
                                                        const int32_t* __restrict__ order,
                                                        int32_t* final_of) {
  const int nn = st->sort_n;
  for (int f = blockIdx.x * kThreads + threadIdx.x; f < nn; f += gridDim.x * kThreads)
    final_of[order[f]] = f;
}

__global__ __launch_bounds__(kThreads) void emit_nodes(const BhState* __restrict__ st, TmpNodes T,
                                                       const int32_t* __restrict__ final_of,
                                                       RefNode* out) {
  const int nn = st->sort_n;
  for (int t = blockIdx.x * kThreads + threadIdx.x; t < nn; t += gridDim.x * kThreads) {
    RefNode nd;
    const float4 b = T.box[t], c = T.com[t];
    nd.cx = b.x;
    nd.cy = b.y;
    nd.cz = b.z;
    nd.half = b.w;
    nd.comx = c.x;
    nd.comy = c.y;
    nd.comz = c.z;
    nd.mass = c.w;
    const bool leaf = (T.info[t] >> 11) & 1u;
    for (int o = 0; o < 8; ++o) {
      const int ch = leaf ? -1 : T.child[t * 8 + o];
      nd.child[o] = ch >= 0 ? final_of[ch] : -1;
    }
    nd.pidx = leaf ? T.pidx[t] : -1;
    nd.leaf = leaf ? 1 : 0;
    nd.pad[0] = nd.pad[1] = nd.pad[2] = 0;
    nd.count = T.count[t];
    out[final_of[t]] = nd;
  }
}

// ------------------------------------------------- exact sequential builder

__device__ void seq_default(RefNode* nd) {
  nd->cx = nd->cy = nd->cz = nd->half = 0.f;
  nd->comx = nd->comy = nd->comz = nd->mass = 0.f;
  for (int o = 0; o < 8; ++o) nd->child[o] = -1;
  nd->pidx = -1;
  nd->leaf = 1;
  nd->pad[0] = nd->pad[1] = nd->pad[2] = 0;
  nd->count = 0;
}

__device__ void seq_merge(RefNode* nd, int pid, float px, float py, float pz, float pm) {
  const float tot = add_rn(nd->mass, pm);  // :355-365
  const float cx = div_rn(add_rn(mul_rn(nd->comx, nd->mass), mul_rn(px, pm)), tot);
  const float cy = div_rn(add_rn(mul_rn(nd->comy, nd->mass), mul_rn(py, pm)), tot);
  const float cz = div_rn(add_rn(mul_rn(nd->comz, nd->mass), mul_rn(pz, pm)), tot);
  nd->comx = cx;
  nd->comy = cy;
  nd->comz = cz;
  nd->mass = tot;
  nd->count += 1;
  if (nd->pidx < 0) nd->pidx = pid;
}

__device__ int seq_leaf(RefNode* nodes, int* count, int max_nodes, int parent, int o, int pid,
                        float px, float py, float pz, float pm) {  // :368-389
  if (*count >= max_nodes) return -1;
  const int id = (*count)++;
  RefNode* p = &nodes[parent];
  p->child[o] = id;
  p->leaf = 0;
  const float hs = mul_rn(p->half, 0.5f);
  RefNode* c = &nodes[id];
  c->cx = add_rn(p->cx, (o & 1) ? hs : -hs);
  c->cy = add_rn(p->cy, (o & 2) ? hs : -hs);
  c->cz = add_rn(p->cz, (o & 4) ? hs : -hs);
  c->half = hs;
  c->leaf = 1;
  c->pad[0] = c->pad[1] = c->pad[2] = 0;
  c->pidx = pid;
  c->count = 1;
  c->mass = pm;
  c->comx = px;
  c->comy = py;
  c->comz = pz;
  for (int q = 0; q < 8; ++q) c->child[q] = -1;
  return id;
}

// One thread replays force_barnes_hut.cu:391-538 (only when the budget may fire).
__global__ void seq_build(BhState* st, const int32_t* __restrict__ sidx, const float* __restrict__ x,
                          const float* __restrict__ y, const float* __restrict__ z,
                          const float* __restrict__ m, RefNode* nodes, int max_nodes) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || !st->use_seq) return;
  const int n = st->n;
  RefNode* root = &nodes[0];
  seq_default(root);
  root->cx = st->root_c[0];
  root->cy = st->root_c[1];
  root->cz = st->root_c[2];
  root->half = st->root_h;
  root->leaf = 0;
  root->count = n;
  int count = 1, deepest = 0, nb = 0, nd = 0;
  for (int r = 0; r < n; ++r) {
    const int pid = sidx[r];
    const float px = x[pid], py = y[pid], pz = z[pid], pm = m[pid];
    int cur = 0, depth = 0;
    bool merged = false, overflow = false;
    while (depth < kMaxDepth) {
      RefNode* q = &nodes[cur];
      if (q->leaf) {
        if (count + kMaxDepth + 1 >= max_nodes) {
          seq_merge(q, pid, px, py, pz, pm);
          merged = true;
          ++nb;
          break;
        }
        const int old = q->pidx;
        const float ox = q->comx, oy = q->comy, oz = q->comz, om = q->mass;
        q->leaf = 0;
        q->pidx = -1;
        q->count = 0;
        q->mass = 0.f;
        q->comx = q->comy = q->comz = 0.f;
        const int oo = octant_of(ox, oy, oz, q->cx, q->cy, q->cz);
        if (q->child[oo] < 0 && seq_leaf(nodes, &count, max_nodes, cur, oo, old, ox, oy, oz, om) < 0) {
          overflow = true;
          break;
        }
      }
      const int o = octant_of(px, py, pz, nodes[cur].cx, nodes[cur].cy, nodes[cur].cz);
      const int ch = nodes[cur].child[o];
      if (ch < 0) {
        if (seq_leaf(nodes, &count, max_nodes, cur, o, pid, px, py, pz, pm) < 0) overflow = true;
        break;
      }
      cur = ch;
      ++depth;
    }
    if (overflow) {  // the reference writes past its node array here
      st->truncated = n - r;
      break;
    }
    if (!merged && depth >= kMaxDepth) {
      int leaf = cur;
      for (int guard = 0; guard <= kMaxDepth && !nodes[leaf].leaf; ++guard) {
        int next = leaf;
        for (int c = 0; c < 8; ++c)
          if (nodes[leaf].child[c] >= 0) {
            next = nodes[leaf].child[c];
            break;
          }
        if (next == leaf) break;
        leaf = next;
      }
      seq_merge(&nodes[leaf], pid, px, py, pz, pm);
      ++nd;
    }
    if (depth > deepest) deepest = depth;
  }
  for (int i = count - 1; i >= 0; --i) {
    RefNode* p = &nodes[i];
    if (p->leaf) continue;
    float tm = 0.f, sx = 0.f, sy = 0.f, sz = 0.f;
    int cnt = 0;
    for (int c = 0; c < 8; ++c) {
      const int ci = p->child[c];
      if (ci < 0) continue;
      const RefNode* q = &nodes[ci];
      tm = add_rn(tm, q->mass);
      sx = add_rn(sx, mul_rn(q->comx, q->mass));
      sy = add_rn(sy, mul_rn(q->comy, q->mass));
      sz = add_rn(sz, mul_rn(q->comz, q->mass));
      cnt += q->count;
    }
    p->mass = tm;
    p->count = cnt;
    if (tm > 0.f) {
      p->comx = div_rn(sx, tm);
      p->comy = div_rn(sy, tm);
      p->comz = div_rn(sz, tm);
    }
  }
  st->final_count = count;
  st->max_depth = deepest;
  st->budget_merges = nb;
  st->depth_merges = nd;
}

// Traversal layout when the sequential builder produced the nodes (use_seq):
// one thread walks them as the reference does (:160-241) and numbers them in
// visiting order, then sizes the subtrees in reverse order (a node's children
// always come after it). pre_of, size and order are scratch of max_nodes ints.
__global__ void seq_preorder(const BhState* __restrict__ st, const RefNode* __restrict__ nodes,
                             int32_t* pre_of, int32_t* size, int32_t* order, float4* rec,
                             TravParams tp) {
  if (!st->use_seq || threadIdx.x != 0 || blockIdx.x != 0) return;
  const int nn = st->final_count;
  if (nn <= 0) return;
  int stack[8 * (kMaxDepth + 2)];
  int sp = 0, cnt = 0;
  stack[sp++] = 0;
  while (sp > 0) {
    const int v = stack[--sp];
    pre_of[v] = cnt;
    order[cnt++] = v;
    if (nodes[v].leaf) continue;
    for (int o = 0; o < 8; ++o) {
      const int c = nodes[v].child[o];
      if (c >= 0 && sp < 8 * (kMaxDepth + 2)) stack[sp++] = c;
    }
  }
  for (int p = cnt - 1; p >= 0; --p) {
    const int v = order[p];
    const RefNode& q = nodes[v];
    int sz = 1;
    if (!q.leaf)
      for (int o = 0; o < 8; ++o)
        if (q.child[o] >= 0) sz += size[q.child[o]];
    size[v] = sz;
    write_trav(rec, p, sz, make_float4(q.comx, q.comy, q.comz, q.mass),
               make_float4(q.cx, q.cy, q.cz, q.half), q.leaf, tp);
  }
}

// ---------------------------------------------------------------- 6b: task order

// The order traversal tasks take the particles in (64 consecutive per task):
// the tree's own order (path order; Morton order after the sequential
// builder), sorted stably by the Hilbert index of each particle's Morton cell
// (the same 1024^3 grid as its Morton code; morton_kernel computes both). Morton and path order jump
// between distant cells at octant boundaries, so some 64-particle tasks span
// two far-apart groups and their walk covers both groups' trees; consecutive
// Hilbert cells are always adjacent, so tasks stay compact (7 % fewer walk
// steps on the two-Plummer million). Any order gives the same forces: each
// lane sums its interactions in the records' order whatever task it is in.
__global__ __launch_bounds__(kThreads) void task_keys(const BhState* __restrict__ st,
                                                      const uint32_t* __restrict__ hkey,
                                                      const int32_t* __restrict__ morton_order,
                                                      const int32_t* __restrict__ path_order,
                                                      uint32_t* key, int32_t* order) {
  const int r = blockIdx.x * kThreads + threadIdx.x;
  if (r >= st->n) return;
  const int p = st->use_seq ? morton_order[r] : path_order[r];
  key[r] = hkey[p];
  order[r] = p;
}

// ---------------------------------------------------------------- 7: traversal

// Per force call: each record's opening thresholds for this theta and its
// G * mass. rec[4p+2].zw = (S_open, S_acc) with S = size2 * (1 / theta2),
// S_open = S (1 - 2^-18) and S_acc = S (1 + 2^-18), each product rounded once;
// rec[4p].w = G * mass rounded once, as the traversal's force term rounds it.
// S_open and S_acc are within 2^-22 (relative) of size2 / theta2 (1 -+ 2^-18).
// The walk's common step tests them against a fused
// d2f = fma(dz, dz, fma(dy, dy, fma(dx, dx, eps2))), three roundings of
// nonnegative terms against the four of the reference's
// d2 = ((dx dx + dy dy) + dz dz) + eps2 (:221), so d2 <= d2f (1 + 7 2^-24)
// and d2 >= d2f (1 - 7 2^-24). Hence d2f < S_open implies d2 < S (1 - 2^-19),
// i.e. size2 / d2 > theta2 (1 + 2^-20), and d2f > S_acc implies
// size2 / d2 < theta2 (1 - 2^-20): either way the rounded division takes the
// same decision (:224). In between, the walk computes the reference's d2
// and divides. Leaves and zero-mass cells (size2 = -1) get negative thresholds
// and are never opened by the test; theta = 0 gives S = +inf, so every cell
// with mass is opened.
__global__ __launch_bounds__(kThreads) void trav_prep(float4* rec, const BhState* __restrict__ st,
                                                      float inv_theta2, float G) {
  const int nn = st->final_count;
  for (int p = blockIdx.x * kThreads + threadIdx.x; p < nn; p += gridDim.x * kThreads) {
    float gm, s_open, s_acc;
    trav_fields(rec[4 * p + 2].y, rec[4 * p + 3].x, inv_theta2, G, &gm, &s_open, &s_acc);
    rec[4 * p].w = gm;
    rec[4 * p + 2].z = s_open;
    rec[4 * p + 2].w = s_acc;
  }
}

// Wave-coherent traversal over the preorder layout. A wave holds 64
// consecutive particles of the task order (task_keys) and walks the union of
// their trees in the reference's visiting order (:160-241). State: the current
// record v and the mask M of lanes that need it. Each lane takes exactly the
// reference's accept/open decisions (in the reference's IEEE arithmetic where a
// decision is close; see trav_prep). If some lanes open v, the
// walk descends to v + 1 (its first child in this order) with those lanes,
// remembering (skip[v], M); otherwise it continues at skip[v]. Arriving at a
// remembered position restores its mask. Positions only grow (the skip is
// clamped to at least v + 1), so the walk takes at most one step per record,
// and the remembered entries are ancestors of v: at most the tree depth
// (<= 33), kept one per lane in VGPRs (lane k holds entry k). A lane's
// interactions are summed in the reference's order.
//
// Each step needs the record the previous step chose, and that load mostly
// misses the scalar cache. So the common step is written out by hand
// (NBX_WALK_FAST) and loads ahead: while
// it works on record v it already fetches both records the next step can
// need, v + 1 (descend) and skip[v] (continue). Three SGPR sets of 12 dwords
// rotate: the step on set X loads v + 1 into Y and skip[v] into Z, and
// continues with Y or Z (A -> B/C, B -> C/A, C -> A/B). Pops change only the
// mask, never v, so the next record is always one of the two. Prefetch
// addresses are clamped to the pad record after the last one (v + 1 <= nn
// inside a step), and every exit waits for the loads in flight. The common
// step is: every lane's theta decision clear of the band where the division
// decides (tested with the fused d2, see trav_prep), and every lane whose
// particle lies inside a leaf is at least as heavy as the leaf (m_eff <= 0 at
// :194-196, the lane's own single-body leaf: no force). Such lanes are handled
// as if they opened the leaf: they take no force, and the walk steps to v + 1
// with them, where the remembered entry (the leaf's end, v + 1) restores the
// mask at once. Any other step leaves the asm and runs as C code below (the
// own-leaf correction of a heavier leaf, the reference's d2 and the IEEE
// division); the walk then re-enters the asm.
//
// Registers: s[20:31], s[32:43], s[44:55] the record sets (dwords 0..11 of a
// record), s[56:57] M, s[58:59] the lanes inside the cell (then the undecided
// lanes), s[60:61] the lanes that open it, s[62:63] the lanes that clearly
// accept it, s[66:69] the lanes that take the force and those that descend,
// s70 = v + 1, s71 = skip clamped to [v + 1, nn]. s66/s67 hold the two load
// offsets until the loads issue (scalar loads read their offsets at issue),
// s58 the pop lane. The record after the skip target (the next step's descend
// target if this step continues) is pulled into the L2 a step ahead by a vector
// load into one scratch VGPR: it counts on the vector memory counter, so the
// step's wait for its scalar loads does not wait for it, and the next step's
// scalar load of it hits the L2 (3 % faster; also loading v + 2 ahead gave only
// 1.3 %, and the record two continues away, from a skip-of-skip field, was 5 %
// slower). s64/s65 stay free: the kernel needs no more than 80 SGPRs. Ordering: the result of v_rsq is first read two instructions later; lane sets pass from vector compares to
// scalar instructions (interlocked) and from scalar instructions to v_cndmask
// (no hazard); readlane takes its lane from a scalar instruction and writelane
// from m0, written one instruction before.
// The common path of one step on record set X (the record at v): the loads of
// the two possible next records into Y (v + 1) and Z (skip), the decisions,
// the force; it ends by continuing at the skip (set Z), which falls through to
// the next block when Z is laid out after X (TAIL empty), or descending, or
// leaving for C. The pop test also ends the walk: with no remembered entry,
// top holds nn. The skip is clamped to [v + 1, nn] and a step runs only at
// v < top <= nn, so both loads stay within the records and the pad record
// after them (at nn), and the walk ends whatever the records hold.
#define NBX_WALK_MAIN(X, Y, Z, c0, c1, c2, gm, b0, b1, b2, hh, sk, so, sa, y8, y4, z8, z4, TAIL) \
  "L_next_" X "_%=:\n"                                                           \
  "s_cmp_ge_u32 %[v], %[top]\n"                                                  \
  "s_cbranch_scc1 L_pop_" X "_%=\n"                                              \
  "s_waitcnt lgkmcnt(0)\n"                                                       \
  "s_add_i32 s70, %[v], 1\n"                                                     \
  "s_max_i32 s71, " sk ", s70\n"                                                 \
  "s_lshl_b32 s66, s70, 6\n"                                                     \
  "s_min_i32 s71, s71, %[nn]\n"                                                  \
  "s_lshl_b32 s67, s71, 6\n"                                                     \
  "s_load_dwordx8 " y8 ", %[rb], s66\n"                                          \
  "s_load_dwordx4 " y4 ", %[rb], s66 offset:32\n"                                \
  "s_load_dwordx8 " z8 ", %[rb], s67\n"                                          \
  "s_load_dwordx4 " z4 ", %[rb], s67 offset:32\n"                                \
  "v_mov_b32 %[po], s67\n"                                                       \
  "global_load_dword %[pf], %[po], %[rb] offset:64\n"                            \
  "v_subrev_f32 %[t0], " b0 ", %[xi]\n"                                          \
  "v_subrev_f32 %[t1], " b1 ", %[yi]\n"                                          \
  "v_subrev_f32 %[t2], " b2 ", %[zi]\n"                                          \
  "v_max3_f32 %[t0], |%[t0]|, |%[t1]|, |%[t2]|\n"                                \
  "v_cmp_le_f32_e64 s[58:59], %[t0], |" hh "|\n"                                 \
  "v_sub_f32 %[dx], " c0 ", %[xi]\n"                                             \
  "v_sub_f32 %[dy], " c1 ", %[yi]\n"                                             \
  "v_sub_f32 %[dz], " c2 ", %[zi]\n"                                             \
  "v_fma_f32 %[d2], %[dx], %[dx], %[eps2]\n"                                     \
  "v_fmac_f32 %[d2], %[dy], %[dy]\n"                                             \
  "v_fmac_f32 %[d2], %[dz], %[dz]\n"                                             \
  "v_rsq_f32 %[inv], %[d2]\n"                                                    \
  "v_cmp_gt_f32_e64 s[60:61], " so ", %[d2]\n"                                   \
  "v_cmp_lt_f32_e64 s[62:63], " sa ", %[d2]\n"                                   \
  "v_mul_f32 %[t1], %[inv], %[inv]\n"                                            \
  "v_mul_f32 %[t1], %[t1], %[inv]\n"                                             \
  "v_mul_f32 %[t1], " gm ", %[t1]\n"                                             \
  "s_and_b64 s[58:59], s[58:59], s[56:57]\n"                                     \
  "s_cbranch_scc1 L_inside_" X "_%=\n"                                           \
  "L_back_" X "_%=:\n"                                                           \
  "s_or_b64 s[58:59], s[60:61], s[62:63]\n"                                      \
  "s_andn2_b64 s[58:59], s[56:57], s[58:59]\n"                                   \
  "s_cbranch_scc1 L_exit_%=\n"                                                   \
  "s_andn2_b64 s[66:67], s[56:57], s[60:61]\n"                                   \
  "s_and_b64 s[68:69], s[56:57], s[60:61]\n"                                     \
  "v_cndmask_b32_e64 %[t1], 0, %[t1], s[66:67]\n"                                \
  "v_fmac_f32 %[sx], %[t1], %[dx]\n"                                             \
  "v_fmac_f32 %[sy], %[t1], %[dy]\n"                                             \
  "v_fmac_f32 %[sz], %[t1], %[dz]\n"                                             \
  "s_cbranch_scc1 L_desc_" X "_%=\n"                                             \
  "s_mov_b32 %[v], s71\n" TAIL

// The rarer paths of a step on set X, laid out after the three common paths:
// descend (remember the skip and the mask, continue at v + 1 in set Y), a lane
// inside the cell (open an internal cell; a leaf no heavier than the lane's
// particle, or leave for C), and the pops (restore the remembered mask; with
// none left the walk is over).
#define NBX_WALK_OOL(X, Y, hh, so)                                               \
  "L_desc_" X "_%=:\n"                                                           \
  "s_mov_b32 m0, %[depth]\n"                                                     \
  "s_add_i32 %[depth], %[depth], 1\n"                                            \
  "v_writelane_b32 %[kat], s71, m0\n"                                            \
  "v_writelane_b32 %[klo], s56, m0\n"                                            \
  "v_writelane_b32 %[khi], s57, m0\n"                                            \
  "s_mov_b32 %[top], s71\n"                                                      \
  "s_mov_b64 s[56:57], s[68:69]\n"                                               \
  "s_mov_b32 %[v], s70\n"                                                        \
  "s_branch L_next_" Y "_%=\n"                                                   \
  "L_inside_" X "_%=:\n"                                                         \
  "s_cmp_lt_i32 " hh ", 0\n"                                                     \
  "s_cbranch_scc0 L_open_" X "_%=\n"                                             \
  "v_cmp_ge_f32_e64 s[66:67], %[mi], -" so "\n"                                  \
  "s_andn2_b64 s[66:67], s[58:59], s[66:67]\n"                                   \
  "s_cbranch_scc1 L_exit_%=\n"                                                   \
  "L_open_" X "_%=:\n"                                                           \
  "s_or_b64 s[60:61], s[60:61], s[58:59]\n"                                      \
  "s_branch L_back_" X "_%=\n"                                                   \
  "L_pop_" X "_%=:\n"                                                            \
  "s_cmp_eq_u32 %[depth], 0\n"                                                   \
  "s_cbranch_scc1 L_exit_%=\n"                                                   \
  "s_add_i32 %[depth], %[depth], -1\n"                                           \
  "v_readlane_b32 s56, %[klo], %[depth]\n"                                       \
  "v_readlane_b32 s57, %[khi], %[depth]\n"                                       \
  "s_add_i32 s58, %[depth], -1\n"                                                \
  "s_cmp_lt_i32 s58, 0\n"                                                        \
  "s_cbranch_scc1 L_root_" X "_%=\n"                                             \
  "v_readlane_b32 %[top], %[kat], s58\n"                                         \
  "s_branch L_next_" X "_%=\n"                                                   \
  "L_root_" X "_%=:\n"                                                           \
  "s_mov_b32 %[top], %[nn]\n"                                                    \
  "s_branch L_next_" X "_%=\n"

// Layout A, C, B: the continue path of A (to C) and of C (to B) falls through.
#define NBX_WALK_FAST                                                                       \
  "s_mov_b64 s[56:57], %[M]\n"                                                              \
  "s_min_i32 s66, %[v], %[nn]\n"                                                            \
  "s_max_i32 s66, s66, 0\n"                                                                 \
  "s_lshl_b32 s66, s66, 6\n"                                                                \
  "s_load_dwordx8 s[20:27], %[rb], s66\n"                                                   \
  "s_load_dwordx4 s[28:31], %[rb], s66 offset:32\n"                                         \
  NBX_WALK_MAIN("A", "B", "C", "s20", "s21", "s22", "s23", "s24", "s25", "s26", "s27", "s28", \
                "s30", "s31", "s[32:39]", "s[40:43]", "s[44:51]", "s[52:55]", "")           \
  NBX_WALK_MAIN("C", "A", "B", "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", \
                "s54", "s55", "s[20:27]", "s[28:31]", "s[32:39]", "s[40:43]", "")           \
  NBX_WALK_MAIN("B", "C", "A", "s32", "s33", "s34", "s35", "s36", "s37", "s38", "s39", "s40", \
                "s42", "s43", "s[44:51]", "s[52:55]", "s[20:27]", "s[28:31]",             \
                "s_branch L_next_A_%=\n")                                                   \
  NBX_WALK_OOL("A", "B", "s27", "s30")                                                      \
  NBX_WALK_OOL("C", "A", "s51", "s54")                                                      \
  NBX_WALK_OOL("B", "C", "s39", "s42")                                                      \
  "L_exit_%=:\n"                                                                            \
  "s_waitcnt vmcnt(0) lgkmcnt(0)\n"                                                         \
  "s_mov_b64 %[M], s[56:57]\n"

// The decisions of one walk step at record v for the wave's lanes M (the C
// form of the asm's common step, covering every case): the lanes that open the
// node (om, descend), the lanes that take its force (im), the force operands
// and the position after the node's subtree. The decisions use the reference's
// d2; the force uses the fused d2 of the asm step, so a lane's sum does not
// depend on which of its steps ran here (nor on which task it was in).
struct WalkStep {
  uint64_t om, im;
  float dx, dy, dz, d2f, gm;  // d2f: the fused d2 the force uses in every step
  int after;
};

__device__ __forceinline__ WalkStep walk_decide(const float4* __restrict__ rec, int v, uint64_t M,
                                                float xi, float yi, float zi, float mi,
                                                uint64_t me, float theta2, float G, float eps2) {
  const char* rp = reinterpret_cast<const char*>(rec) + (uint32_t)v * 64u;
  const float4 c4 = *reinterpret_cast<const float4*>(rp);
  const float4 b4 = *reinterpret_cast<const float4*>(rp + 16);
  const float4 a4 = *reinterpret_cast<const float4*>(rp + 32);
  const float mass = reinterpret_cast<const float4*>(rp + 48)->x;
  const int after = max(__float_as_int(a4.x), v + 1);
  const float size2 = a4.y, s_open = a4.z, s_acc = a4.w;
  const float h = fabsf(b4.w);
  const float far = fmaxf(fmaxf(fabsf(xi - b4.x), fabsf(yi - b4.y)), fabsf(zi - b4.z));
  const uint64_t cm = __ballot(far <= h) & M;  // :180-182, particle inside the cell
  float dx = sub_rn(c4.x, xi), dy = sub_rn(c4.y, yi), dz = sub_rn(c4.z, zi);
  float gm = c4.w;
  uint64_t o = 0;     // lanes that open the node
  uint64_t gone = 0;  // lanes whose own leaf holds no other mass
  if (cm) {
    if (__builtin_signbit(b4.w)) {  // :190-200, the particle's own leaf
      const bool own = (cm & me) != 0;
      const float meff = own ? sub_rn(mass, mi) : mass;
      const bool fix = own && meff > 0.0f;
      if (__ballot(fix)) {  // remove the particle from a leaf that holds others
        const float cx = div_rn(sub_rn(mul_rn(c4.x, mass), mul_rn(xi, mi)), meff);
        const float cy = div_rn(sub_rn(mul_rn(c4.y, mass), mul_rn(yi, mi)), meff);
        const float cz = div_rn(sub_rn(mul_rn(c4.z, mass), mul_rn(zi, mi)), meff);
        dx = fix ? sub_rn(cx, xi) : dx;
        dy = fix ? sub_rn(cy, yi) : dy;
        dz = fix ? sub_rn(cz, zi) : dz;
      }
      gm = own ? mul_rn(G, meff) : gm;
      gone = __ballot(meff <= 0.0f) & cm;
    } else {
      o = cm;  // :222, an internal node holding the particle is always opened
    }
  }
  const float d2 = add_rn(add_rn(add_rn(mul_rn(dx, dx), mul_rn(dy, dy)), mul_rn(dz, dz)), eps2);
  // :218-224, open when contains || !(size2 / d2 < theta2): the thresholds
  // decide outside a narrow band (trav_prep); inside it the IEEE division.
  o |= __ballot(d2 < s_open);
  const uint64_t am = M & ~(o | __ballot(d2 > s_acc));
  if (am) {
    const bool mine = (am & me) != 0;
    o |= __ballot(mine && !(div_rn(size2, d2) < theta2)) & am;
  }
  WalkStep w;
  w.om = o & M;
  w.im = M & ~o & ~gone;  // lanes that take the node's force
  w.dx = dx;
  w.dy = dy;
  w.dz = dz;
  w.d2f = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
  w.gm = gm;
  w.after = after;
  return w;
}

// One task: the 64 particles [64 blk, 64 blk + 64) of the order.
__device__ __forceinline__ void walk_task(int blk, int lane, const float4* __restrict__ rec,
                                          const BhState* __restrict__ st,
                                          const int32_t* __restrict__ order,
                                          const float* __restrict__ x, const float* __restrict__ y,
                                          const float* __restrict__ z, const float* __restrict__ m,
                                          int n, float theta2, float G, float eps2, float* ax,
                                          float* ay, float* az) {
  const int r = blk * 64 + lane;
  const bool valid = r < n;
  const int ip = valid ? order[r] : 0;
  const float xi = x[ip], yi = y[ip], zi = z[ip], mi = m[ip];
  float sx = 0.f, sy = 0.f, sz = 0.f;
  uint64_t M = __ballot(valid);
  const int nn = st->final_count;
  const uint64_t me = 1ull << lane;
  int keep_at = 0;              // lane k: position of remembered entry k
  uint32_t keep_lo = 0, keep_hi = 0;  // lane k: its mask
  int depth = 0;
  int top = nn;  // position of the innermost remembered entry (nn: none)
  uint32_t pf = 0;  // the asm's L2 prefetches land here (drained before the asm exits)
  int v = 0;
  for (;;) {
    // The compiler cannot prove the walk state uniform after a C step, nor the
    // asm's scalar outputs: readfirstlane states it on both sides.
    v = __builtin_amdgcn_readfirstlane(v);
    top = __builtin_amdgcn_readfirstlane(top);
    depth = __builtin_amdgcn_readfirstlane(depth);
    M = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(M >> 32)) << 32) |
        (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)M);
    float t0, t1, t2, dxa, dya, dza, d2a, inva;
    uint32_t po;
    asm volatile(NBX_WALK_FAST
                 : [v] "+s"(v), [top] "+s"(top), [depth] "+s"(depth), [M] "+s"(M),
                   [sx] "+v"(sx), [sy] "+v"(sy), [sz] "+v"(sz), [kat] "+v"(keep_at),
                   [klo] "+v"(keep_lo), [khi] "+v"(keep_hi), [t0] "=&v"(t0), [t1] "=&v"(t1),
                   [t2] "=&v"(t2), [dx] "=&v"(dxa), [dy] "=&v"(dya), [dz] "=&v"(dza),
                   [d2] "=&v"(d2a), [inv] "=&v"(inva), [pf] "+v"(pf), [po] "=&v"(po)
                 : [rb] "s"(rec), [nn] "s"(nn), [eps2] "v"(eps2), [xi] "v"(xi), [yi] "v"(yi),
                   [zi] "v"(zi), [mi] "v"(mi)
                 : "s20", "s21", "s22", "s23", "s24", "s25", "s26", "s27", "s28", "s29", "s30",
                   "s31", "s32", "s33", "s34", "s35", "s36", "s37", "s38", "s39", "s40", "s41",
                   "s42", "s43", "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52",
                   "s53", "s54", "s55", "s56", "s57", "s58", "s59", "s60", "s61", "s62", "s63",
                   "s66", "s67", "s68", "s69", "s70", "s71", "m0", "scc", "memory");
    v = __builtin_amdgcn_readfirstlane(v);
    top = __builtin_amdgcn_readfirstlane(top);
    depth = __builtin_amdgcn_readfirstlane(depth);
    M = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(M >> 32)) << 32) |
        (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)M);
    if (v >= nn) break;
    // A step the asm leaves to C: record v (v != top here).
    const WalkStep w = walk_decide(rec, v, M, xi, yi, zi, mi, me, theta2, G, eps2);
    const int after = w.after;
    const uint64_t om = w.om, im = w.im;
    const float dx = w.dx, dy = w.dy, dz = w.dz, gm = w.gm;
    const float inv = __builtin_amdgcn_rsqf(w.d2f);
    const float fi = gm * (inv * inv * inv);
    const float f = ((im >> lane) & 1) ? fi : 0.0f;
    sx = __builtin_fmaf(f, dx, sx);
    sy = __builtin_fmaf(f, dy, sy);
    sz = __builtin_fmaf(f, dz, sz);
    if (om) {  // :234-240, descend; remember where this subtree ends
      // writelane takes its lane from m0 (one scalar operand per vector
      // instruction); the s_nop separates the m0 write from its first read.
      asm volatile(
          "s_mov_b32 m0, %[d]\n"
          "s_nop 0\n"
          "v_writelane_b32 %[kat], %[aft], m0\n"
          "v_writelane_b32 %[klo], %[mlo], m0\n"
          "v_writelane_b32 %[khi], %[mhi], m0\n"
          : [kat] "+v"(keep_at), [klo] "+v"(keep_lo), [khi] "+v"(keep_hi)
          : [aft] "s"(__builtin_amdgcn_readfirstlane(after)),
            [mlo] "s"(__builtin_amdgcn_readfirstlane((int)(uint32_t)M)),
            [mhi] "s"(__builtin_amdgcn_readfirstlane((int)(uint32_t)(M >> 32))),
            [d] "s"(__builtin_amdgcn_readfirstlane(depth))
          : "m0");
      ++depth;
      top = min(after, nn);
      M = om;
      v = v + 1;
    } else {
      v = after;
    }
  }
  if (valid) {
    ax[ip] = sx;
    ay[ip] = sy;
    az[ip] = sz;
  }
}

// Persistent workgroups of kTravWaves waves, one grid of resident waves.
// Tasks are 64 consecutive particles of the task order (task_keys: the tree's
// own order re-sorted by Hilbert index), so bodies of one cell share a wave
// however many share a Morton code, and a task's bodies are close together.
// Workgroups are dealt to the 8 XCDs round-robin; each XCD owns one contiguous
// eighth of the tasks, so that its L2 holds the part of the tree its waves
// walk, and its workgroups take that eighth's tasks in order from a counter,
// kTravWaves adjacent tasks per round, one per wave. The waves of a workgroup
// share a CU and its scalar cache, and adjacent tasks walk mostly the same
// records at about the same time: 8-wave rounds raise the scalar-cache hits
// from 8 to 12 % and save 2-3 % of the traversal, more than the waves lose
// waiting at the end of a round (1-wave workgroups taking one task at a time
// keep the tail shortest but share nothing). Every workgroup exits once its
// XCD's counter is past the range. (Taking the other XCDs' leftovers was
// slower: those walks miss the L2 that holds their part of the tree.)
constexpr int kTravWaves = 8;
__global__ __launch_bounds__(64 * kTravWaves) void bh_traverse(
    const float4* __restrict__ rec, const BhState* __restrict__ st,
    const int32_t* __restrict__ task_order, int* queue, int ntasks, const float* __restrict__ x, const float* __restrict__ y,
    const float* __restrict__ z, const float* __restrict__ m, int n, float theta2, float G,
    float eps2, float* ax, float* ay, float* az) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  __shared__ int round_base;
  // The walk's asm clobbers s20-s71, which keeps the kernel at 78 SGPRs: on
  // gfx950, 8 waves per SIMD stay resident up to 78 SGPRs and only 7 from 82
  // on (measured with spinning waves), although the compiler's occupancy and
  // the occupancy query report 8 either way. The per-task values below are only
  // used as per-lane addresses, so they are held in VGPRs (an empty asm hides
  // their uniformity): in SGPRs they would raise the count past that limit,
  // and a latency-bound walk pays for the lost wave in full.
  auto vreg = [](const void* p) {
    uint64_t v = reinterpret_cast<uint64_t>(p);
    asm volatile("" : "+v"(v));
    return v;
  };
  const int32_t* order =
      reinterpret_cast<const int32_t*>(vreg(task_order));
  const float* vx = reinterpret_cast<const float*>(vreg(x));
  const float* vy = reinterpret_cast<const float*>(vreg(y));
  const float* vz = reinterpret_cast<const float*>(vreg(z));
  const float* vm = reinterpret_cast<const float*>(vreg(m));
  float* vax = reinterpret_cast<float*>(vreg(ax));
  float* vay = reinterpret_cast<float*>(vreg(ay));
  float* vaz = reinterpret_cast<float*>(vreg(az));
  const int xq = (int)blockIdx.x % 8;
  const int per = ntasks / 8, rem = ntasks % 8;
  int b = xq < rem ? xq * (per + 1) : rem * (per + 1) + (xq - rem) * per;
  int len = per + (xq < rem ? 1 : 0);
  asm volatile("" : "+v"(b), "+v"(len));
  int* ctr = reinterpret_cast<int*>(vreg(queue + 16 * xq));
  asm volatile("" : "+v"(theta2), "+v"(G), "+v"(eps2));  // operands of vector instructions only
  for (;;) {
    if (threadIdx.x == 0) round_base = atomicAdd(ctr, kTravWaves);
    __syncthreads();
    const int c = __builtin_amdgcn_readfirstlane(round_base);
    __syncthreads();  // every wave has read round_base before the next round writes it
    if (c >= __builtin_amdgcn_readfirstlane(len)) break;  // the same for the whole workgroup
    const int t = c + wave;
    if (t < __builtin_amdgcn_readfirstlane(len))
      walk_task(__builtin_amdgcn_readfirstlane(b) + t, lane, rec, st, order, vx, vy, vz, vm, n,
                theta2, G, eps2, vax, vay, vaz);
  }
}

__global__ void trav_queue_reset(int* queue) {
  if (threadIdx.x < 8) queue[16 * threadIdx.x] = 0;
}

// Diagnostics: the walk of bh_traverse in C (pops included) over the same tasks
// with the same decisions (walk_decide), counted instead of summed: records the
// waves visit (one per step), particle-node force evaluations, openings and
// active lane-steps, added to counts[0..3]. One wave per workgroup; the stack
// of remembered (position, mask) entries sits in LDS, written by every lane
// with the same values.
__global__ __launch_bounds__(64) void bh_walk_census(
    const float4* __restrict__ rec, const BhState* __restrict__ st,
    const int32_t* __restrict__ order, int ntasks,
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    const float* __restrict__ m, int n, float theta2, float G, float eps2,
    unsigned long long* counts) {
  __shared__ int at[kMaxDepth + 8];
  __shared__ unsigned long long mask[kMaxDepth + 8];
  const int lane = threadIdx.x;
  const int nn = st->final_count;
  const uint64_t me = 1ull << lane;
  unsigned long long steps = 0, forces = 0, opens = 0, active = 0;
  for (int blk = blockIdx.x; blk < ntasks; blk += gridDim.x) {
    const int r = blk * 64 + lane;
    const bool valid = r < n;
    const int ip = valid ? order[r] : 0;
    const float xi = x[ip], yi = y[ip], zi = z[ip], mi = m[ip];
    uint64_t M = __ballot(valid);
    int depth = 0, v = 0;
    while (v < nn) {
      if (depth > 0 && v == at[depth - 1]) {
        --depth;
        M = mask[depth];
        continue;
      }
      const WalkStep w = walk_decide(rec, v, M, xi, yi, zi, mi, me, theta2, G, eps2);
      ++steps;
      forces += __popcll(w.im);
      opens += __popcll(w.om);
      active += __popcll(M);
      if (w.om && depth < kMaxDepth + 8) {
        at[depth] = w.after;
        mask[depth] = M;
        ++depth;
        M = w.om;
        v = v + 1;
      } else {
        v = w.after;
      }
    }
  }
  if (lane == 0) {
    atomicAdd(&counts[0], steps);
    atomicAdd(&counts[1], forces);
    atomicAdd(&counts[2], opens);
    atomicAdd(&counts[3], active);
  }
}

__global__ void census_clear(unsigned long long* counts) {
  if (threadIdx.x < 4) counts[threadIdx.x] = 0;
}

}  // namespace

}  // namespace nbx

// ---------------------------------------------------------------- object

struct nbx_bh {
  size_t max_particles = 0;
  int max_nodes = 0;
  nbx::BhState* st = nullptr;
  float* bbox_part = nullptr;
  int bbox_blocks = 0;
  uint32_t* codes = nullptr;
  int32_t* idx = nullptr;
  uint4* path = nullptr;
  float4* sbody = nullptr;  // particles in sorted order: x, y, z, m
  nbx::TmpNodes T{};
  nbx::PathOrder O{};
  int pos_blocks = 0;  // blocks of kPosBlock positions at max_particles
  uint32_t* nkeys = nullptr;
  int32_t* nvals = nullptr;
  int32_t* final_of = nullptr;
  nbx::RefNode* nodes = nullptr;
  float4* trav_rec = nullptr;  // 64 bytes per cell, traversal order
  uint32_t* hkey = nullptr;       // Hilbert index of each particle's Morton cell
  uint32_t* task_key = nullptr;   // the same, by position of the task order (sort keys)
  int32_t* task_order = nullptr;  // particles in the order traversal tasks take them
  nbx::RadixSort sort_particles;
  nbx::RadixSort sort_nodes;
  bool built = false;
  size_t built_n = 0;
  // (1 / theta^2, G) the traversal records hold; set by the last trav_prep
  nbx::TravParams prep{0.0f, 0.0f, 0};
  bool nodes_pending = false;  // the reference node array is emitted on demand
  int* trav_queue = nullptr;     // 8 task counters, 64 bytes apart
  unsigned long long* census = nullptr;  // nbx_bh_walk_census counters
  int trav_groups = 0;           // resident traversal workgroups on this device
  int force_seq = 0;
  int task_tree_order = 0;  // 1: tasks in the tree's own order (no Hilbert sort)
  bool timing = false;
  hipEvent_t ev[10] = {};
  double phase_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<void*> allocs;
};

namespace nbx {
namespace {

template <typename T>
int dalloc(nbx_bh* t, T** p, size_t count) {
  void* q = nullptr;
  if (hipMalloc(&q, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(NBX_ERR_NOMEM, "Barnes-Hut buffers: allocation of %zu bytes failed",
                sizeof(T) * count);
  }
  t->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return NBX_OK;
}

void bh_release(nbx_bh* t) {
  for (void* p : t->allocs) (void)hipFree(p);
  t->allocs.clear();
  t->sort_particles.release();
  t->sort_nodes.release();
  for (auto& e : t->ev)
    if (e) (void)hipEventDestroy(e);
}

int grid_for(size_t n, int cap = 2048) {
  return (int)std::max<size_t>(1, std::min<size_t>((n + kThreads - 1) / kThreads, cap));
}

int key_bits_for(size_t n) {
  int b = 0;
  while ((size_t(1) << b) <= n) ++b;  // k + 1 <= n
  return b + 7;
}

void mark(nbx_bh* t, int k, hipStream_t s) {
  if (t->timing) (void)hipEventRecord(t->ev[k], s);
}

}  // namespace
}  // namespace nbx

using namespace nbx;

extern "C" {

int nbx_bh_create(size_t max_particles, nbx_bh** out) {
  if (!out) return fail(NBX_ERR_INVALID, "nbx_bh_create: null output");
  *out = nullptr;
  if (max_particles == 0) return fail(NBX_ERR_INVALID, "Barnes-Hut tree needs max_particles > 0");
  if (key_bits_for(max_particles) > 32)
    return fail(NBX_ERR_INVALID, "Barnes-Hut tree supports at most 2^25-1 particles (got %zu)",
                max_particles);
  nbx_bh* t = new (std::nothrow) nbx_bh();
  if (!t) return fail(NBX_ERR_NOMEM, "nbx_bh_create: host allocation");
  const size_t P = max_particles;
  const size_t NN = 2 * P;  // max_nodes_ = 2 * max_particles (:253)
  t->max_particles = P;
  t->max_nodes = (int)NN;
  t->bbox_blocks = grid_for(P, 1024);
  int rc = NBX_OK;
  auto A = [&](auto** p, size_t c) {
    if (rc == NBX_OK) rc = dalloc(t, p, c);
  };
  A(&t->st, 1);
  A(&t->bbox_part, (size_t)t->bbox_blocks * 6);
  A(&t->codes, P);
  A(&t->idx, P);
  A(&t->path, P);
  A(&t->sbody, P);
  A(&t->T.t1, NN);
  A(&t->T.t2, NN);
  A(&t->T.k, NN);
  A(&t->T.count, NN);
  A(&t->T.pidx, NN);
  A(&t->T.info, NN);
  A(&t->T.child, NN * 8);
  A(&t->T.box, NN);
  A(&t->T.com, NN);
  A(&t->T.skip, NN);
  A(&t->T.sowner, NN);
  A(&t->T.size, NN);
  A(&t->T.pre, NN);
  t->pos_blocks = (int)(P / kPosBlock + 1);
  const size_t supers = (size_t)t->pos_blocks / kSuperBlocks + 1;
  A(&t->O.key, P);
  A(&t->O.key_hi, P);
  A(&t->O.rank, P);
  A(&t->O.spath, P);
  A(&t->O.lcp, (size_t)t->pos_blocks * kPosBlock + 64);
  A(&t->O.lmin1, (size_t)t->pos_blocks * (kPosBlock / 64) + 64);
  A(&t->O.lmin2, supers * kSuperBlocks + 64);
  A(&t->O.lmin3, (supers / 64 + 1) * 64 + 64);
  A(&t->O.first_pre, P + 1);
  A(&t->O.bsum, (size_t)t->pos_blocks);
  A(&t->O.lvl, (size_t)lvl_base((int)(P / kLocalBlock + 1), kMaxLevel + 1) + 64);
  A(&t->O.runs, P);
  A(&t->O.lrun, P / (kRefineMax + 1) + 1);
  A(&t->O.lkey[0], P);
  A(&t->O.lkey[1], P);
  A(&t->O.lval[0], P);
  A(&t->O.lval[1], P);
  A(&t->O.bstat, (size_t)t->pos_blocks * 3);
  A(&t->O.ppart, P);
  A(&t->nkeys, NN);
  A(&t->nvals, NN);
  A(&t->final_of, NN);
  A(&t->nodes, NN);
  A(&t->trav_rec, (NN + 2) * 4);  // + two pad records the walk's loads ahead may read
  A(&t->hkey, P);
  A(&t->task_key, P);
  A(&t->task_order, P);
  A(&t->trav_queue, 8 * 16);
  A(&t->census, 4);
  if (rc == NBX_OK) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(bh_traverse), 64 * kTravWaves, 0) !=
            hipSuccess)
      rc = fail(NBX_ERR_HIP, "Barnes-Hut: occupancy query failed");
    t->trav_groups = std::max(8, cus * per_cu);
  }
  if (rc == NBX_OK) rc = t->sort_particles.init((int)P);
  if (rc == NBX_OK) rc = t->sort_nodes.init((int)NN);
  for (auto& e : t->ev)
    if (rc == NBX_OK && hipEventCreate(&e) != hipSuccess) rc = fail(NBX_ERR_HIP, "hipEventCreate");
  if (rc != NBX_OK) {
    bh_release(t);
    delete t;
    return rc;
  }
  *out = t;
  return NBX_OK;
}

int nbx_bh_destroy(nbx_bh* t) {
  if (!t) return NBX_OK;
  bh_release(t);
  delete t;
  return NBX_OK;
}

int nbx_bh_build(nbx_bh* t, const nbx_particles* d, nbx_stream stream) {
  if (!t) return fail(NBX_ERR_INVALID, "nbx_bh_build: null tree");
  NBX_TRY(check_particles(d, "nbx_bh_build"));
  if (d->count > t->max_particles)
    return fail(NBX_ERR_CAPACITY, "tree built for %zu particles, got %zu", t->max_particles,
                d->count);
  hipStream_t s = as_stream(stream);
  const int n = (int)d->count;
  mark(t, 0, s);
  const int bb = std::min(t->bbox_blocks, grid_for(n, 1024));
  hipLaunchKernelGGL(bbox_partial, dim3(bb), dim3(kThreads), 0, s, d->pos_x, d->pos_y, d->pos_z,
                     n, t->bbox_part);
  hipLaunchKernelGGL(bbox_final, dim3(1), dim3(kThreads), 0, s, t->bbox_part, bb, n, t->st);
  NBX_LAUNCHED();
  mark(t, 1, s);
  // the packed particles borrow the path-order buffer, which path_lcp_kernel fills later
  float4* packed = reinterpret_cast<float4*>(t->O.spath);
  hipLaunchKernelGGL(morton_kernel, dim3(div_up(n, kThreads)), dim3(kThreads), 0, s, d->pos_x,
                     d->pos_y, d->pos_z, d->mass, n, t->st, t->codes, t->idx, packed, t->hkey);
  NBX_LAUNCHED();
  mark(t, 2, s);
  NBX_TRY(t->sort_particles.sort(t->codes, t->idx, &t->st->n, 30, s));
  mark(t, 3, s);
  hipLaunchKernelGGL(path_kernel, dim3(div_up(n, kThreads)), dim3(kThreads), 0, s, packed,
                     t->idx, n, t->st, t->path, t->sbody, t->O.key, t->O.key_hi, t->O.rank);
  NBX_LAUNCHED();
  mark(t, 4, s);
  // path order: stable sort by digits 10..19, then by digits 0..9
  NBX_TRY(t->sort_particles.sort(t->O.key, t->O.rank, &t->st->n, 30, s));
  hipLaunchKernelGGL(path_gather_key, dim3(div_up(n, kThreads)), dim3(kThreads), 0, s, &t->st->n,
                     t->O);
  NBX_TRY(t->sort_particles.sort(t->O.key, t->O.rank, &t->st->n, 30, s));
  hipLaunchKernelGGL(path_lcp_kernel, dim3(div_up((size_t)n + 64, kThreads)), dim3(kThreads), 0, s,
                     t->st, t->path, t->O);
  hipLaunchKernelGGL(path_refine, dim3(grid_for((size_t)n * 64, 1024)), dim3(kThreads), 0, s,
                     t->st, t->path, t->O);
  hipLaunchKernelGGL(path_refine_long, dim3(64), dim3(kLongThreads), 0, s, t->st, t->path, t->O);
  const int pb = (int)div_up(n, kPosBlock);
  hipLaunchKernelGGL(cell_count, dim3(n / kPosBlock + 1), dim3(kThreads), 0, s, t->st, t->O);
  hipLaunchKernelGGL(cell_offsets, dim3(1), dim3(1024), 0, s, t->st, t->O, t->T, t->max_nodes);
  NBX_LAUNCHED();
  if (t->force_seq) {
    // diagnostic: route through the exact sequential builder
    const int one = 1, zero = 0;
    NBX_HIP(hipMemcpyAsync(&t->st->use_seq, &one, sizeof(int), hipMemcpyHostToDevice, s));
    NBX_HIP(hipMemcpyAsync(&t->st->sort_n, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    NBX_HIP(hipStreamSynchronize(s));
  }
  hipLaunchKernelGGL(cell_first_pre, dim3(pb), dim3(kThreads), 0, s, t->st, t->O);
  hipLaunchKernelGGL(cell_emit, dim3(div_up(n, kThreads)), dim3(kThreads), 0, s, t->st, t->O,
                     t->T, t->idx, t->sbody, t->trav_rec, t->prep);
  hipLaunchKernelGGL(cell_ends, dim3(grid_for((size_t)t->max_nodes, 4096)), dim3(kThreads), 0, s,
                     t->st, t->O, t->T);
  hipLaunchKernelGGL(cell_link, dim3(grid_for((size_t)t->max_nodes, 4096)), dim3(kThreads), 0, s,
                     t->st, t->T);
  NBX_LAUNCHED();
  mark(t, 5, s);
  hipLaunchKernelGGL(cell_local, dim3(div_up(n, kLocalBlock)), dim3(kThreads), 0, s, t->st, t->O,
                     t->T, t->trav_rec, t->prep);
  {
    const int nbl = (int)div_up(n, kLocalBlock);
    for (int L = 2; L <= root_level(nbl); ++L)
      hipLaunchKernelGGL(cell_group, dim3(div_up(lvl_units(nbl, L), kFan)), dim3(kThreads), 0, s,
                         t->st, t->O, t->T, t->trav_rec, t->prep, L);
  }
  NBX_LAUNCHED();
  mark(t, 6, s);
  t->nodes_pending = true;  // the reference node array is numbered and emitted on demand
  hipLaunchKernelGGL(seq_build, dim3(1), dim3(64), 0, s, t->st, t->idx, d->pos_x, d->pos_y,
                     d->pos_z, d->mass, t->nodes, t->max_nodes);
  hipLaunchKernelGGL(seq_preorder, dim3(1), dim3(64), 0, s, t->st, t->nodes, t->T.pre, t->T.size,
                     t->T.k, t->trav_rec, t->prep);
  hipLaunchKernelGGL(task_keys, dim3(div_up(n, kThreads)), dim3(kThreads), 0, s, t->st, t->hkey,
                     t->idx, t->O.ppart, t->task_key, t->task_order);
  NBX_LAUNCHED();
  if (!t->task_tree_order) NBX_TRY(t->sort_particles.sort(t->task_key, t->task_order, &t->st->n, 32, s));
  mark(t, 7, s);
  t->built = true;
  t->built_n = d->count;
  return NBX_OK;
}

// The per-call fields of the traversal records (thresholds for theta, G * m),
// rewritten only when (theta, G) differ from the last call's.
static int ensure_prep(nbx_bh* t, float theta2, float G, hipStream_t s) {
  const nbx::TravParams want{1.0f / theta2, G, 1};
  if (!t->prep.valid || std::memcmp(&want, &t->prep, sizeof(want)) != 0) {
    hipLaunchKernelGGL(trav_prep, dim3(grid_for((size_t)t->max_nodes, 2048)), dim3(kThreads), 0,
                       s, t->trav_rec, t->st, want.inv_theta2, want.G);
    NBX_LAUNCHED();
    t->prep = want;
  }
  return NBX_OK;
}

int nbx_bh_forces(nbx_bh* t, nbx_particles* d, float theta, float G, float eps, nbx_stream stream) {
  if (!t) return fail(NBX_ERR_INVALID, "nbx_bh_forces: null tree");
  NBX_TRY(check_particles(d, "nbx_bh_forces"));
  if (!t->built || t->built_n != d->count)
    return fail(NBX_ERR_STATE, "nbx_bh_forces: build the tree for these particles first");
  hipStream_t s = as_stream(stream);
  const int n = (int)d->count;
  mark(t, 8, s);
  const float theta2 = theta * theta;
  NBX_TRY(ensure_prep(t, theta2, G, s));
  const int ntasks = (int)div_up(n, 64);
  const int groups =
      std::max(8, std::min((ntasks + kTravWaves - 1) / kTravWaves, t->trav_groups) & ~7);
  hipLaunchKernelGGL(trav_queue_reset, dim3(1), dim3(64), 0, s, t->trav_queue);
  hipLaunchKernelGGL(bh_traverse, dim3(groups), dim3(64 * kTravWaves), 0, s, t->trav_rec, t->st,
                     t->task_order, t->trav_queue, ntasks, d->pos_x, d->pos_y, d->pos_z, d->mass, n,
                     theta2, G, eps * eps, d->acc_x, d->acc_y, d->acc_z);
  NBX_LAUNCHED();
  mark(t, 9, s);
  return NBX_OK;
}

// The reference's insertion (force_barnes_hut.cu:368-389) creates a leaf in an
// empty child without checking its node budget; inputs exist (deep pairs whose
// chains use up the budget, then bodies in the chains' empty octants) where it
// writes past its 2 * max_particles nodes. The sequential builder stops there
// instead, and the synchronous queries report it.
static int check_truncated(const BhState& h) {
  if (h.truncated)
    return fail(NBX_ERR_CAPACITY,
                "Barnes-Hut node budget exhausted: %d of %d bodies left out of the tree "
                "(the reference's insertion writes past its node array for this input)",
                h.truncated, h.n);
  return NBX_OK;
}

int nbx_bh_node_count(nbx_bh* t, int32_t* node_count, int32_t* max_depth) {
  if (!t) return fail(NBX_ERR_INVALID, "nbx_bh_node_count: null tree");
  if (!t->built) {
    if (node_count) *node_count = 0;
    if (max_depth) *max_depth = 0;
    return NBX_OK;
  }
  BhState h;
  NBX_HIP(hipDeviceSynchronize());
  NBX_HIP(hipMemcpy(&h, t->st, sizeof(h), hipMemcpyDeviceToHost));
  NBX_TRY(check_truncated(h));
  if (node_count) *node_count = h.final_count;
  if (max_depth) *max_depth = h.max_depth;
  return NBX_OK;
}

int nbx_bh_copy_nodes(nbx_bh* t, void* host_nodes, size_t capacity_nodes) {
  if (!t || !host_nodes) return fail(NBX_ERR_INVALID, "nbx_bh_copy_nodes: null argument");
  int32_t cnt = 0;
  NBX_TRY(nbx_bh_node_count(t, &cnt, nullptr));
  if ((size_t)cnt > capacity_nodes)
    return fail(NBX_ERR_CAPACITY, "%d nodes do not fit in %zu", cnt, capacity_nodes);
  if (t->nodes_pending) {
    const int gn = grid_for((size_t)t->max_nodes);
    hipLaunchKernelGGL(node_keys, dim3(gn), dim3(kThreads), 0, 0, t->st, t->T, t->nkeys, t->nvals);
    NBX_LAUNCHED();
    NBX_TRY(t->sort_nodes.sort(t->nkeys, t->nvals, &t->st->sort_n,
                               key_bits_for(t->max_particles), 0));
    hipLaunchKernelGGL(invert_perm, dim3(gn), dim3(kThreads), 0, 0, t->st, t->nvals, t->final_of);
    hipLaunchKernelGGL(emit_nodes, dim3(gn), dim3(kThreads), 0, 0, t->st, t->T, t->final_of,
                       t->nodes);
    NBX_LAUNCHED();
    NBX_HIP(hipDeviceSynchronize());
    t->nodes_pending = false;
  }
  if (cnt) NBX_HIP(hipMemcpy(host_nodes, t->nodes, sizeof(RefNode) * cnt, hipMemcpyDeviceToHost));
  return NBX_OK;
}

int nbx_bh_copy_sort(nbx_bh* t, float bbox[6], uint32_t* codes, int32_t* indices, size_t n) {
  if (!t) return fail(NBX_ERR_INVALID, "nbx_bh_copy_sort: null tree");
  if (!t->built) return fail(NBX_ERR_STATE, "nbx_bh_copy_sort: no build yet");
  if ((codes || indices) && n != t->built_n)
    return fail(NBX_ERR_INVALID, "nbx_bh_copy_sort: n=%zu but the build had %zu", n, t->built_n);
  NBX_HIP(hipDeviceSynchronize());
  if (bbox) {
    BhState h;
    NBX_HIP(hipMemcpy(&h, t->st, sizeof(h), hipMemcpyDeviceToHost));
    for (int a = 0; a < 3; ++a) {
      bbox[a] = h.bmin[a];
      bbox[3 + a] = h.bmax[a];
    }
  }
  if (codes) NBX_HIP(hipMemcpy(codes, t->codes, 4 * n, hipMemcpyDeviceToHost));
  if (indices) NBX_HIP(hipMemcpy(indices, t->idx, 4 * n, hipMemcpyDeviceToHost));
  return NBX_OK;
}

int nbx_bh_build_info(nbx_bh* t, int32_t* builder, int32_t* depth32_leaves) {
  if (!t) return fail(NBX_ERR_INVALID, "nbx_bh_build_info: null tree");
  BhState h;
  memset(&h, 0, sizeof(h));
  if (t->built) {
    NBX_HIP(hipDeviceSynchronize());
    NBX_HIP(hipMemcpy(&h, t->st, sizeof(h), hipMemcpyDeviceToHost));
    NBX_TRY(check_truncated(h));
  }
  if (builder) *builder = h.use_seq;
  if (depth32_leaves) *depth32_leaves = h.use_seq ? h.depth_merges : h.multi_count;
  return NBX_OK;
}

int nbx_bh_walk_census(nbx_bh* t, const nbx_particles* d, float theta, float G, float eps,
                       int64_t counts[4]) {
  if (!t || !counts) return fail(NBX_ERR_INVALID, "nbx_bh_walk_census: null argument");
  NBX_TRY(check_particles(d, "nbx_bh_walk_census"));
  if (!t->built || t->built_n != d->count)
    return fail(NBX_ERR_STATE, "nbx_bh_walk_census: build the tree for these particles first");
  const int n = (int)d->count;
  NBX_TRY(ensure_prep(t, theta * theta, G, nullptr));
  const int ntasks = (int)div_up(n, 64);
  hipLaunchKernelGGL(census_clear, dim3(1), dim3(64), 0, nullptr, t->census);
  hipLaunchKernelGGL(bh_walk_census, dim3(std::min(ntasks, 8192)), dim3(64), 0, nullptr,
                     t->trav_rec, t->st, t->task_order, ntasks, d->pos_x, d->pos_y,
                     d->pos_z, d->mass, n, theta * theta, G, eps * eps, t->census);
  NBX_LAUNCHED();
  unsigned long long h[4];
  NBX_HIP(hipMemcpy(h, t->census, sizeof(h), hipMemcpyDeviceToHost));
  for (int k = 0; k < 4; ++k) counts[k] = (int64_t)h[k];
  return NBX_OK;
}

int nbx_bh_set_timing(nbx_bh* t, int enable) {
  if (!t) return fail(NBX_ERR_INVALID, "nbx_bh_set_timing: null tree");
  t->timing = enable != 0;
  return NBX_OK;
}

int nbx_bh_set_task_order(nbx_bh* t, int mode) {
  if (!t || mode < 0 || mode > 1)
    return fail(NBX_ERR_INVALID, "nbx_bh_set_task_order: bad argument");
  t->task_tree_order = mode;
  return NBX_OK;
}

int nbx_bh_set_builder(nbx_bh* t, int mode) {
  if (!t || mode < 0 || mode > 1) return fail(NBX_ERR_INVALID, "nbx_bh_set_builder: bad argument");
  t->force_seq = mode;
  return NBX_OK;
}

int nbx_bh_phase_ms(nbx_bh* t, double ms[8]) {
  if (!t || !ms) return fail(NBX_ERR_INVALID, "nbx_bh_phase_ms: null argument");
  if (!t->timing) return fail(NBX_ERR_STATE, "nbx_bh_phase_ms: timing is off");
  NBX_HIP(hipDeviceSynchronize());
  const int from[8] = {0, 1, 2, 3, 4, 5, 6, 8};
  const int to[8] = {1, 2, 3, 4, 5, 6, 7, 9};
  for (int k = 0; k < 8; ++k) {
    float v = 0.f;
    if (hipEventElapsedTime(&v, t->ev[from[k]], t->ev[to[k]]) != hipSuccess) {
      (void)hipGetLastError();
      v = -1.f;
    }
    ms[k] = v;
  }
  return NBX_OK;
}

}  // extern "C"
