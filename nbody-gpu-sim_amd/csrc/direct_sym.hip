// Pair-symmetric direct N^2 forces: each unordered pair is evaluated once and
// applied to both bodies (a_i += m_j f r_ij, a_j -= m_i f r_ij), which halves
// the v_rsq_f32 work and most of the arithmetic of the one-sided kernel
// (direct_core.hpp). Same semantics as src/cuda/force_direct.cu:10-85.
//
// Decomposition. Bodies are grouped in super-blocks of S = 64 * IPL * W: a
// workgroup of W waves, IPL targets per lane. A tile (A, d) pairs super-block A
// with super-block C = (A + d) mod K (K super-blocks):
//   * d = 0 (diagonal): every ordered pair of A, targets only (one-sided);
//   * d = 1 .. D_A: each unordered pair of super-blocks once. D_A = floor(K/2),
//     except that for even K the tile at distance K/2 belongs to rows A < K/2.
// Each tile is one workgroup; diagonal tiles come first so the dispatcher
// starts the heavier ones early.
//
// Within a tile, each wave holds its IPL*64 targets in VGPRs and walks C in
// chunks of 64 bodies, one per lane. A chunk takes 64 steps: every lane
// evaluates its IPL pairs with the body it holds, then the chunk's positions
// and j-side accumulators rotate one lane (ds_bpermute, off the VALU). After
// 64 steps every body has met every target of the wave and sits in its home
// lane again. No atomics, no cross-lane reductions inside a chunk.
//
// Per chunk, the W waves' j-side sums are added in wave order through LDS and
// written as one FP32 partial; per tile, the i-side FP64 sums are written as
// one FP32 partial. sym_reduce_kernel sums every body's partials in a fixed
// order and applies G (and the Verlet kick), so results are deterministic.
//
// Rounds. The partials of a whole evaluation take (2 floor(K/2) + 1) * 3 * K * S
// floats (808 MB at 262,144 bodies, 13 GB at a million). When the workspace is
// smaller, the tiles run in rounds of consecutive distances: round 0 holds
// d = 0 .. R, each later round R more. After each round's tile kernel, a
// reduction adds the round's partials to every body's FP64 carry, in the same
// canonical order as one round would (sym_round_add), so the accelerations are
// bit-identical whatever R the workspace allows.
//
// Ragged N: bodies past N are padded at kPadPos with zero mass. A pair of a
// real body (|x| << kPadPos) and a pad has inv^3 below the FP32 range, so its
// contribution is an exact zero on both sides.
#include <algorithm>

#include "common.hpp"
#include "direct_sym.hpp"
#include "direct_sym_core.hpp"

namespace nbx {
namespace {

template <int IPL, int W, bool CHECKED>
__global__ __launch_bounds__(64 * W) void direct_sym_kernel(const float* __restrict__ x,
                                                            const float* __restrict__ y,
                                                            const float* __restrict__ z,
                                                            const float* __restrict__ m,
                                                            SymGeom g, float eps2) {
  __shared__ float red[2][W][3][64];
  int A, d;
  if (!sym_tile_of(g, (int)blockIdx.x, A, d)) return;  // whole workgroup: the other row's tile
  sym_run_tile<IPL, W, CHECKED>(x, y, z, m, g, A, d, eps2, red);
}

template <bool KICK>
__global__ __launch_bounds__(256) void sym_reduce_kernel(SymGeom g, double G, float* ax,
                                                         float* ay, float* az, KickArgs kick,
                                                         double* carry, int last) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.n) return;
  sym_reduce_body<KICK>(g, b, G, ax, ay, az, kick, carry, last != 0);
}

}  // namespace

size_t direct_sym_partial_bytes(size_t n, int S) { return sym_partial_floats(n, S) * sizeof(float); }

size_t direct_sym_carry_bytes(size_t n, int S) {
  const size_t K = (n + S - 1) / S;
  return 3 * K * (size_t)S * sizeof(double);
}

int direct_sym_round_len(size_t n, int S, size_t part_bytes) {
  const size_t K = (n + S - 1) / S;
  const size_t Dmax = K / 2;
  const size_t slots = part_bytes / (3 * K * (size_t)S * sizeof(float));
  if (slots >= 2 * Dmax + 1) return (int)Dmax;
  if (slots < 3) return -1;
  return (int)((slots - 1) / 2);
}

int direct_sym_rounds(size_t n, int S, int R) {
  const int Dmax = (int)(((n + S - 1) / S) / 2);
  if (R >= Dmax) return 1;
  return 1 + (Dmax - R + R - 1) / R;
}

template <int IPL, int W>
int direct_sym_launch(const nbx_particles* d, double G, float eps2, bool checked,
                      const KickArgs* kick, const SymWork& w, hipStream_t s) {
  constexpr int S = 64 * IPL * W;
  const int n = (int)d->count;
  const int R = direct_sym_round_len((size_t)n, S, w.part_bytes);
  const SymGeom g = sym_geom(n, S, w.part);
  if (R < 0 || !w.part)
    return fail(NBX_ERR_CAPACITY,
                "pair-symmetric scratch of %zu bytes is below the %zu needed for %d bodies",
                w.part_bytes, (size_t)3 * 3 * g.npad * sizeof(float), n);
  const bool multi = R < g.Dmax;
  if (multi && (!w.carry || w.carry_bytes < direct_sym_carry_bytes((size_t)n, S)))
    return fail(NBX_ERR_CAPACITY, "pair-symmetric evaluation of %d bodies in rounds needs a "
                "%zu-byte carry", n, direct_sym_carry_bytes((size_t)n, S));
  KickArgs k{};
  if (kick) k = *kick;
  const unsigned rb = div_up((size_t)n, 256);
  for (int lo = 0;;) {
    const int hi = std::min(g.Dmax, lo == 0 ? R : lo + R - 1);
    const bool last = hi == g.Dmax;
    const SymGeom gr = sym_round(g, lo, hi, w.part);
    const unsigned tiles = (unsigned)sym_tiles(gr);
    if (checked)
      hipLaunchKernelGGL((direct_sym_kernel<IPL, W, true>), dim3(tiles), dim3(64 * W), 0, s,
                         d->pos_x, d->pos_y, d->pos_z, d->mass, gr, eps2);
    else
      hipLaunchKernelGGL((direct_sym_kernel<IPL, W, false>), dim3(tiles), dim3(64 * W), 0, s,
                         d->pos_x, d->pos_y, d->pos_z, d->mass, gr, eps2);
    NBX_LAUNCHED();
    if (kick)
      hipLaunchKernelGGL(sym_reduce_kernel<true>, dim3(rb), dim3(256), 0, s, gr, G, d->acc_x,
                         d->acc_y, d->acc_z, k, w.carry, last ? 1 : 0);
    else
      hipLaunchKernelGGL(sym_reduce_kernel<false>, dim3(rb), dim3(256), 0, s, gr, G, d->acc_x,
                         d->acc_y, d->acc_z, k, w.carry, last ? 1 : 0);
    NBX_LAUNCHED();
    if (last) break;
    lo = hi + 1;
  }
  return NBX_OK;
}

#define NBX_SYM_INST(I, W)                                                                  \
  template int direct_sym_launch<I, W>(const nbx_particles*, double, float, bool,          \
                                       const KickArgs*, const SymWork&, hipStream_t);
NBX_SYM_VARIANTS(NBX_SYM_INST)
#undef NBX_SYM_INST

}  // namespace nbx
