// Pair-symmetric direct N^2 forces (direct_sym.hip): each unordered pair is
// evaluated once and applied to both bodies. Used for whole-system force
// evaluations that bring a workspace (nbx_direct); target sub-ranges and
// calls without a workspace use the one-sided kernel of direct_core.hpp.
#pragma once

#include <hip/hip_runtime.h>

#include "direct_core.hpp"
#include "nbx.h"

namespace nbx {

// Shapes X(IPL, W) of the symmetric kernel: IPL targets per lane, W waves per
// workgroup; super-blocks of 64 * IPL * W bodies (256 .. 2,048). New shapes go
// at the end: the variant indices are part of the interface.
#define NBX_SYM_VARIANTS(X) X(8, 4) X(4, 4) X(4, 8) X(8, 2) X(4, 2) X(2, 4) X(2, 2) X(4, 1) X(6, 4) X(6, 2) X(6, 1) X(8, 1)

// Device scratch of a symmetric evaluation: FP32 partial sums of one round of
// tile distances, and the FP64 carry of every body's sum between rounds
// (3 doubles per padded body; unused when the evaluation is one round).
struct SymWork {
  float* part = nullptr;
  size_t part_bytes = 0;
  double* carry = nullptr;
  size_t carry_bytes = 0;
};

// Bytes of FP32 partial sums a symmetric evaluation of n bodies needs as one
// round with super-blocks of S: (2 * floor(K/2) + 1) * 3 * K * S floats, K = ceil(n/S).
size_t direct_sym_partial_bytes(size_t n, int S);
// Bytes of the FP64 carry for n bodies in super-blocks of S (24 per padded body).
size_t direct_sym_carry_bytes(size_t n, int S);
// Tile distances per round that part_bytes of partial scratch allows: floor(K/2)
// when the whole evaluation is one round, otherwise the largest R >= 1 whose
// 2R + 1 partial slots fit; -1 if not even R = 1 fits.
int direct_sym_round_len(size_t n, int S, size_t part_bytes);
// Rounds of an evaluation with round length R (1 for one round).
int direct_sym_rounds(size_t n, int S, int R);

// Forces on all d->count bodies (plus the optional kick), accelerations
// written to d->acc_*: per round, the tile kernel and the reduction on s.
// Fails with NBX_ERR_CAPACITY when w is too small (direct_sym_round_len < 0,
// or several rounds without a large enough carry).
template <int IPL, int W>
int direct_sym_launch(const nbx_particles* d, double G, float eps2, bool checked,
                      const KickArgs* kick, const SymWork& w, hipStream_t s);

}  // namespace nbx

// Workspace of the pair-symmetric kernels, owned by the caller through the C ABI
// (nbx_direct_create / nbx_direct_destroy): allocated once, never resized.
struct nbx_direct {
  int device = 0;
  size_t max_particles = 0;
  nbx::SymWork work;
};
