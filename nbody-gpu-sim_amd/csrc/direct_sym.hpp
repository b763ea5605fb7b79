// Pair-symmetric direct N^2 forces (direct_sym.hip): each unordered pair is
// evaluated once and applied to both bodies. Used for whole-system force
// evaluations; target sub-ranges and batched ensembles use the one-sided
// kernel of direct_core.hpp.
#pragma once

#include <hip/hip_runtime.h>

#include "direct_core.hpp"
#include "nbx.h"

namespace nbx {

// Shapes X(IPL, W) of the symmetric kernel: IPL targets per lane, W waves per
// workgroup; super-blocks of 64 * IPL * W bodies.
#define NBX_SYM_VARIANTS(X) X(8, 4) X(4, 4) X(4, 8) X(8, 2) X(4, 2) X(2, 4) X(2, 2) X(4, 1)

// Bytes of FP32 partial sums a symmetric evaluation of n bodies needs with
// super-blocks of S bodies: (2 * floor(K/2) + 1) * 3 * K * S floats, K = ceil(n/S).
size_t direct_sym_partial_bytes(size_t n, int S);

// Largest partial buffer the symmetric path takes; above it (n of about a
// million bodies and more) whole-system calls use the one-sided kernel.
constexpr size_t kSymMaxPartialBytes = size_t(8) << 30;

// Forces on all d->count bodies (plus the optional kick), accelerations
// written to d->acc_*. Launches the tile kernel and the reduction on s.
template <int IPL, int W>
int direct_sym_launch(const nbx_particles* d, double G, float eps2, bool checked,
                      const KickArgs* kick, hipStream_t s);

}  // namespace nbx
