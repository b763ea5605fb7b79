// Internal launch helpers of the Velocity Verlet kernels (verlet.hip).
#pragma once

#include "common.hpp"

namespace nbx {

// storeAccelerations (if store) + updatePositions in one launch.
int verlet_drift_launch(nbx_particles* d, float dt, bool store, hipStream_t s);
int verlet_kick_launch(nbx_particles* d, float dt, hipStream_t s);

}  // namespace nbx

namespace nbx {
// FP64 KE / momentum / PE; lo arrays (may be null) hold double-single position tails.
int energy_launch(const nbx_particles* d, const float* xlo, const float* ylo, const float* zlo,
                  float G, float eps2, double* ke, double* pe, double* mom, hipStream_t s);
}  // namespace nbx
