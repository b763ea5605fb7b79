// Force operators over the C ABI (reference: src/cuda/force_direct.cu:100-117,
// src/cuda/force_barnes_hut.cu:573-583, factory force_spatial_hash.cu:391-404).
#include <cmath>

#include "nbody/barnes_hut_tree.hpp"
#include "nbody/error_handling.hpp"
#include "nbody/force_calculator.hpp"
#include "nbx.h"

namespace nbody {

static nbx_particles* raw(ParticleData* d) { return reinterpret_cast<nbx_particles*>(d); }

DirectForceCalculator::DirectForceCalculator(int block_size) : block_size_(block_size) {}

void DirectForceCalculator::computeForces(ParticleData* d_particles) {
  NBODY_CHECK(nbx_direct_forces(raw(d_particles), G_, softening_eps2_, block_size_, nullptr));
}

BarnesHutCalculator::BarnesHutCalculator(float theta) : theta_(theta) {}
BarnesHutCalculator::~BarnesHutCalculator() = default;

void BarnesHutCalculator::computeForces(ParticleData* d_particles) {
  if (!tree_) tree_ = std::make_unique<BarnesHutTree>(d_particles->count);
  tree_->build(d_particles);
  tree_->computeForces(d_particles, theta_, G_, softening_eps_);
}

std::unique_ptr<ForceCalculator> createForceCalculator(ForceMethod method,
                                                       const SimulationConfig& config) {
  switch (method) {
    case ForceMethod::DIRECT_N2:
      return std::make_unique<DirectForceCalculator>(config.cuda_block_size);
    case ForceMethod::BARNES_HUT:
      return std::make_unique<BarnesHutCalculator>(config.barnes_hut_theta);
    case ForceMethod::SPATIAL_HASH:
      throw ValidationException("The spatial-hash force method is not built in this engine");
  }
  throw ValidationException("Unknown force method");
}

Vec3 computeGravitationalAccelerationCPU(const Vec3& p1, const Vec3& p2, float m2, float G,
                                         float eps) {
  const Vec3 r = p2 - p1;
  const float dist2 = r.length2() + eps * eps;
  const float inv = 1.0f / std::sqrt(dist2);
  const float inv3 = inv * inv * inv;
  const float f = G * m2 * inv3;
  return r * f;
}

}  // namespace nbody
