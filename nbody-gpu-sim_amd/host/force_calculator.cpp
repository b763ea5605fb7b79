// Force operators over the C ABI (reference: src/cuda/force_direct.cu:100-117,
// src/cuda/force_barnes_hut.cu:573-583, src/cuda/force_spatial_hash.cu:370-404).
#include <algorithm>
#include <cmath>
#include <string>

#include "nbody/barnes_hut_tree.hpp"
#include "nbody/error_handling.hpp"
#include "nbody/force_calculator.hpp"
#include "nbody/performance_observability.hpp"
#include "nbody/spatial_hash_grid.hpp"
#include "nbx.h"

namespace nbody {

static nbx_particles* raw(ParticleData* d) { return reinterpret_cast<nbx_particles*>(d); }

DirectForceCalculator::DirectForceCalculator(int block_size) : block_size_(block_size) {}

DirectForceCalculator::~DirectForceCalculator() { (void)nbx_direct_destroy(workspace_); }

void DirectForceCalculator::reserve(size_t max_particles) {
  if (workspace_ && max_particles <= capacity_) return;
  nbx_direct* ws = nullptr;
  NBODY_CHECK(nbx_direct_create(max_particles, 0, &ws));
  (void)nbx_direct_destroy(workspace_);
  workspace_ = ws;
  capacity_ = max_particles;
}

// The block size is checked as launchDirectForceKernel's launch would reject it
// (force_direct.cu:88-98); the tile shape itself comes from the autotuner.
static void check_block_size(int b) {
  if (b < 1 || b > 1024)
    throw ValidationException("block size must be between 1 and 1024 (got " +
                              std::to_string(b) + ")");
}

void DirectForceCalculator::computeForces(ParticleData* d_particles) {
  NBODY_PROFILE_SCOPE("force.direct_n2");
  check_block_size(block_size_);
  reserve(d_particles->count);
  NBODY_CHECK(nbx_direct_compute(workspace_, raw(d_particles), G_, softening_eps2_, nullptr));
}

void DirectForceCalculator::computeForcesKick(ParticleData* d_particles, float dt) {
  NBODY_PROFILE_SCOPE("force.direct_n2");
  check_block_size(block_size_);
  reserve(d_particles->count);
  NBODY_CHECK(
      nbx_direct_compute_kick(workspace_, raw(d_particles), G_, softening_eps2_, dt, nullptr));
}

BarnesHutCalculator::BarnesHutCalculator(float theta) : theta_(theta) {}
BarnesHutCalculator::~BarnesHutCalculator() = default;

void BarnesHutCalculator::computeForces(ParticleData* d_particles) {
  if (!tree_) tree_ = std::make_unique<BarnesHutTree>(d_particles->count);
  tree_->build(d_particles);
  tree_->computeForces(d_particles, theta_, G_, softening_eps_);
}

SpatialHashCalculator::SpatialHashCalculator(float cell_size, float cutoff_radius)
    : cell_size_(cell_size), cutoff_radius_(cutoff_radius) {}
SpatialHashCalculator::~SpatialHashCalculator() = default;

// :376-388. The grid is remade when the effective cell changes, and also when the
// particle count outgrows it (the reference keeps the first count's buffers).
void SpatialHashCalculator::computeForces(ParticleData* d_particles) {
  NBODY_PROFILE_SCOPE("force.spatial_hash");
  const float effective_cell = std::max(cell_size_, cutoff_radius_);
  if (!grid_ || grid_->getCellSize() != effective_cell ||
      d_particles->count > grid_->getMaxParticles())
    grid_ = std::make_unique<SpatialHashGrid>(d_particles->count, effective_cell);
  grid_->build(d_particles);
  grid_->computeForces(d_particles, cutoff_radius_, G_, softening_eps_);
}

std::unique_ptr<ForceCalculator> createForceCalculator(ForceMethod method,
                                                       const SimulationConfig& config) {
  switch (method) {
    case ForceMethod::DIRECT_N2:
      return std::make_unique<DirectForceCalculator>(config.cuda_block_size);
    case ForceMethod::BARNES_HUT:
      return std::make_unique<BarnesHutCalculator>(config.barnes_hut_theta);
    case ForceMethod::SPATIAL_HASH:
      return std::make_unique<SpatialHashCalculator>(config.spatial_hash_cell_size,
                                                     config.spatial_hash_cutoff);
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
