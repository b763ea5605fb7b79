// Integrator over the C ABI (reference: src/cuda/integrator.cu:123-244).
#include "nbody/error_handling.hpp"
#include "nbody/force_calculator.hpp"
#include "nbody/integrator.hpp"
#include "nbody/performance_observability.hpp"
#include "nbx.h"

namespace nbody {

static nbx_particles* raw(ParticleData* d) { return reinterpret_cast<nbx_particles*>(d); }
static const nbx_particles* raw(const ParticleData* d) {
  return reinterpret_cast<const nbx_particles*>(d);
}

void launchUpdatePositionsKernel(ParticleData* d, float dt, int) {
  NBODY_CHECK(nbx_update_positions(raw(d), dt, nullptr));
}
void launchUpdateVelocitiesKernel(ParticleData* d, float dt, int) {
  NBODY_CHECK(nbx_update_velocities(raw(d), dt, nullptr));
}
void launchStoreAccelerationsKernel(ParticleData* d, int) {
  NBODY_CHECK(nbx_store_old_accelerations(raw(d), nullptr));
}

Integrator::Integrator(int block_size) : block_size_(block_size) {}
Integrator::~Integrator() = default;

void Integrator::integrate(ParticleData* d, ForceCalculator* force_calc, float dt) {
  NBODY_PROFILE_SCOPE("integrator.velocity_verlet");
  if (auto* direct = dynamic_cast<DirectForceCalculator*>(force_calc)) {
    // same four reference kernels, fused into two launches (drift; forces + kick)
    NBODY_CHECK(nbx_verlet_drift(raw(d), dt, nullptr));
    direct->computeForcesKick(d, dt);
    return;
  }
  NBODY_CHECK(nbx_verlet_drift(raw(d), dt, nullptr));
  force_calc->computeForces(d);
  NBODY_CHECK(nbx_update_velocities(raw(d), dt, nullptr));
}

void Integrator::updatePositions(ParticleData* d, float dt) {
  launchUpdatePositionsKernel(d, dt, block_size_);
}
void Integrator::updateVelocities(ParticleData* d, float dt) {
  launchUpdateVelocitiesKernel(d, dt, block_size_);
}
void Integrator::storeOldAccelerations(ParticleData* d) {
  launchStoreAccelerationsKernel(d, block_size_);
}

EnergyF64 Integrator::computeEnergyF64(const ParticleData* d, float G, float eps, bool potential) {
  EnergyF64 e;
  NBODY_CHECK(nbx_energy(raw(d), G, eps * eps, &e.kinetic, potential ? &e.potential : nullptr,
                         e.momentum, nullptr));
  return e;
}

float Integrator::computeKineticEnergy(const ParticleData* d) {
  return (float)computeEnergyF64(d, 1.0f, 1.0f, false).kinetic;
}
float Integrator::computePotentialEnergy(const ParticleData* d, float G, float eps) {
  return (float)computeEnergyF64(d, G, eps, true).potential;
}
float Integrator::computeTotalEnergy(const ParticleData* d, float G, float eps) {
  const EnergyF64 e = computeEnergyF64(d, G, eps, true);
  return (float)(e.kinetic + e.potential);
}

}  // namespace nbody
