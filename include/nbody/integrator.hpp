// Velocity Verlet integrator of the engine's C++ API, mirroring the reference's
// include/nbody/integrator.hpp (build-workbench/nbody-gpu-sim). Energies are
// reduced in FP64 on the GPU and returned rounded to float, as the reference's
// signatures require; computeEnergyF64 exposes the full values.
#pragma once

#include "nbody/types.hpp"

namespace nbody {

struct EnergyF64 {
  double kinetic = 0.0;
  double potential = 0.0;
  double momentum[3] = {0.0, 0.0, 0.0};
};

class Integrator {
 public:
  explicit Integrator(int block_size = 256);
  ~Integrator();
  Integrator(const Integrator&) = delete;
  Integrator& operator=(const Integrator&) = delete;
  Integrator(Integrator&&) = delete;
  Integrator& operator=(Integrator&&) = delete;

  // store a -> a_old; x += v dt + a dt^2/2; forces; v += (a_old + a) dt/2
  void integrate(ParticleData* d_particles, ForceCalculator* force_calc, float dt);
  void updatePositions(ParticleData* d_particles, float dt);
  void updateVelocities(ParticleData* d_particles, float dt);
  void storeOldAccelerations(ParticleData* d_particles);

  float computeKineticEnergy(const ParticleData* d_particles);
  float computePotentialEnergy(const ParticleData* d_particles, float G, float eps);
  float computeTotalEnergy(const ParticleData* d_particles, float G, float eps);
  EnergyF64 computeEnergyF64(const ParticleData* d_particles, float G, float eps,
                             bool potential = true);

  void ensureScratchBuffer(size_t) {}  // energies use the library's own scratch
  void setBlockSize(int size) { block_size_ = size; }
  int getBlockSize() const noexcept { return block_size_; }

 private:
  int block_size_;
};

void launchUpdatePositionsKernel(ParticleData* d_particles, float dt, int block_size);
void launchUpdateVelocitiesKernel(ParticleData* d_particles, float dt, int block_size);
void launchStoreAccelerationsKernel(ParticleData* d_particles, int block_size);

}  // namespace nbody
