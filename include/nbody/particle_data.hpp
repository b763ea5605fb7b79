// Particle memory management and host-side initial conditions, mirroring the
// reference's include/nbody/particle_data.hpp (build-workbench/nbody-gpu-sim).
#pragma once

#include <random>

#include "nbody/types.hpp"

namespace nbody {

class ParticleDataManager {
 public:
  // 13 device arrays; accelerations and old accelerations zeroed.
  static void allocateDevice(ParticleData& data, size_t count);
  static void freeDevice(ParticleData& data);
  static void allocateHost(ParticleData& data, size_t count);
  static void freeHost(ParticleData& data);
  // Positions, velocities, accelerations and masses (not acc_old).
  static void copyToDevice(ParticleData& d_data, const ParticleData& h_data);
  static void copyToHost(ParticleData& h_data, const ParticleData& d_data);
};

// The reference's generators (std::mt19937 + uniform_real_distribution<float>).
class ParticleInitializer {
 public:
  static void initUniform(ParticleData& h_data, const UniformDistParams& params,
                          unsigned int seed = 42);
  static void initSpherical(ParticleData& h_data, const SphericalDistParams& params,
                            unsigned int seed = 42);
  static void initDisk(ParticleData& h_data, const DiskDistParams& params, unsigned int seed = 42);
  static void zeroVelocities(ParticleData& h_data);
  static void zeroAccelerations(ParticleData& h_data);

 private:
  static std::mt19937 createRNG(unsigned int seed);
};

}  // namespace nbody
