// Status-to-exception mapping and the input validators of the C++ API
// (reference: src/utils/error_handling.cpp:26-138).
#include <cmath>

#include "nbody/error_handling.hpp"
#include "nbody/serialization.hpp"
#include "nbody/types.hpp"
#include "nbx.h"

namespace nbody {

void throwOnStatus(int status, const char* file, int line) {
  if (status == NBX_OK) return;
  const std::string msg = nbx_last_error();
  switch (status) {
    case NBX_ERR_INVALID:
    case NBX_ERR_STATE:
      throw ValidationException(msg);
    case NBX_ERR_NOMEM:
    case NBX_ERR_CAPACITY:
      throw ResourceException(msg, 0, 0);
    default:
      throw DeviceException(msg, file, line);
  }
}

void validateResourceRequirements(size_t particle_count) {
  size_t total = 0;
  NBODY_CHECK(nbx_device_info(nullptr, 0, nullptr, nullptr, &total));
  // 13 floats per particle, doubled for acceleration structures (:31-38)
  const size_t required = particle_count * sizeof(float) * 13 * 2;
  if ((double)required > (double)total * 0.8)
    throw ResourceException("Insufficient GPU memory", required, (size_t)((double)total * 0.8));
}

void validateSimulationConfig(const SimulationConfig& config) {
  validateParticleCountRange(config.particle_count);
  validateTimeStep(config.dt);
  validateSoftening(config.softening);
  if (config.force_method == ForceMethod::BARNES_HUT) validateTheta(config.barnes_hut_theta);
  if (config.G <= 0 || std::isnan(config.G) || std::isinf(config.G))
    throw ValidationException("Gravitational constant must be positive and finite");
  if (config.force_method == ForceMethod::SPATIAL_HASH) {
    if (config.spatial_hash_cell_size <= 0 || std::isnan(config.spatial_hash_cell_size) ||
        std::isinf(config.spatial_hash_cell_size))
      throw ValidationException("Spatial hash cell size must be positive and finite");
    if (config.spatial_hash_cutoff <= 0 || std::isnan(config.spatial_hash_cutoff) ||
        std::isinf(config.spatial_hash_cutoff))
      throw ValidationException("Spatial hash cutoff must be positive and finite");
    if (config.spatial_hash_cutoff > config.spatial_hash_cell_size)
      throw ValidationException("Spatial hash cutoff must not exceed cell size");
  }
  if (config.cuda_block_size <= 0 || config.cuda_block_size > 1024)
    throw ValidationException("CUDA block size must be between 1 and 1024");
  if ((config.cuda_block_size & (config.cuda_block_size - 1)) != 0)
    throw ValidationException("CUDA block size must be a power of two");
}

void validateParticleCountRange(size_t count) {
  if (count == 0) throw ValidationException("Particle count must be greater than 0");
  if (count > MAX_PARTICLE_COUNT)
    throw ValidationException("Particle count exceeds maximum supported (100M)");
}

void validateParticleCount(size_t count) {
  validateParticleCountRange(count);
  validateResourceRequirements(count);
}

void validateTimeStep(float dt) {
  if (dt <= 0) throw ValidationException("Time step must be positive");
  if (std::isnan(dt) || std::isinf(dt)) throw ValidationException("Time step must be a finite number");
  if (dt > 1.0f) throw ValidationException("Time step is too large (max 1.0)");
}

void validateSoftening(float eps) {
  if (std::isnan(eps) || std::isinf(eps))
    throw ValidationException("Softening parameter must be a finite number");
  if (eps <= 0) throw ValidationException("Softening parameter must be positive");
}

void validateTheta(float theta) {
  if (theta < 0 || theta > 2.0f) throw ValidationException("Barnes-Hut theta must be between 0 and 2");
  if (std::isnan(theta) || std::isinf(theta))
    throw ValidationException("Barnes-Hut theta must be a finite number");
}

}  // namespace nbody
