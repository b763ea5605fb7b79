// ParticleDataManager over the C ABI and the reference's host generators
// (reference: src/cuda/particle_init.cu:11-255).
#include <cstring>
#include <new>

#include "nbody/error_handling.hpp"
#include "nbody/particle_data.hpp"
#include "nbx.h"

namespace nbody {

static_assert(sizeof(ParticleData) == sizeof(nbx_particles), "ParticleData must match nbx_particles");

static nbx_particles* raw(ParticleData& d) { return reinterpret_cast<nbx_particles*>(&d); }
static const nbx_particles* raw(const ParticleData& d) {
  return reinterpret_cast<const nbx_particles*>(&d);
}

void ParticleDataManager::allocateDevice(ParticleData& data, size_t count) {
  data = ParticleData();
  NBODY_CHECK(nbx_particles_alloc(raw(data), count));
}

void ParticleDataManager::freeDevice(ParticleData& data) {
  NBODY_CHECK(nbx_particles_free(raw(data)));
  data = ParticleData();
}

static float** fields(ParticleData& d, int k) {
  float** f[13] = {&d.pos_x, &d.pos_y, &d.pos_z, &d.vel_x, &d.vel_y, &d.vel_z, &d.acc_x,
                   &d.acc_y, &d.acc_z, &d.acc_old_x, &d.acc_old_y, &d.acc_old_z, &d.mass};
  return f[k];
}

void ParticleDataManager::allocateHost(ParticleData& data, size_t count) {
  data = ParticleData();
  data.count = count;
  for (int k = 0; k < 13; ++k) {
    float* p = new (std::nothrow) float[count ? count : 1];
    if (!p) {
      freeHost(data);
      throw ResourceException("Failed to allocate host memory for particles",
                              count * 13 * sizeof(float), 0);
    }
    *fields(data, k) = p;
  }
}

void ParticleDataManager::freeHost(ParticleData& data) {
  for (int k = 0; k < 13; ++k) {
    delete[] *fields(data, k);
    *fields(data, k) = nullptr;
  }
  data.count = 0;
}

void ParticleDataManager::copyToDevice(ParticleData& d_data, const ParticleData& h_data) {
  NBODY_CHECK(nbx_particles_upload(raw(d_data), raw(h_data)));
}

void ParticleDataManager::copyToHost(ParticleData& h_data, const ParticleData& d_data) {
  NBODY_CHECK(nbx_particles_download(raw(h_data), raw(d_data)));
}

std::mt19937 ParticleInitializer::createRNG(unsigned int seed) { return std::mt19937(seed); }

void ParticleInitializer::initUniform(ParticleData& h, const UniformDistParams& p,
                                      unsigned int seed) {
  auto rng = createRNG(seed);
  std::uniform_real_distribution<float> dx(p.min_bounds.x, p.max_bounds.x);
  std::uniform_real_distribution<float> dy(p.min_bounds.y, p.max_bounds.y);
  std::uniform_real_distribution<float> dz(p.min_bounds.z, p.max_bounds.z);
  std::uniform_real_distribution<float> dm(p.min_mass, p.max_mass);
  for (size_t i = 0; i < h.count; i++) {
    h.pos_x[i] = dx(rng);
    h.pos_y[i] = dy(rng);
    h.pos_z[i] = dz(rng);
    h.vel_x[i] = h.vel_y[i] = h.vel_z[i] = 0.0f;
    h.mass[i] = dm(rng);
  }
  zeroAccelerations(h);
}

void ParticleInitializer::initSpherical(ParticleData& h, const SphericalDistParams& p,
                                        unsigned int seed) {
  auto rng = createRNG(seed);
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  std::uniform_real_distribution<float> dm(p.min_mass, p.max_mass);
  for (size_t i = 0; i < h.count; i++) {
    const float r = std::cbrt(u(rng)) * p.radius;
    const float theta = u(rng) * 2.0f * 3.14159265f;
    const float phi = std::acos(2.0f * u(rng) - 1.0f);
    h.pos_x[i] = p.center.x + r * std::sin(phi) * std::cos(theta);
    h.pos_y[i] = p.center.y + r * std::sin(phi) * std::sin(theta);
    h.pos_z[i] = p.center.z + r * std::cos(phi);
    h.vel_x[i] = h.vel_y[i] = h.vel_z[i] = 0.0f;
    h.mass[i] = dm(rng);
  }
  zeroAccelerations(h);
}

void ParticleInitializer::initDisk(ParticleData& h, const DiskDistParams& p, unsigned int seed) {
  auto rng = createRNG(seed);
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  std::uniform_real_distribution<float> dm(p.min_mass, p.max_mass);
  for (size_t i = 0; i < h.count; i++) {
    const float r = std::sqrt(u(rng)) * p.radius;
    const float theta = u(rng) * 2.0f * 3.14159265f;
    const float z = (u(rng) - 0.5f) * p.thickness;
    h.pos_x[i] = p.center.x + r * std::cos(theta);
    h.pos_y[i] = p.center.y + r * std::sin(theta);
    h.pos_z[i] = p.center.z + z;
    const float v = p.rotation_speed * std::sqrt(r);
    h.vel_x[i] = -v * std::sin(theta);
    h.vel_y[i] = v * std::cos(theta);
    h.vel_z[i] = 0.0f;
    h.mass[i] = dm(rng);
  }
  zeroAccelerations(h);
}

void ParticleInitializer::zeroVelocities(ParticleData& h) {
  for (size_t i = 0; i < h.count; i++) h.vel_x[i] = h.vel_y[i] = h.vel_z[i] = 0.0f;
}

void ParticleInitializer::zeroAccelerations(ParticleData& h) {
  for (size_t i = 0; i < h.count; i++) {
    h.acc_x[i] = h.acc_y[i] = h.acc_z[i] = 0.0f;
    h.acc_old_x[i] = h.acc_old_y[i] = h.acc_old_z[i] = 0.0f;
  }
}

}  // namespace nbody
