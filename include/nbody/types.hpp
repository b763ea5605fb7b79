// Data model of the engine's C++ API: the names, layouts and defaults of the
// reference's include/nbody/types.hpp (build-workbench/nbody-gpu-sim), so code
// written against that header compiles against this one.
#pragma once

#include <cmath>
#include <cstddef>

// Integer 3-vector of the cell coordinates (SpatialHashGrid::getCellIndex): HIP's
// own int3 when the including file is compiled by hipcc, a plain struct of the
// same layout otherwise (the reference does the same with CUDA's, types.hpp:6-21).
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else
struct int3 {
  int x, y, z;
};
inline constexpr int3 make_int3(int x, int y, int z) { return {x, y, z}; }
#endif

namespace nbody {

struct ParticleData;
class ForceCalculator;
class Integrator;
class BarnesHutTree;
class SpatialHashGrid;
class ParticleSystem;

// Values 0/1/2 are part of the .nbody checkpoint format (serialization.cpp:52-55
// of the reference).
enum class ForceMethod { DIRECT_N2, BARNES_HUT, SPATIAL_HASH };

inline const char* forceMethodToString(ForceMethod method) {
  switch (method) {
    case ForceMethod::DIRECT_N2: return "direct_n2";
    case ForceMethod::BARNES_HUT: return "barnes_hut";
    case ForceMethod::SPATIAL_HASH: return "spatial_hash";
  }
  return "unknown";
}

enum class InitDistribution { UNIFORM, SPHERICAL, DISK };

// Float 3-vector with the reference's evaluation order (dot = (x*x + y*y) + z*z,
// division by s is a true division).
struct Vec3 {
  float x, y, z;
  Vec3() : x(0), y(0), z(0) {}
  Vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
  Vec3 operator+(const Vec3& v) const { return Vec3(x + v.x, y + v.y, z + v.z); }
  Vec3 operator-(const Vec3& v) const { return Vec3(x - v.x, y - v.y, z - v.z); }
  Vec3 operator*(float s) const { return Vec3(x * s, y * s, z * s); }
  Vec3 operator/(float s) const { return Vec3(x / s, y / s, z / s); }
  Vec3& operator+=(const Vec3& v) {
    x += v.x;
    y += v.y;
    z += v.z;
    return *this;
  }
  Vec3& operator-=(const Vec3& v) {
    x -= v.x;
    y -= v.y;
    z -= v.z;
    return *this;
  }
  float dot(const Vec3& v) const { return x * v.x + y * v.y + z * v.z; }
  float length2() const { return dot(*this); }
  float length() const { return std::sqrt(length2()); }
  Vec3 normalized() const {
    const float l = length();
    return l > 0 ? *this / l : Vec3();
  }
};

inline Vec3 operator*(float s, const Vec3& v) { return v * s; }

// SoA particle set: 13 FP32 arrays (52 B per particle). Device instances hold
// HIP device pointers; the layout is identical to nbx_particles (include/nbx.h).
struct ParticleData {
  float* pos_x = nullptr;
  float* pos_y = nullptr;
  float* pos_z = nullptr;
  float* vel_x = nullptr;
  float* vel_y = nullptr;
  float* vel_z = nullptr;
  float* acc_x = nullptr;
  float* acc_y = nullptr;
  float* acc_z = nullptr;
  float* acc_old_x = nullptr;
  float* acc_old_y = nullptr;
  float* acc_old_z = nullptr;
  float* mass = nullptr;
  size_t count = 0;
};

struct SimulationConfig {
  size_t particle_count = 10000;
  InitDistribution init_distribution = InitDistribution::SPHERICAL;
  ForceMethod force_method = ForceMethod::DIRECT_N2;
  float dt = 0.001f;
  float G = 1.0f;
  float softening = 0.1f;
  float barnes_hut_theta = 0.5f;
  float spatial_hash_cell_size = 2.0f;
  float spatial_hash_cutoff = 2.0f;
  int cuda_block_size = 256;  // name kept for source compatibility: workgroup size hint
};

struct UniformDistParams {
  Vec3 min_bounds;
  Vec3 max_bounds;
  float min_mass = 1.0f;
  float max_mass = 1.0f;
};

struct SphericalDistParams {
  Vec3 center;
  float radius = 10.0f;
  float min_mass = 1.0f;
  float max_mass = 1.0f;
};

struct DiskDistParams {
  Vec3 center;
  float radius = 10.0f;
  float thickness = 1.0f;
  float min_mass = 1.0f;
  float max_mass = 1.0f;
  float rotation_speed = 1.0f;
};

}  // namespace nbody
