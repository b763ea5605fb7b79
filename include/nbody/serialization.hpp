// The .nbody checkpoint format. Version 1 is byte-compatible with the
// reference's include/nbody/serialization.hpp (build-workbench/nbody-gpu-sim):
// a 56-byte little-endian header followed by pos_x, pos_y, pos_z, vel_x, vel_y,
// vel_z, mass as FP32 arrays. Version 2 (precision-form state) has the same
// header with version = 2 and the same seven arrays, followed by pos_lo_x,
// pos_lo_y, pos_lo_z: a version-1 reader, which accepts only version 1, rejects
// it with "Unsupported file version" instead of misreading it.
#pragma once

#include <cstdint>
#include <iosfwd>
#include <string>
#include <vector>

#include "nbody/error_handling.hpp"
#include "nbody/simulation_state.hpp"

namespace nbody {

constexpr uint32_t NBODY_MAGIC = 0x4E424F44;
constexpr uint32_t NBODY_VERSION = 1;
constexpr uint32_t NBODY_VERSION_DOUBLE_SINGLE = 2;
constexpr uint64_t MAX_PARTICLE_COUNT = 100'000'000;

struct FileHeader {
  uint32_t magic;
  uint32_t version;
  uint64_t particle_count;
  float simulation_time;
  float dt;
  float G;
  float softening;
  uint32_t force_method;
  uint32_t reserved[4];
};
static_assert(sizeof(FileHeader) == 56, "FileHeader layout is part of the file format");

class Serializer {
 public:
  static void save(const std::string& filename, const SimulationState& state);
  static SimulationState load(const std::string& filename);
  static void save(std::ostream& out, const SimulationState& state);
  static SimulationState load(std::istream& in);
  // True for a readable header of version 1 or 2.
  static bool validateFile(const std::string& filename);
  static bool validateStream(std::istream& in);

 private:
  static void writeHeader(std::ostream& out, const SimulationState& state);
  static FileHeader readHeader(std::istream& in);
  static void writeFloatArray(std::ostream& out, const std::vector<float>& data);
  static std::vector<float> readFloatArray(std::istream& in, size_t count);
};

}  // namespace nbody
