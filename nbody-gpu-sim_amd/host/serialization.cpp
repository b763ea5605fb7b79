// .nbody reader/writer: version 1 as the reference (src/utils/serialization.cpp:9-145,
// src/utils/simulation_state.cpp:7-60), version 2 for precision-form state.
#include <cmath>
#include <cstring>
#include <fstream>

#include "nbody/performance_observability.hpp"
#include "nbody/serialization.hpp"

namespace nbody {

void Serializer::save(const std::string& filename, const SimulationState& state) {
  std::ofstream file(filename, std::ios::binary);
  if (!file) throw IOException("Failed to open file for writing: " + filename);
  save(file, state);
  if (!file) throw IOException("Failed to write checkpoint: " + filename);
}

SimulationState Serializer::load(const std::string& filename) {
  std::ifstream file(filename, std::ios::binary);
  if (!file) throw IOException("Failed to open file for reading: " + filename);
  return load(file);
}

void Serializer::save(std::ostream& out, const SimulationState& state) {
  NBODY_PROFILE_SCOPE("serialization.save");
  if (state.hasLowParts() &&
      (state.pos_lo_x.size() != state.particle_count || state.pos_lo_y.size() != state.particle_count ||
       state.pos_lo_z.size() != state.particle_count))
    throw ValidationException("SimulationState low-part vectors do not match particle_count");
  writeHeader(out, state);
  for (const auto* a : {&state.pos_x, &state.pos_y, &state.pos_z, &state.vel_x, &state.vel_y,
                        &state.vel_z, &state.mass})
    writeFloatArray(out, *a);
  if (state.hasLowParts())
    for (const auto* a : {&state.pos_lo_x, &state.pos_lo_y, &state.pos_lo_z}) writeFloatArray(out, *a);
}

SimulationState Serializer::load(std::istream& in) {
  NBODY_PROFILE_SCOPE("serialization.load");
  const FileHeader header = readHeader(in);
  if (header.particle_count > MAX_PARTICLE_COUNT)
    throw ValidationException("Particle count (" + std::to_string(header.particle_count) +
                              ") exceeds maximum allowed (" + std::to_string(MAX_PARTICLE_COUNT) +
                              ")");
  if (header.force_method > static_cast<uint32_t>(ForceMethod::SPATIAL_HASH))
    throw ValidationException("Invalid force method in file header: " +
                              std::to_string(header.force_method));
  SimulationState st;
  st.particle_count = header.particle_count;
  st.simulation_time = header.simulation_time;
  st.dt = header.dt;
  st.G = header.G;
  st.softening = header.softening;
  st.force_method = static_cast<ForceMethod>(header.force_method);
  st.pos_x = readFloatArray(in, st.particle_count);
  st.pos_y = readFloatArray(in, st.particle_count);
  st.pos_z = readFloatArray(in, st.particle_count);
  st.vel_x = readFloatArray(in, st.particle_count);
  st.vel_y = readFloatArray(in, st.particle_count);
  st.vel_z = readFloatArray(in, st.particle_count);
  st.mass = readFloatArray(in, st.particle_count);
  if (header.version == NBODY_VERSION_DOUBLE_SINGLE) {
    st.pos_lo_x = readFloatArray(in, st.particle_count);
    st.pos_lo_y = readFloatArray(in, st.particle_count);
    st.pos_lo_z = readFloatArray(in, st.particle_count);
  }
  return st;
}

bool Serializer::validateFile(const std::string& filename) {
  std::ifstream file(filename, std::ios::binary);
  if (!file) return false;
  return validateStream(file);
}

bool Serializer::validateStream(std::istream& in) {
  try {
    const FileHeader h = readHeader(in);
    return h.magic == NBODY_MAGIC &&
           (h.version == NBODY_VERSION || h.version == NBODY_VERSION_DOUBLE_SINGLE);
  } catch (const std::exception&) {
    return false;
  }
}

void Serializer::writeHeader(std::ostream& out, const SimulationState& state) {
  FileHeader h;
  std::memset(&h, 0, sizeof(h));  // deterministic bytes, padding included
  h.magic = NBODY_MAGIC;
  h.version = state.hasLowParts() ? NBODY_VERSION_DOUBLE_SINGLE : NBODY_VERSION;
  h.particle_count = state.particle_count;
  h.simulation_time = state.simulation_time;
  h.dt = state.dt;
  h.G = state.G;
  h.softening = state.softening;
  h.force_method = static_cast<uint32_t>(state.force_method);
  out.write(reinterpret_cast<const char*>(&h), sizeof(h));
}

FileHeader Serializer::readHeader(std::istream& in) {
  FileHeader h;
  in.read(reinterpret_cast<char*>(&h), sizeof(h));
  if (!in || in.gcount() != sizeof(h))
    throw ValidationException("Failed to read file header: file may be truncated or corrupted");
  if (h.magic != NBODY_MAGIC) throw ValidationException("Invalid file format: wrong magic number");
  if (h.version != NBODY_VERSION && h.version != NBODY_VERSION_DOUBLE_SINGLE)
    throw ValidationException("Unsupported file version");
  return h;
}

void Serializer::writeFloatArray(std::ostream& out, const std::vector<float>& data) {
  out.write(reinterpret_cast<const char*>(data.data()),
            static_cast<std::streamsize>(data.size() * sizeof(float)));
}

std::vector<float> Serializer::readFloatArray(std::istream& in, size_t count) {
  std::vector<float> data(count);
  in.read(reinterpret_cast<char*>(data.data()), static_cast<std::streamsize>(count * sizeof(float)));
  if (!in || in.gcount() != static_cast<std::streamsize>(count * sizeof(float)))
    throw ValidationException("Failed to read particle data: file may be truncated or corrupted");
  return data;
}

bool SimulationState::operator==(const SimulationState& o) const {
  auto near = [](float a, float b) { return std::abs(a - b) <= 1e-6f; };
  if (particle_count != o.particle_count || !near(simulation_time, o.simulation_time) ||
      !near(dt, o.dt) || !near(G, o.G) || !near(softening, o.softening) ||
      force_method != o.force_method)
    return false;
  const std::vector<float>* a[7] = {&pos_x, &pos_y, &pos_z, &vel_x, &vel_y, &vel_z, &mass};
  const std::vector<float>* b[7] = {&o.pos_x, &o.pos_y, &o.pos_z, &o.vel_x,
                                    &o.vel_y, &o.vel_z, &o.mass};
  for (int k = 0; k < 7; ++k)
    if (a[k]->size() != particle_count || b[k]->size() != particle_count) return false;
  for (int k = 0; k < 7; ++k)
    for (size_t i = 0; i < particle_count; ++i)
      if (!near((*a[k])[i], (*b[k])[i])) return false;
  return pos_lo_x == o.pos_lo_x && pos_lo_y == o.pos_lo_y && pos_lo_z == o.pos_lo_z;
}

void SimulationState::serialize(std::ostream& out) const { Serializer::save(out, *this); }
SimulationState SimulationState::deserialize(std::istream& in) { return Serializer::load(in); }

}  // namespace nbody
