// Checkpointable simulation state, mirroring the reference's
// include/nbody/simulation_state.hpp (build-workbench/nbody-gpu-sim).
#pragma once

#include <iosfwd>
#include <vector>

#include "nbody/types.hpp"

namespace nbody {

struct SimulationState {
  std::vector<float> pos_x, pos_y, pos_z;
  std::vector<float> vel_x, vel_y, vel_z;
  std::vector<float> mass;
  // Precision form only: low parts of double-single positions (x = pos_x + pos_lo_x).
  // Empty for plain FP32 state; non-empty state is written as .nbody version 2.
  std::vector<float> pos_lo_x, pos_lo_y, pos_lo_z;
  size_t particle_count = 0;
  float simulation_time = 0.0f;
  float dt = 0.0f;
  float G = 0.0f;
  float softening = 0.0f;
  ForceMethod force_method = ForceMethod::DIRECT_N2;

  void serialize(std::ostream& out) const;
  static SimulationState deserialize(std::istream& in);
  bool hasLowParts() const noexcept { return !pos_lo_x.empty(); }
  // Element-wise within 1e-6 (the reference's operator==); low parts must match exactly.
  bool operator==(const SimulationState& other) const;
};

}  // namespace nbody
