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
  size_t particle_count = 0;
  float simulation_time = 0.0f;
  float dt = 0.0f;
  float G = 0.0f;
  float softening = 0.0f;
  ForceMethod force_method = ForceMethod::DIRECT_N2;

  void serialize(std::ostream& out) const;
  static SimulationState deserialize(std::istream& in);
  // Element-wise within 1e-6 (the reference's operator==).
  bool operator==(const SimulationState& other) const;
};

}  // namespace nbody
