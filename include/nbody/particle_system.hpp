// One simulation run behind the reference's ParticleSystem interface
// (include/nbody/particle_system.hpp of build-workbench/nbody-gpu-sim), without
// the OpenGL interop: the same public methods, defaults and validation. The
// state lives on the GPU; every public method that changes the particle set
// recomputes a(0) so the next Verlet step has it. Not thread-safe.
//
// Beyond the reference: a precision form for the direct method
// (setPrecisionMode), in which positions are double-single hi + lo pairs and the
// pair sums are compensated; its state and checkpoints carry the low parts
// (.nbody version 2).
#pragma once

#include <memory>
#include <string>

#include "nbody/force_calculator.hpp"
#include "nbody/integrator.hpp"
#include "nbody/particle_data.hpp"
#include "nbody/simulation_state.hpp"
#include "nbody/types.hpp"

namespace nbody {

class ParticleSystem {
 public:
  ParticleSystem();
  ~ParticleSystem();
  ParticleSystem(const ParticleSystem&) = delete;
  ParticleSystem& operator=(const ParticleSystem&) = delete;
  ParticleSystem(ParticleSystem&&) = delete;
  ParticleSystem& operator=(ParticleSystem&&) = delete;

  void initialize(const SimulationConfig& config);
  void initializeWithDistribution(size_t particle_count, InitDistribution dist);

  void update(float dt);
  void pause();
  void resume();
  void reset();
  bool isPaused() const noexcept;

  void setForceMethod(ForceMethod method);
  void setGravitationalConstant(float G);
  void setSofteningParameter(float eps);
  void setTimeStep(float dt);
  void setBarnesHutTheta(float theta);
  void setSpatialHashCellSize(float size);
  void setSpatialHashCutoff(float cutoff);
  // Precision form (direct method only; ValidationException otherwise). Enabling
  // keeps the current positions (low parts zero); disabling rounds hi + lo to FP32.
  void setPrecisionMode(bool enabled);

  ForceMethod getForceMethod() const noexcept;
  float getGravitationalConstant() const noexcept;
  float getSofteningParameter() const noexcept;
  float getTimeStep() const noexcept;
  float getSimulationTime() const noexcept;
  size_t getParticleCount() const noexcept;
  bool getPrecisionMode() const noexcept;

  // Device arrays (the high parts of the positions in the precision form).
  ParticleData* getDeviceData();
  const ParticleData* getDeviceData() const;
  void copyToHost(ParticleData& h_particles) const;

  void saveState(const std::string& filename) const;
  void loadState(const std::string& filename);
  SimulationState getState() const;
  // Accepts plain state and precision-form state (low parts present, direct method).
  void setState(const SimulationState& state);

  float computeKineticEnergy() const;
  float computePotentialEnergy() const;
  float computeTotalEnergy() const;

 private:
  struct Run;
  std::unique_ptr<Run> run_;
};

}  // namespace nbody
