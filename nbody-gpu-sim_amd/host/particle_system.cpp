// ParticleSystem (reference: src/core/particle_system.cpp:12-330, without the
// OpenGL interop): owns device and host buffers, the force operator and the
// integrator; a(0) is computed on (re)initialisation so the first Verlet step
// has it.
#include <algorithm>
#include <cmath>

#include "nbody/error_handling.hpp"
#include "nbody/particle_system.hpp"
#include "nbody/performance_observability.hpp"
#include "nbody/serialization.hpp"

namespace nbody {

ParticleSystem::ParticleSystem()
    : particle_count_(0), dt_(0.001f), G_(1.0f), softening_(0.1f), simulation_time_(0.0f),
      force_method_(ForceMethod::DIRECT_N2), is_paused_(false), is_initialized_(false) {}

ParticleSystem::~ParticleSystem() {
  try {
    freeMemory();
  } catch (...) {
  }
}

void ParticleSystem::allocateMemory(size_t count) {
  particle_count_ = count;
  ParticleDataManager::allocateDevice(d_particles_, count);
  ParticleDataManager::allocateHost(h_particles_, count);
}

void ParticleSystem::freeMemory() {
  force_calculator_.reset();
  if (d_particles_.pos_x) ParticleDataManager::freeDevice(d_particles_);
  if (h_particles_.pos_x) ParticleDataManager::freeHost(h_particles_);
  is_initialized_ = false;
}

void ParticleSystem::finishInitialization() {
  ParticleDataManager::copyToDevice(d_particles_, h_particles_);
  createForceCalculator();
  integrator_ = std::make_unique<Integrator>(config_.cuda_block_size);
  force_calculator_->computeForces(&d_particles_);
  is_paused_ = false;
  is_initialized_ = true;
}

void ParticleSystem::initialize(const SimulationConfig& config) {
  validateSimulationConfig(config);
  validateResourceRequirements(config.particle_count);
  config_ = config;
  dt_ = config.dt;
  G_ = config.G;
  softening_ = config.softening;
  force_method_ = config.force_method;
  is_paused_ = false;
  freeMemory();
  allocateMemory(config.particle_count);
  switch (config.init_distribution) {
    case InitDistribution::UNIFORM: {
      UniformDistParams p;
      p.min_bounds = Vec3(-10, -10, -10);
      p.max_bounds = Vec3(10, 10, 10);
      ParticleInitializer::initUniform(h_particles_, p);
      break;
    }
    case InitDistribution::SPHERICAL: {
      SphericalDistParams p;
      p.center = Vec3(0, 0, 0);
      p.radius = 10.0f;
      ParticleInitializer::initSpherical(h_particles_, p);
      break;
    }
    case InitDistribution::DISK: {
      DiskDistParams p;
      p.center = Vec3(0, 0, 0);
      p.radius = 10.0f;
      p.thickness = 1.0f;
      p.rotation_speed = 0.5f;
      ParticleInitializer::initDisk(h_particles_, p);
      break;
    }
  }
  simulation_time_ = 0.0f;
  finishInitialization();
}

void ParticleSystem::initializeWithDistribution(size_t particle_count, InitDistribution dist) {
  SimulationConfig c;
  c.particle_count = particle_count;
  c.init_distribution = dist;
  initialize(c);
}

void ParticleSystem::createForceCalculator() {
  force_calculator_ = nbody::createForceCalculator(force_method_, config_);
  force_calculator_->setGravitationalConstant(G_);
  force_calculator_->setSofteningParameter(softening_);
}

void ParticleSystem::update(float dt) {
  if (!is_initialized_ || is_paused_) return;
  validateTimeStep(dt);
  NBODY_PROFILE_SCOPE("simulation.update");
  integrator_->integrate(&d_particles_, force_calculator_.get(), dt);
  simulation_time_ += dt;
}

void ParticleSystem::reset() {
  if (!is_initialized_) return;
  initialize(config_);
}

void ParticleSystem::setForceMethod(ForceMethod method) {
  if (force_method_ != method) {
    force_method_ = method;
    config_.force_method = method;
    if (is_initialized_) createForceCalculator();
  }
}

void ParticleSystem::setGravitationalConstant(float G) {
  if (G <= 0 || std::isnan(G) || std::isinf(G))
    throw ValidationException("Gravitational constant must be positive and finite");
  G_ = G;
  config_.G = G;
  if (force_calculator_) force_calculator_->setGravitationalConstant(G);
}

void ParticleSystem::setSofteningParameter(float eps) {
  validateSoftening(eps);
  softening_ = eps;
  config_.softening = eps;
  if (force_calculator_) force_calculator_->setSofteningParameter(eps);
}

void ParticleSystem::setTimeStep(float dt) {
  validateTimeStep(dt);
  dt_ = dt;
  config_.dt = dt;
}

void ParticleSystem::setBarnesHutTheta(float theta) {
  validateTheta(theta);
  config_.barnes_hut_theta = theta;
  if (force_method_ == ForceMethod::BARNES_HUT)
    if (auto* bh = dynamic_cast<BarnesHutCalculator*>(force_calculator_.get())) bh->setTheta(theta);
}

// :192-224 of the reference's particle_system.cpp.
void ParticleSystem::setSpatialHashCellSize(float size) {
  if (size <= 0 || std::isnan(size) || std::isinf(size))
    throw ValidationException("Spatial hash cell size must be positive and finite");
  if (size < config_.spatial_hash_cutoff)
    throw ValidationException("Spatial hash cell size must not be smaller than the cutoff");
  config_.spatial_hash_cell_size = size;
  if (force_method_ == ForceMethod::SPATIAL_HASH)
    if (auto* sh = dynamic_cast<SpatialHashCalculator*>(force_calculator_.get()))
      sh->setCellSize(size);
}

void ParticleSystem::setSpatialHashCutoff(float cutoff) {
  if (cutoff <= 0 || std::isnan(cutoff) || std::isinf(cutoff))
    throw ValidationException("Spatial hash cutoff must be positive and finite");
  if (cutoff > config_.spatial_hash_cell_size)
    throw ValidationException("Spatial hash cutoff must not exceed the cell size");
  config_.spatial_hash_cutoff = cutoff;
  if (force_method_ == ForceMethod::SPATIAL_HASH)
    if (auto* sh = dynamic_cast<SpatialHashCalculator*>(force_calculator_.get()))
      sh->setCutoffRadius(cutoff);
}

void ParticleSystem::copyToHost(ParticleData& h) const {
  ParticleDataManager::copyToHost(h, d_particles_);
}

SimulationState ParticleSystem::getState() const {
  SimulationState st;
  st.particle_count = particle_count_;
  st.simulation_time = simulation_time_;
  st.dt = dt_;
  st.G = G_;
  st.softening = softening_;
  st.force_method = force_method_;
  ParticleData h;
  ParticleDataManager::allocateHost(h, particle_count_);
  try {
    ParticleDataManager::copyToHost(h, d_particles_);
  } catch (...) {
    ParticleDataManager::freeHost(h);
    throw;
  }
  st.pos_x.assign(h.pos_x, h.pos_x + particle_count_);
  st.pos_y.assign(h.pos_y, h.pos_y + particle_count_);
  st.pos_z.assign(h.pos_z, h.pos_z + particle_count_);
  st.vel_x.assign(h.vel_x, h.vel_x + particle_count_);
  st.vel_y.assign(h.vel_y, h.vel_y + particle_count_);
  st.vel_z.assign(h.vel_z, h.vel_z + particle_count_);
  st.mass.assign(h.mass, h.mass + particle_count_);
  ParticleDataManager::freeHost(h);
  return st;
}

void ParticleSystem::setState(const SimulationState& st) {
  const size_t n = st.particle_count;
  if (st.pos_x.size() != n || st.pos_y.size() != n || st.pos_z.size() != n ||
      st.vel_x.size() != n || st.vel_y.size() != n || st.vel_z.size() != n || st.mass.size() != n)
    throw ValidationException("SimulationState data vector sizes do not match particle_count");
  SimulationConfig c = config_;
  c.particle_count = n;
  c.dt = st.dt;
  c.G = st.G;
  c.softening = st.softening;
  c.force_method = st.force_method;
  validateSimulationConfig(c);
  validateResourceRequirements(n);
  freeMemory();
  allocateMemory(n);
  std::copy(st.pos_x.begin(), st.pos_x.end(), h_particles_.pos_x);
  std::copy(st.pos_y.begin(), st.pos_y.end(), h_particles_.pos_y);
  std::copy(st.pos_z.begin(), st.pos_z.end(), h_particles_.pos_z);
  std::copy(st.vel_x.begin(), st.vel_x.end(), h_particles_.vel_x);
  std::copy(st.vel_y.begin(), st.vel_y.end(), h_particles_.vel_y);
  std::copy(st.vel_z.begin(), st.vel_z.end(), h_particles_.vel_z);
  std::copy(st.mass.begin(), st.mass.end(), h_particles_.mass);
  ParticleInitializer::zeroAccelerations(h_particles_);
  simulation_time_ = st.simulation_time;
  dt_ = st.dt;
  G_ = st.G;
  softening_ = st.softening;
  force_method_ = st.force_method;
  config_ = c;
  finishInitialization();
}

void ParticleSystem::saveState(const std::string& filename) const {
  Serializer::save(filename, getState());
}

void ParticleSystem::loadState(const std::string& filename) { setState(Serializer::load(filename)); }

float ParticleSystem::computeKineticEnergy() const {
  return integrator_ ? integrator_->computeKineticEnergy(&d_particles_) : 0.0f;
}
float ParticleSystem::computePotentialEnergy() const {
  return integrator_ ? integrator_->computePotentialEnergy(&d_particles_, G_, softening_) : 0.0f;
}
float ParticleSystem::computeTotalEnergy() const {
  return computeKineticEnergy() + computePotentialEnergy();
}

}  // namespace nbody
