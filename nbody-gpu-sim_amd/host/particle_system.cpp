// ParticleSystem over the engine's C++ operators and C ABI. The public
// behaviour follows src/core/particle_system.cpp of the reference (defaults,
// validation order and messages, a(0) on every new particle set, pause/resume,
// reset, state round trips), without the OpenGL interop.
//
// Structure: a Run holds everything of one simulation (settings, device
// particle set, operators). initialize() samples a SimulationState from the
// configured distribution and installs it through the same path as setState(),
// so a fresh run and a restored one are set up identically.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nbody/error_handling.hpp"
#include "nbody/particle_system.hpp"
#include "nbody/performance_observability.hpp"
#include "nbody/serialization.hpp"
#include "nbx.h"

namespace nbody {

namespace {

// Initial conditions of each distribution, with the reference's parameters
// (src/core/particle_system.cpp:80-104): a +-10 cube, a sphere of radius 10, a
// disk of radius 10, thickness 1 and rotation speed 0.5, centred at the origin;
// masses 1 (the parameter structs' defaults); seed 42.
void sample_distribution(InitDistribution dist, ParticleData& h) {
  switch (dist) {
    case InitDistribution::UNIFORM: {
      UniformDistParams cube;
      cube.min_bounds = Vec3(-10.0f, -10.0f, -10.0f);
      cube.max_bounds = Vec3(10.0f, 10.0f, 10.0f);
      ParticleInitializer::initUniform(h, cube);
      return;
    }
    case InitDistribution::SPHERICAL: {
      SphericalDistParams ball;
      ball.center = Vec3(0.0f, 0.0f, 0.0f);
      ball.radius = 10.0f;
      ParticleInitializer::initSpherical(h, ball);
      return;
    }
    case InitDistribution::DISK: {
      DiskDistParams disk;
      disk.center = Vec3(0.0f, 0.0f, 0.0f);
      disk.radius = 10.0f;
      disk.thickness = 1.0f;
      disk.rotation_speed = 0.5f;
      ParticleInitializer::initDisk(h, disk);
      return;
    }
  }
  throw ValidationException("Unknown initial distribution");
}

// Host particle arrays released on scope exit.
struct HostArrays {
  ParticleData h{};
  explicit HostArrays(size_t n) { ParticleDataManager::allocateHost(h, n); }
  ~HostArrays() {
    try {
      ParticleDataManager::freeHost(h);
    } catch (...) {
    }
  }
  HostArrays(const HostArrays&) = delete;
  HostArrays& operator=(const HostArrays&) = delete;
};

void check_sizes(const SimulationState& st) {
  const size_t n = st.particle_count;
  for (const auto* v : {&st.pos_x, &st.pos_y, &st.pos_z, &st.vel_x, &st.vel_y, &st.vel_z,
                        &st.mass})
    if (v->size() != n)
      throw ValidationException("SimulationState data vector sizes do not match particle_count");
  if (st.hasLowParts())
    for (const auto* v : {&st.pos_lo_x, &st.pos_lo_y, &st.pos_lo_z})
      if (v->size() != n)
        throw ValidationException("SimulationState low-part vectors do not match particle_count");
}

void require_direct(ForceMethod m) {
  if (m != ForceMethod::DIRECT_N2)
    throw ValidationException("The precision form runs with the direct force method only");
}

}  // namespace

struct ParticleSystem::Run {
  SimulationConfig config;  // method, dt, G, softening, theta, grid, block size
  size_t n = 0;
  float time = 0.0f;
  bool paused = false;
  bool live = false;        // particles on the device, a(0) computed
  bool precision = false;   // the precision form is selected (survives re-initialisation)
  ParticleData dev{};
  float* lo[3] = {nullptr, nullptr, nullptr};  // precision form: low parts of the positions
  nbx_direct* ws = nullptr;                    // precision form: pair-symmetric workspace
  std::unique_ptr<ForceCalculator> force;
  std::unique_ptr<Integrator> integrator;

  ~Run() {
    try {
      drop();
    } catch (...) {
    }
  }

  float eps2() const { return config.softening * config.softening; }

  nbx_ds_particles ds_view() const {
    static_assert(sizeof(nbx_particles) == sizeof(ParticleData), "same 13-pointer layout");
    nbx_ds_particles v;
    std::memcpy(&v.p, &dev, sizeof(nbx_particles));
    v.pos_lo_x = lo[0];
    v.pos_lo_y = lo[1];
    v.pos_lo_z = lo[2];
    return v;
  }

  void drop() {
    live = false;
    force.reset();
    integrator.reset();
    if (ws) (void)nbx_direct_destroy(ws);
    ws = nullptr;
    for (float*& p : lo) {
      if (p) (void)nbx_free(p);
      p = nullptr;
    }
    if (dev.pos_x) ParticleDataManager::freeDevice(dev);
    dev = ParticleData{};
    n = 0;
  }

  void make_force() {
    force = nbody::createForceCalculator(config.force_method, config);
    force->setGravitationalConstant(config.G);
    force->setSofteningParameter(config.softening);
    if (auto* direct = dynamic_cast<DirectForceCalculator*>(force.get())) direct->reserve(n);
  }

  // a(0) for the particle set now on the device.
  void accelerations() {
    if (precision) {
      nbx_ds_particles v = ds_view();
      NBODY_CHECK(nbx_ds_forces(ws, &v, config.G, eps2(), nullptr));
    } else {
      force->computeForces(&dev);
    }
  }

  // Puts st's particles on the device with the current settings and computes a(0).
  void install(const SimulationState& st) {
    drop();
    n = st.particle_count;
    ParticleDataManager::allocateDevice(dev, n);
    std::vector<float> zeros(n, 0.0f);
    ParticleData view{};
    view.pos_x = const_cast<float*>(st.pos_x.data());
    view.pos_y = const_cast<float*>(st.pos_y.data());
    view.pos_z = const_cast<float*>(st.pos_z.data());
    view.vel_x = const_cast<float*>(st.vel_x.data());
    view.vel_y = const_cast<float*>(st.vel_y.data());
    view.vel_z = const_cast<float*>(st.vel_z.data());
    view.acc_x = view.acc_y = view.acc_z = zeros.data();
    view.mass = const_cast<float*>(st.mass.data());
    view.count = n;
    ParticleDataManager::copyToDevice(dev, view);
    if (precision) {
      const std::vector<float>* parts[3] = {&st.pos_lo_x, &st.pos_lo_y, &st.pos_lo_z};
      for (int k = 0; k < 3; ++k) {
        NBODY_CHECK(nbx_malloc(reinterpret_cast<void**>(&lo[k]), n * sizeof(float)));
        const float* src = st.hasLowParts() ? parts[k]->data() : zeros.data();
        NBODY_CHECK(nbx_memcpy_htod(lo[k], src, n * sizeof(float)));
      }
      NBODY_CHECK(nbx_direct_create(n, 0, &ws));
    }
    time = st.simulation_time;
    make_force();
    integrator = std::make_unique<Integrator>(config.cuda_block_size);
    accelerations();
    paused = false;
    live = true;
  }
};

ParticleSystem::ParticleSystem() : run_(std::make_unique<Run>()) {}
ParticleSystem::~ParticleSystem() = default;

void ParticleSystem::initialize(const SimulationConfig& config) {
  validateSimulationConfig(config);
  validateResourceRequirements(config.particle_count);
  if (run_->precision) require_direct(config.force_method);
  SimulationState st;
  st.particle_count = config.particle_count;
  {
    HostArrays host(config.particle_count);
    sample_distribution(config.init_distribution, host.h);
    const ParticleData& h = host.h;
    const size_t n = config.particle_count;
    st.pos_x.assign(h.pos_x, h.pos_x + n);
    st.pos_y.assign(h.pos_y, h.pos_y + n);
    st.pos_z.assign(h.pos_z, h.pos_z + n);
    st.vel_x.assign(h.vel_x, h.vel_x + n);
    st.vel_y.assign(h.vel_y, h.vel_y + n);
    st.vel_z.assign(h.vel_z, h.vel_z + n);
    st.mass.assign(h.mass, h.mass + n);
  }
  run_->config = config;
  run_->install(st);
}

void ParticleSystem::initializeWithDistribution(size_t particle_count, InitDistribution dist) {
  SimulationConfig config;
  config.particle_count = particle_count;
  config.init_distribution = dist;
  initialize(config);
}

void ParticleSystem::update(float dt) {
  Run& r = *run_;
  if (!r.live || r.paused) return;
  validateTimeStep(dt);
  NBODY_PROFILE_SCOPE("simulation.update");
  if (r.precision) {
    nbx_ds_particles v = r.ds_view();
    NBODY_CHECK(nbx_ds_step(r.ws, &v, r.config.G, r.eps2(), dt, nullptr));
  } else {
    r.integrator->integrate(&r.dev, r.force.get(), dt);
  }
  r.time += dt;
}

void ParticleSystem::pause() { run_->paused = true; }
void ParticleSystem::resume() { run_->paused = false; }
bool ParticleSystem::isPaused() const noexcept { return run_->paused; }

void ParticleSystem::reset() {
  if (run_->live) initialize(run_->config);
}

void ParticleSystem::setForceMethod(ForceMethod method) {
  Run& r = *run_;
  if (method == r.config.force_method) return;
  if (r.precision) require_direct(method);
  r.config.force_method = method;
  r.make_force();  // the accelerations of the previous method stay until the next step
}

void ParticleSystem::setGravitationalConstant(float G) {
  if (!(G > 0.0f) || std::isinf(G))
    throw ValidationException("Gravitational constant must be positive and finite");
  run_->config.G = G;
  if (run_->force) run_->force->setGravitationalConstant(G);
}

void ParticleSystem::setSofteningParameter(float eps) {
  validateSoftening(eps);
  run_->config.softening = eps;
  if (run_->force) run_->force->setSofteningParameter(eps);
}

void ParticleSystem::setTimeStep(float dt) {
  validateTimeStep(dt);
  run_->config.dt = dt;
}

void ParticleSystem::setBarnesHutTheta(float theta) {
  validateTheta(theta);
  run_->config.barnes_hut_theta = theta;
  if (auto* bh = dynamic_cast<BarnesHutCalculator*>(run_->force.get())) bh->setTheta(theta);
}

void ParticleSystem::setSpatialHashCellSize(float size) {
  if (!(size > 0.0f) || std::isinf(size))
    throw ValidationException("Spatial hash cell size must be positive and finite");
  if (size < run_->config.spatial_hash_cutoff)
    throw ValidationException("Spatial hash cell size must not be smaller than the cutoff");
  run_->config.spatial_hash_cell_size = size;
  if (auto* sh = dynamic_cast<SpatialHashCalculator*>(run_->force.get())) sh->setCellSize(size);
}

void ParticleSystem::setSpatialHashCutoff(float cutoff) {
  if (!(cutoff > 0.0f) || std::isinf(cutoff))
    throw ValidationException("Spatial hash cutoff must be positive and finite");
  if (cutoff > run_->config.spatial_hash_cell_size)
    throw ValidationException("Spatial hash cutoff must not exceed the cell size");
  run_->config.spatial_hash_cutoff = cutoff;
  if (auto* sh = dynamic_cast<SpatialHashCalculator*>(run_->force.get()))
    sh->setCutoffRadius(cutoff);
}

void ParticleSystem::setPrecisionMode(bool enabled) {
  Run& r = *run_;
  if (enabled == r.precision) return;
  if (enabled) require_direct(r.config.force_method);
  if (!r.live) {
    r.precision = enabled;
    return;
  }
  SimulationState st = getState();
  if (!enabled) {  // round hi + lo to the nearest FP32
    std::vector<float>* hi[3] = {&st.pos_x, &st.pos_y, &st.pos_z};
    const std::vector<float>* low[3] = {&st.pos_lo_x, &st.pos_lo_y, &st.pos_lo_z};
    for (int k = 0; k < 3; ++k)
      for (size_t i = 0; i < st.particle_count; ++i)
        (*hi[k])[i] = (float)((double)(*hi[k])[i] + (double)(*low[k])[i]);
    st.pos_lo_x.clear();
    st.pos_lo_y.clear();
    st.pos_lo_z.clear();
  }
  r.precision = enabled;
  const bool paused = r.paused;
  r.install(st);
  r.paused = paused;
}

ForceMethod ParticleSystem::getForceMethod() const noexcept { return run_->config.force_method; }
float ParticleSystem::getGravitationalConstant() const noexcept { return run_->config.G; }
float ParticleSystem::getSofteningParameter() const noexcept { return run_->config.softening; }
float ParticleSystem::getTimeStep() const noexcept { return run_->config.dt; }
float ParticleSystem::getSimulationTime() const noexcept { return run_->time; }
size_t ParticleSystem::getParticleCount() const noexcept { return run_->n; }
bool ParticleSystem::getPrecisionMode() const noexcept { return run_->precision; }

ParticleData* ParticleSystem::getDeviceData() { return &run_->dev; }
const ParticleData* ParticleSystem::getDeviceData() const { return &run_->dev; }

void ParticleSystem::copyToHost(ParticleData& h) const {
  ParticleDataManager::copyToHost(h, run_->dev);
}

SimulationState ParticleSystem::getState() const {
  const Run& r = *run_;
  SimulationState st;
  st.particle_count = r.n;
  st.simulation_time = r.time;
  st.dt = r.config.dt;
  st.G = r.config.G;
  st.softening = r.config.softening;
  st.force_method = r.config.force_method;
  if (r.n == 0) return st;
  HostArrays host(r.n);
  ParticleDataManager::copyToHost(host.h, r.dev);
  const ParticleData& h = host.h;
  st.pos_x.assign(h.pos_x, h.pos_x + r.n);
  st.pos_y.assign(h.pos_y, h.pos_y + r.n);
  st.pos_z.assign(h.pos_z, h.pos_z + r.n);
  st.vel_x.assign(h.vel_x, h.vel_x + r.n);
  st.vel_y.assign(h.vel_y, h.vel_y + r.n);
  st.vel_z.assign(h.vel_z, h.vel_z + r.n);
  st.mass.assign(h.mass, h.mass + r.n);
  if (r.precision) {
    std::vector<float>* parts[3] = {&st.pos_lo_x, &st.pos_lo_y, &st.pos_lo_z};
    for (int k = 0; k < 3; ++k) {
      parts[k]->resize(r.n);
      NBODY_CHECK(nbx_memcpy_dtoh(parts[k]->data(), r.lo[k], r.n * sizeof(float)));
    }
  }
  return st;
}

void ParticleSystem::setState(const SimulationState& st) {
  check_sizes(st);
  Run& r = *run_;
  SimulationConfig c = r.config;
  c.particle_count = st.particle_count;
  c.dt = st.dt;
  c.G = st.G;
  c.softening = st.softening;
  c.force_method = st.force_method;
  validateSimulationConfig(c);
  validateResourceRequirements(st.particle_count);
  if (st.hasLowParts() || r.precision) require_direct(c.force_method);
  r.config = c;
  if (st.hasLowParts()) r.precision = true;
  r.install(st);
}

void ParticleSystem::saveState(const std::string& filename) const {
  Serializer::save(filename, getState());
}

void ParticleSystem::loadState(const std::string& filename) { setState(Serializer::load(filename)); }

float ParticleSystem::computeKineticEnergy() const {
  const Run& r = *run_;
  if (!r.live) return 0.0f;
  if (!r.precision) return r.integrator->computeKineticEnergy(&r.dev);
  nbx_ds_particles v = r.ds_view();
  double ke = 0.0;
  NBODY_CHECK(nbx_ds_energy(&v, r.config.G, r.eps2(), &ke, nullptr, nullptr, nullptr));
  return (float)ke;
}

float ParticleSystem::computePotentialEnergy() const {
  const Run& r = *run_;
  if (!r.live) return 0.0f;
  if (!r.precision)
    return r.integrator->computePotentialEnergy(&r.dev, r.config.G, r.config.softening);
  nbx_ds_particles v = r.ds_view();
  double pe = 0.0;
  NBODY_CHECK(nbx_ds_energy(&v, r.config.G, r.eps2(), nullptr, &pe, nullptr, nullptr));
  return (float)pe;
}

float ParticleSystem::computeTotalEnergy() const {
  return computeKineticEnergy() + computePotentialEnergy();
}

}  // namespace nbody
