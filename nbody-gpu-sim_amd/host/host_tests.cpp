// Tests of the C++ API (the reference's include/nbody interfaces on this
// engine). A small self-contained harness: `host_tests --cpu` runs the
// host-only cases, `host_tests --gpu` the cases that launch kernels.
// Cases follow the reference's tests (test_serialization.cpp, test_app_cli.cpp,
// test_validation.cpp, test_force_calculation.cpp, test_integrator.cpp,
// test_barnes_hut.cpp, test_simulation_control.cpp) plus exact resume.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "nbody/app_cli.hpp"
#include "nbody/barnes_hut_tree.hpp"
#include "nbody/error_handling.hpp"
#include "nbody/force_calculator.hpp"
#include "nbody/integrator.hpp"
#include "nbody/particle_data.hpp"
#include "nbody/particle_system.hpp"
#include "nbody/performance_observability.hpp"
#include "nbody/serialization.hpp"
#include "nbody/spatial_hash_grid.hpp"

using namespace nbody;

namespace {

struct Case {
  const char* name;
  bool gpu;
  std::function<void()> fn;
};
std::vector<Case>& registry() {
  static std::vector<Case> r;
  return r;
}
struct Reg {
  Reg(const char* n, bool g, std::function<void()> f) { registry().push_back({n, g, f}); }
};
struct Failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define TEST_CPU(name) \
  static void name();  \
  static Reg reg_##name(#name, false, name); \
  static void name()
#define TEST_GPU(name) \
  static void name();  \
  static Reg reg_##name(#name, true, name); \
  static void name()
#define CHECK(cond)                                                                        \
  do {                                                                                     \
    if (!(cond)) throw Failure(std::string(__FILE__) + ":" + std::to_string(__LINE__) + \
                               ": " #cond);                                                \
  } while (0)
#define CHECK_NEAR(a, b, tol) CHECK(std::fabs((double)(a) - (double)(b)) <= (double)(tol))
#define CHECK_THROWS(expr, T)   \
  do {                          \
    bool thrown_ = false;       \
    try {                       \
      expr;                     \
    } catch (const T&) {        \
      thrown_ = true;           \
    }                           \
    CHECK(thrown_);             \
  } while (0)

SimulationState make_state(size_t n, unsigned seed) {
  SimulationState s;
  s.particle_count = n;
  s.simulation_time = 1.5f;
  s.dt = 0.001f;
  s.G = 1.0f;
  s.softening = 0.05f;
  s.force_method = ForceMethod::BARNES_HUT;
  std::vector<float>* v[7] = {&s.pos_x, &s.pos_y, &s.pos_z, &s.vel_x, &s.vel_y, &s.vel_z, &s.mass};
  unsigned x = seed;
  for (auto* a : v) {
    a->resize(n);
    for (auto& e : *a) {
      x = x * 1664525u + 1013904223u;
      e = (float)(x >> 8) / 16777216.0f + 0.01f;
    }
  }
  return s;
}

struct Scoped {
  ParticleData d, h;
  explicit Scoped(size_t n) {
    ParticleDataManager::allocateDevice(d, n);
    ParticleDataManager::allocateHost(h, n);
    for (size_t i = 0; i < n; ++i) {
      h.pos_x[i] = h.pos_y[i] = h.pos_z[i] = h.vel_x[i] = h.vel_y[i] = h.vel_z[i] = 0.0f;
      h.acc_x[i] = h.acc_y[i] = h.acc_z[i] = h.acc_old_x[i] = h.acc_old_y[i] = h.acc_old_z[i] = 0.0f;
      h.mass[i] = 1.0f;
    }
  }
  ~Scoped() {
    ParticleDataManager::freeDevice(d);
    ParticleDataManager::freeHost(h);
  }
  void set(size_t i, Vec3 p, Vec3 v, float m) {
    h.pos_x[i] = p.x;
    h.pos_y[i] = p.y;
    h.pos_z[i] = p.z;
    h.vel_x[i] = v.x;
    h.vel_y[i] = v.y;
    h.vel_z[i] = v.z;
    h.mass[i] = m;
  }
  void up() { ParticleDataManager::copyToDevice(d, h); }
  void down() { ParticleDataManager::copyToHost(h, d); }
};

// ------------------------------------------------------------------ CPU cases

TEST_CPU(serialization_round_trip_sizes) {
  for (size_t n : {0, 1, 10, 1000}) {
    const SimulationState s = make_state(n, 7 + (unsigned)n);
    std::stringstream buf;
    Serializer::save(buf, s);
    CHECK(buf.str().size() == 56 + 7 * 4 * n);
    const SimulationState r = Serializer::load(buf);
    CHECK(r == s);
    CHECK(std::memcmp(r.pos_x.data(), s.pos_x.data(), 4 * n) == 0);
  }
}

TEST_CPU(serialization_header_bytes) {
  std::stringstream a, b;
  Serializer::save(a, make_state(3, 1));
  Serializer::save(b, make_state(3, 1));
  CHECK(a.str() == b.str());
  uint32_t magic, version;
  uint64_t count;
  std::memcpy(&magic, a.str().data(), 4);
  std::memcpy(&version, a.str().data() + 4, 4);
  std::memcpy(&count, a.str().data() + 8, 8);
  CHECK(magic == 0x4E424F44u && version == 1 && count == 3);
  uint32_t method;
  std::memcpy(&method, a.str().data() + 32, 4);
  CHECK(method == 1);
  for (int k = 36; k < 56; ++k) CHECK(a.str()[k] == 0);  // reserved + padding zeroed
}

TEST_CPU(serialization_rejects_corrupt_input) {
  std::stringstream good;
  Serializer::save(good, make_state(4, 2));
  const std::string g = good.str();
  auto load = [](std::string bytes) {
    std::stringstream s(bytes);
    return Serializer::load(s);
  };
  std::string bad = g;
  bad[0] ^= 0x5a;
  CHECK_THROWS(load(bad), ValidationException);
  bad = g;
  bad[4] = 3;  // unknown version
  CHECK_THROWS(load(bad), ValidationException);
  bad = g;
  bad[4] = 2;  // a v1 body under a v2 header: the low parts are missing
  CHECK_THROWS(load(bad), ValidationException);
  CHECK_THROWS(load(g.substr(0, 40)), ValidationException);
  CHECK_THROWS(load(g.substr(0, g.size() - 1)), ValidationException);
  bad = g;
  const uint64_t huge = MAX_PARTICLE_COUNT + 1;
  std::memcpy(&bad[8], &huge, 8);
  CHECK_THROWS(load(bad), ValidationException);
  bad = g;
  const uint32_t m = 3;
  std::memcpy(&bad[32], &m, 4);
  CHECK_THROWS(load(bad), ValidationException);
  std::stringstream ok(g), trunc(g.substr(0, 10));
  CHECK(Serializer::validateStream(ok));
  CHECK(!Serializer::validateStream(trunc));
  CHECK_THROWS(Serializer::load(std::string("/nonexistent/dir/file.nbody")), IOException);
}

// Version 2 (precision-form state): the v1 header and arrays, then the three
// low-part arrays. A reader of version 1 only (the reference's, which checks
// version == 1 right after the magic) sees version 2 and rejects the file.
TEST_CPU(serialization_version2_low_parts) {
  for (size_t n : {1, 10, 1000}) {
    SimulationState s = make_state(n, 11 + (unsigned)n);
    for (auto* v : {&s.pos_lo_x, &s.pos_lo_y, &s.pos_lo_z}) {
      v->resize(n);
      for (size_t i = 0; i < n; ++i) (*v)[i] = 1e-9f * (float)(i % 13) - 3e-9f;
    }
    std::stringstream buf;
    Serializer::save(buf, s);
    const std::string bytes = buf.str();
    CHECK(bytes.size() == 56 + 10 * 4 * n);
    uint32_t version;
    std::memcpy(&version, bytes.data() + 4, 4);
    CHECK(version == NBODY_VERSION_DOUBLE_SINGLE && version != NBODY_VERSION);
    // the first seven arrays sit where version 1 puts them
    SimulationState plain = s;
    plain.pos_lo_x.clear();
    plain.pos_lo_y.clear();
    plain.pos_lo_z.clear();
    std::stringstream v1;
    Serializer::save(v1, plain);
    CHECK(v1.str().size() == 56 + 7 * 4 * n);
    CHECK(std::memcmp(v1.str().data() + 8, bytes.data() + 8, v1.str().size() - 8) == 0);
    std::stringstream in(bytes);
    const SimulationState r = Serializer::load(in);
    CHECK(r == s && r.hasLowParts());
    CHECK(std::memcmp(r.pos_lo_y.data(), s.pos_lo_y.data(), 4 * n) == 0);
    std::stringstream check(bytes);
    CHECK(Serializer::validateStream(check));
  }
  SimulationState bad = make_state(5, 3);
  bad.pos_lo_x.assign(4, 0.0f);
  bad.pos_lo_y.assign(5, 0.0f);
  bad.pos_lo_z.assign(5, 0.0f);
  std::stringstream out;
  CHECK_THROWS(Serializer::save(out, bad), ValidationException);
}

TEST_CPU(cli_parsing) {
  {
    const char* argv[] = {"nbody_sim"};
    const AppCliOptions o = parseAppCliOptions(1, argv);
    CHECK(o.particle_count == 10000 && o.force_method == ForceMethod::DIRECT_N2);
    CHECK(!o.benchmark_mode && o.benchmark_steps == 120);
  }
  for (const char* m : {"direct", "direct-n2", "direct_n2"}) {
    const char* argv[] = {"nbody_sim", "--method", m};
    CHECK(parseAppCliOptions(3, argv).force_method == ForceMethod::DIRECT_N2);
  }
  for (const char* m : {"barnes-hut", "barnes_hut"}) {
    const char* argv[] = {"nbody_sim", "--method", m};
    CHECK(parseAppCliOptions(3, argv).force_method == ForceMethod::BARNES_HUT);
  }
  {
    const char* argv[] = {"nbody_sim", "5000", "--dt", "0.002", "--theta", "0.7",
                          "--benchmark-steps", "9", "--export", "a.nbody", "--import", "b.nbody"};
    const AppCliOptions o = parseAppCliOptions(12, argv);
    CHECK(o.particle_count == 5000 && std::fabs(o.dt - 0.002f) < 1e-9f);
    CHECK(o.benchmark_mode && o.benchmark_steps == 9 && o.export_path == "a.nbody" &&
          o.import_path == "b.nbody");
    CHECK(std::fabs(o.barnes_hut_theta - 0.7f) < 1e-7f);
  }
  const char* bad1[] = {"nbody_sim", "--frobnicate"};
  CHECK_THROWS(parseAppCliOptions(2, bad1), ValidationException);
  const char* bad2[] = {"nbody_sim", "--dt"};
  CHECK_THROWS(parseAppCliOptions(2, bad2), ValidationException);
  const char* bad3[] = {"nbody_sim", "--particles", "abc"};
  CHECK_THROWS(parseAppCliOptions(3, bad3), ValidationException);
  const char* bad4[] = {"nbody_sim", "--method", "fmm"};
  CHECK_THROWS(parseAppCliOptions(3, bad4), ValidationException);
  const char* bad5[] = {"nbody_sim", "--benchmark-steps", "0"};
  CHECK_THROWS(parseAppCliOptions(3, bad5), ValidationException);
  const char* bad6[] = {"nbody_sim", "--softening", "0"};
  CHECK_THROWS(parseAppCliOptions(3, bad6), ValidationException);
  const char* bad7[] = {"nbody_sim", "--theta", "3"};
  CHECK_THROWS(parseAppCliOptions(3, bad7), ValidationException);
  const char* bad8[] = {"nbody_sim", "--particles", "0"};
  CHECK_THROWS(parseAppCliOptions(3, bad8), ValidationException);
  CHECK(appCliUsage().find("--method") != std::string::npos);
}

TEST_CPU(validators) {
  CHECK_THROWS(validateTimeStep(0.0f), ValidationException);
  CHECK_THROWS(validateTimeStep(2.0f), ValidationException);
  CHECK_THROWS(validateTimeStep(NAN), ValidationException);
  validateTimeStep(0.01f);
  CHECK_THROWS(validateSoftening(-1.0f), ValidationException);
  CHECK_THROWS(validateSoftening(INFINITY), ValidationException);
  CHECK_THROWS(validateTheta(-0.1f), ValidationException);
  CHECK_THROWS(validateParticleCountRange(0), ValidationException);
  CHECK_THROWS(validateParticleCountRange(MAX_PARTICLE_COUNT + 1), ValidationException);
  SimulationConfig c;
  validateSimulationConfig(c);
  c.cuda_block_size = 300;
  CHECK_THROWS(validateSimulationConfig(c), ValidationException);
  c.cuda_block_size = 256;
  c.G = 0;
  CHECK_THROWS(validateSimulationConfig(c), ValidationException);
}

TEST_CPU(cpu_pair_function) {
  Vec3 a = computeGravitationalAccelerationCPU(Vec3(0, 0, 0), Vec3(1, 0, 0), 1.0f, 1.0f, 0.0f);
  CHECK(a.x > 0 && std::fabs(a.y) < 1e-6 && std::fabs(a.length() - 1.0f) < 1e-5);
  a = computeGravitationalAccelerationCPU(Vec3(0, 0, 0), Vec3(0.001f, 0, 0), 1.0f, 1.0f, 0.1f);
  CHECK(std::isfinite(a.x));
  a = computeGravitationalAccelerationCPU(Vec3(0, 0, 0), Vec3(1, 1, 1), 1.0f, 1.0f, 0.01f);
  const Vec3 dir = a.normalized();
  CHECK_NEAR(dir.x, 1.0 / std::sqrt(3.0), 1e-5);
}

TEST_CPU(morton_utilities) {
  for (unsigned v = 0; v < 1024; ++v) {
    unsigned naive = 0;
    for (int b = 0; b < 10; ++b) naive |= ((v >> b) & 1u) << (3 * b);
    CHECK(expandBits(v) == naive);
  }
  CHECK(computeMortonCode(0, 0, 0) == 0);
  CHECK(computeMortonCode(1, 1, 1) == 0x3FFFFFFFu);
  CHECK(computeMortonCode(-5, 2, 0.5f) == ((expandBits(1023) << 1) | expandBits(511)));
}

TEST_CPU(factory_and_spatial_hash) {
  SimulationConfig c;
  auto d = createForceCalculator(ForceMethod::DIRECT_N2, c);
  CHECK(d->getMethod() == ForceMethod::DIRECT_N2);
  c.barnes_hut_theta = 0.3f;
  auto b = createForceCalculator(ForceMethod::BARNES_HUT, c);
  CHECK(b->getMethod() == ForceMethod::BARNES_HUT);
  CHECK(dynamic_cast<BarnesHutCalculator*>(b.get())->getTheta() == 0.3f);
  c.spatial_hash_cell_size = 3.0f;
  c.spatial_hash_cutoff = 1.5f;
  auto sh = createForceCalculator(ForceMethod::SPATIAL_HASH, c);
  CHECK(sh->getMethod() == ForceMethod::SPATIAL_HASH);
  auto* shc = dynamic_cast<SpatialHashCalculator*>(sh.get());
  CHECK(shc && shc->getCellSize() == 3.0f && shc->getCutoffRadius() == 1.5f);
  CHECK(shc->getGrid() == nullptr);  // made on the first computeForces
  d->setSofteningParameter(0.2f);
  CHECK(d->getSofteningParameter() == 0.2f);
  CHECK(std::string(forceMethodToString(ForceMethod::BARNES_HUT)) == "barnes_hut");
}

// tests/test_spatial_hash.cpp:34-47 and hashCell's wrap of negative indices.
TEST_CPU(spatial_hash_cell_index) {
  const int3 a = SpatialHashGrid::getCellIndex(0.5f, 0.5f, 0.5f, 2.0f);
  CHECK(a.x == 0 && a.y == 0 && a.z == 0);
  const int3 b = SpatialHashGrid::getCellIndex(2.5f, 4.5f, 6.5f, 2.0f);
  CHECK(b.x == 1 && b.y == 2 && b.z == 3);
  const int3 c = SpatialHashGrid::getCellIndex(-0.5f, -2.0f, -2.5f, 2.0f);
  CHECK(c.x == -1 && c.y == -1 && c.z == -2);
  CHECK(SpatialHashGrid::hashCell(make_int3(-1, -1, -1), make_int3(3, 4, 5)) == 2 + 3 * 3 + 4 * 12);
  CHECK(SpatialHashGrid::hashCell(make_int3(4, 9, 5), make_int3(3, 4, 5)) == 1 + 1 * 3 + 0);
}

TEST_CPU(benchmark_json_schema) {
  BenchmarkRunRecord r;
  r.benchmark_name = "force.direct_n2";
  r.particle_count = 4096;
  r.iterations = 3;
  r.metrics["wall_time_ms"] = 1.5;
  r.parameters["dt"] = 0.001;
  r.phase_timings.push_back({"phase \"a\"", Milliseconds(2.0), 3});
  const std::string j = serializeBenchmarkRunRecords({r});
  CHECK(j.rfind("{\"benchmarks\":[{\"benchmark_name\":\"force.direct_n2\",", 0) == 0);
  CHECK(j.find("\"force_method\":\"direct_n2\"") != std::string::npos);
  CHECK(j.find("\"metrics\":{\"wall_time_ms\":1.5}") != std::string::npos);
  CHECK(j.find("\\\"a\\\"") != std::string::npos);
  CHECK(j.find("\"samples\":3}]}]}") != std::string::npos);
}

// tests/test_performance_observability.cpp:7-24 and :26-39.
TEST_CPU(benchmark_record_fields) {
  BenchmarkRunRecord r;
  r.benchmark_name = "force.direct_n2";
  r.force_method = ForceMethod::DIRECT_N2;
  r.particle_count = 4096;
  r.iterations = 12;
  r.metrics["wall_time_ms"] = 1.25;
  r.metrics["throughput_particles_per_second"] = 32768.0;
  r.parameters["cuda_block_size"] = 256.0;
  const std::string j = serializeBenchmarkRunRecord(r);
  CHECK(j.find("\"benchmark_name\":\"force.direct_n2\"") != std::string::npos);
  CHECK(j.find("\"particle_count\":4096") != std::string::npos);
  CHECK(j.find("\"force_method\":\"direct_n2\"") != std::string::npos);
  CHECK(j.find("\"wall_time_ms\":1.25") != std::string::npos);
  CHECK(j.find("\"cuda_block_size\":256") != std::string::npos);
}

TEST_CPU(scoped_phase_profile_accumulates) {
  PhaseProfiler profiler;
  {
    ScopedPhaseProfile profile(profiler, "simulation.update");
    std::this_thread::sleep_for(std::chrono::milliseconds(1));
  }
  {
    ScopedPhaseProfile profile(profiler, "simulation.update");
  }
  {
    ScopedPhaseProfile profile(profiler, "force.direct_n2");
  }
  const auto snap = profiler.snapshot();
  CHECK(snap.size() == 2u);
  CHECK(snap[0].name == "simulation.update" && snap[0].samples == 2u);
  CHECK(snap[0].total_duration.count() >= 0.5);
  CHECK(snap[1].name == "force.direct_n2" && snap[1].samples == 1u);
  const auto drained = profiler.consume();
  CHECK(drained.size() == 2u && profiler.snapshot().empty());
}

TEST_CPU(generators_follow_reference_ranges) {
  ParticleData h;
  ParticleDataManager::allocateHost(h, 500);
  SphericalDistParams p;
  p.radius = 5.0f;
  p.min_mass = 0.5f;
  p.max_mass = 2.0f;
  ParticleInitializer::initSpherical(h, p, 42);
  for (size_t i = 0; i < 500; ++i) {
    const float r = std::sqrt(h.pos_x[i] * h.pos_x[i] + h.pos_y[i] * h.pos_y[i] + h.pos_z[i] * h.pos_z[i]);
    CHECK(r <= 5.0001f && h.mass[i] >= 0.5f && h.mass[i] <= 2.0f && h.vel_x[i] == 0.0f);
    CHECK(h.acc_old_z[i] == 0.0f);
  }
  ParticleDataManager::freeHost(h);
}

// ------------------------------------------------------------------ GPU cases

TEST_GPU(direct_single_particle_zero) {
  Scoped s(1);
  s.set(0, Vec3(1, 2, 3), Vec3(), 5.0f);
  s.up();
  DirectForceCalculator calc(256);
  calc.setSofteningParameter(0.1f);
  calc.computeForces(&s.d);
  s.down();
  CHECK(s.h.acc_x[0] == 0.0f && s.h.acc_y[0] == 0.0f && s.h.acc_z[0] == 0.0f);
}

TEST_GPU(direct_two_body_and_grid32) {
  Scoped s(2);
  s.set(0, Vec3(-0.5f, 0, 0), Vec3(), 1.0f);
  s.set(1, Vec3(0.5f, 0, 0), Vec3(), 1.0f);
  s.up();
  DirectForceCalculator calc;
  calc.setSofteningParameter(0.01f);
  calc.computeForces(&s.d);
  s.down();
  CHECK_NEAR(s.h.acc_x[0], -s.h.acc_x[1], 1e-5);
  CHECK(s.h.acc_x[0] > 0);

  Scoped g(32);
  for (size_t i = 0; i < 32; i++)
    g.set(i, Vec3((float)((i * 7) % 11) * 0.3f, (float)((i * 3) % 13) * 0.3f,
                  (float)((i * 5) % 9) * 0.3f),
          Vec3(), 0.5f + 0.1f * (float)(i % 5));
  g.up();
  DirectForceCalculator c2;
  c2.setSofteningParameter(0.1f);
  c2.computeForces(&g.d);
  g.down();
  for (size_t i = 0; i < 32; i++) {
    Vec3 ref;
    for (size_t j = 0; j < 32; j++) {
      if (i == j) continue;
      ref += computeGravitationalAccelerationCPU(Vec3(g.h.pos_x[i], g.h.pos_y[i], g.h.pos_z[i]),
                                                 Vec3(g.h.pos_x[j], g.h.pos_y[j], g.h.pos_z[j]),
                                                 g.h.mass[j], 1.0f, 0.1f);
    }
    CHECK(std::fabs(g.h.acc_x[i] - ref.x) <= 1e-3f * (std::fabs(ref.x) + 1e-3f));
    CHECK(std::fabs(g.h.acc_y[i] - ref.y) <= 1e-3f * (std::fabs(ref.y) + 1e-3f));
    CHECK(std::fabs(g.h.acc_z[i] - ref.z) <= 1e-3f * (std::fabs(ref.z) + 1e-3f));
  }
}

TEST_GPU(integrator_closed_forms_and_energy) {
  Scoped s(1);
  s.set(0, Vec3(1, -2, 0.5f), Vec3(0.3f, 0, -0.2f), 1.0f);
  s.h.acc_x[0] = 2.0f;
  s.h.acc_y[0] = -1.0f;
  s.h.acc_z[0] = 4.0f;
  s.up();
  Integrator in;
  const float dt = 0.01f;
  in.updatePositions(&s.d, dt);
  s.down();
  CHECK_NEAR(s.h.pos_x[0], 1.0f + 0.3f * dt + 0.5f * 2.0f * dt * dt, 1e-6);
  CHECK_NEAR(s.h.pos_z[0], 0.5f - 0.2f * dt + 0.5f * 4.0f * dt * dt, 1e-6);

  Scoped k(2);
  k.set(0, Vec3(), Vec3(1, 0, 0), 1.0f);
  k.set(1, Vec3(), Vec3(0, 2, 0), 2.0f);
  k.up();
  CHECK_NEAR(in.computeKineticEnergy(&k.d), 4.5f, 1e-4);

  Scoped o(2);
  const float r = 5.0f, v = std::sqrt(1.0f / (4.0f * r));
  o.set(0, Vec3(-r, 0, 0), Vec3(0, -v, 0), 1.0f);
  o.set(1, Vec3(r, 0, 0), Vec3(0, v, 0), 1.0f);
  o.up();
  DirectForceCalculator f;
  f.setSofteningParameter(0.01f);
  f.computeForces(&o.d);
  const float e0 = in.computeTotalEnergy(&o.d, 1.0f, 0.01f);
  for (int s2 = 0; s2 < 100; ++s2) in.integrate(&o.d, &f, 0.001f);
  const float e1 = in.computeTotalEnergy(&o.d, 1.0f, 0.01f);
  CHECK(std::fabs(e1 - e0) / std::fabs(e0) < 1e-3f);
}

TEST_GPU(barnes_hut_tree_kats) {
  Scoped s(50);
  SphericalDistParams p;
  p.radius = 5.0f;
  ParticleInitializer::initSpherical(s.h, p, 42);
  s.up();
  BarnesHutTree t(50);
  t.build(&s.d);
  CHECK(t.getNodeCount() > 0);
  // getNodes() holds the nodes of the last build without an explicit copy
  float root_mass = t.getNodes()[0].total_mass;
  CHECK(std::fabs(root_mass - 50.0f) < 1e-3f);
  CHECK(t.verifyMassConservation(&s.h));
  for (size_t i = 0; i < 50; ++i) s.h.mass[i] = 2.0f;
  s.up();
  t.build(&s.d);
  CHECK(std::fabs(t.getNodes()[0].total_mass - 100.0f) < 1e-3f);
  t.copyNodesToHost();
  CHECK(t.verifyMassConservation(&s.h));
  for (int mode = 0; mode < 2; ++mode) {  // coincident, collinear
    Scoped c(16);
    for (size_t i = 0; i < 16; ++i)
      c.set(i, mode == 0 ? Vec3(1, 2, 3) : Vec3(0.0001f * (float)i, 0, 0), Vec3(), 1.0f);
    c.up();
    BarnesHutTree tc(16);
    tc.build(&c.d);
    tc.copyNodesToHost();
    CHECK(tc.verifyMassConservation(&c.h));
    CHECK(tc.getNodeCount() <= 32);
  }
}

TEST_GPU(barnes_hut_vs_direct) {
  Scoped s(500);
  SphericalDistParams p;
  p.radius = 10.0f;
  ParticleInitializer::initSpherical(s.h, p, 42);
  s.up();
  DirectForceCalculator d;
  d.setSofteningParameter(0.1f);
  d.computeForces(&s.d);
  s.down();
  std::vector<float> ax(s.h.acc_x, s.h.acc_x + 500), ay(s.h.acc_y, s.h.acc_y + 500),
      az(s.h.acc_z, s.h.acc_z + 500);
  auto err = [&](float theta) {
    BarnesHutCalculator b(theta);
    b.setSofteningParameter(0.1f);
    b.computeForces(&s.d);
    s.down();
    double e = 0, worst = 0;
    for (size_t i = 0; i < 500; ++i) {
      const double dx = s.h.acc_x[i] - ax[i], dy = s.h.acc_y[i] - ay[i], dz = s.h.acc_z[i] - az[i];
      const double m = std::sqrt((double)ax[i] * ax[i] + (double)ay[i] * ay[i] + (double)az[i] * az[i]);
      e += std::sqrt(dx * dx + dy * dy + dz * dz);
      worst = std::max(worst, std::sqrt(dx * dx + dy * dy + dz * dz) / m);
    }
    return std::make_pair(e, worst);
  };
  const auto fine = err(0.1f), mid = err(0.3f), coarse = err(0.8f);
  CHECK(fine.second < 0.02);
  CHECK(mid.first <= 1.1 * coarse.first);
}

// tests/test_spatial_hash.cpp:18-32, 49-73, 81-127, 133-167.
TEST_GPU(spatial_hash_grid_kats) {
  {
    Scoped s(100);
    UniformDistParams p;
    p.min_bounds = Vec3(-10, -10, -10);
    p.max_bounds = Vec3(10, 10, 10);
    ParticleInitializer::initUniform(s.h, p);
    s.up();
    SpatialHashGrid grid(100, 2.0f);
    grid.build(&s.d);
    CHECK(grid.getTotalCells() > 0);
    const int3 dims = grid.getGridDims();
    CHECK(dims.x * dims.y * dims.z == grid.getTotalCells());
    CHECK(grid.verifyCellAssignment(&s.h));
    for (float cell : {0.6f, 1.3f, 2.0f, 4.7f, 9.9f}) {  // each particle in exactly one cell
      SpatialHashGrid g(100, cell);
      g.build(&s.d);
      std::vector<int> cs, ce, pc, si;
      g.copyCellDataToHost(cs, ce, pc, si);
      std::vector<int> seen(100, 0);
      for (int c = 0; c < g.getTotalCells(); ++c)
        for (int k = cs[c]; k < ce[c]; ++k) seen[si[k]] += 1;
      for (int v : seen) CHECK(v == 1);
      CHECK(g.verifyCellAssignment(&s.h));
    }
  }
  {
    Scoped s(100);
    UniformDistParams p;
    p.min_bounds = Vec3(-5, -5, -5);
    p.max_bounds = Vec3(5, 5, 5);
    ParticleInitializer::initUniform(s.h, p);
    s.up();
    SpatialHashCalculator calc(2.0f, 2.0f);
    calc.setGravitationalConstant(1.0f);
    calc.setSofteningParameter(0.1f);
    calc.computeForces(&s.d);
    s.down();
    for (size_t i = 0; i < 100; ++i)
      CHECK(std::isfinite(s.h.acc_x[i]) && std::isfinite(s.h.acc_y[i]) &&
            std::isfinite(s.h.acc_z[i]));
    CHECK(calc.getGrid() != nullptr && calc.getGrid()->getCellSize() == 2.0f);
  }
  const float G = 1.0f, eps = 0.01f, cutoff = 2.0f, delta = 0.25f;
  for (int inside = 0; inside < 2; ++inside) {
    Scoped s(2);
    const float r = inside ? cutoff - delta : cutoff + delta;
    s.set(0, Vec3(0, 0, 0), Vec3(), 1.0f);
    s.set(1, Vec3(r, 0, 0), Vec3(), 1.0f);
    s.up();
    SpatialHashCalculator calc(2.0f, cutoff);
    calc.setGravitationalConstant(G);
    calc.setSofteningParameter(eps);
    calc.computeForces(&s.d);
    s.down();
    if (!inside) {
      CHECK(s.h.acc_x[0] == 0.0f && s.h.acc_y[0] == 0.0f && s.h.acc_z[0] == 0.0f);
      CHECK(s.h.acc_x[1] == 0.0f);
    } else {
      const float d2 = r * r + eps * eps;
      const float expected = G * 1.0f * r / (d2 * std::sqrt(d2));
      CHECK_NEAR(s.h.acc_x[0], expected, 1e-4f);
      CHECK_NEAR(s.h.acc_x[1], -expected, 1e-4f);
    }
  }
  {
    Scoped s(2);
    s.set(0, Vec3(-1e4f, 0, 0), Vec3(), 1.0f);
    s.set(1, Vec3(1e4f, 1e4f, 1e4f), Vec3(), 1.0f);
    s.up();
    SpatialHashGrid grid(2, 0.01f);
    bool refused = false;
    try {
      grid.build(&s.d);
    } catch (const std::runtime_error& e) {
      refused = std::string(e.what()).find("too large") != std::string::npos;
    }
    CHECK(refused);
  }
}

TEST_GPU(particle_system_pause_and_methods) {
  ParticleSystem ps;
  SimulationConfig c;
  c.particle_count = 50;
  ps.initialize(c);
  const SimulationState before = ps.getState();
  ps.pause();
  ps.update(0.01f);
  CHECK(ps.getState() == before);
  CHECK(ps.getSimulationTime() == 0.0f);
  ps.resume();
  ps.update(0.01f);
  CHECK(!(ps.getState() == before));
  ps.setForceMethod(ForceMethod::BARNES_HUT);
  ps.setBarnesHutTheta(0.4f);
  ps.update(0.01f);
  CHECK(ps.getForceMethod() == ForceMethod::BARNES_HUT);
  ps.setForceMethod(ForceMethod::SPATIAL_HASH);
  CHECK_THROWS(ps.setSpatialHashCutoff(5.0f), ValidationException);    // > cell size 2
  CHECK_THROWS(ps.setSpatialHashCellSize(1.0f), ValidationException);  // < cutoff 2
  ps.setSpatialHashCellSize(4.0f);
  ps.setSpatialHashCutoff(3.0f);
  ps.update(0.01f);
  CHECK(ps.getForceMethod() == ForceMethod::SPATIAL_HASH);
  SimulationState bad = ps.getState();
  bad.mass.pop_back();
  CHECK_THROWS(ps.setState(bad), ValidationException);
}

// Accelerations are not stored in checkpoints; a(0) is recomputed on load.
// With deterministic kernels a resumed run equals the uninterrupted one bit for bit.
TEST_GPU(checkpoint_exact_resume) {
  for (ForceMethod m : {ForceMethod::DIRECT_N2, ForceMethod::BARNES_HUT}) {
    ParticleSystem a;
    SimulationConfig c;
    c.particle_count = 3000;
    c.force_method = m;
    c.softening = 0.05f;
    a.initialize(c);
    for (int k = 0; k < 5; ++k) a.update(0.001f);
    const std::string path = "/tmp/nbx_resume_test.nbody";
    a.saveState(path);
    for (int k = 0; k < 7; ++k) a.update(0.001f);
    ParticleSystem b;
    b.initialize(c);
    b.loadState(path);
    for (int k = 0; k < 7; ++k) b.update(0.001f);
    const SimulationState sa = a.getState(), sb = b.getState();
    CHECK(std::memcmp(sa.pos_x.data(), sb.pos_x.data(), 4 * sa.particle_count) == 0);
    CHECK(std::memcmp(sa.vel_z.data(), sb.vel_z.data(), 4 * sa.particle_count) == 0);
    CHECK(std::fabs(sa.simulation_time - sb.simulation_time) < 1e-6f);
    std::remove(path.c_str());
  }
}

// The precision form behind ParticleSystem: a double-single run checkpointed as
// .nbody version 2 resumes bit for bit, low parts included.
TEST_GPU(precision_checkpoint_exact_resume) {
  ParticleSystem a;
  SimulationConfig c;
  c.particle_count = 3000;
  c.softening = 0.05f;
  a.setPrecisionMode(true);
  a.initialize(c);
  CHECK(a.getPrecisionMode());
  for (int k = 0; k < 5; ++k) a.update(0.001f);
  const std::string path = "/tmp/nbx_resume_ds.nbody";
  a.saveState(path);
  {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    CHECK((size_t)f.tellg() == 56 + 10 * 4 * c.particle_count);
  }
  for (int k = 0; k < 7; ++k) a.update(0.001f);
  ParticleSystem b;
  b.initialize(c);
  CHECK(!b.getPrecisionMode());
  b.loadState(path);
  CHECK(b.getPrecisionMode());
  for (int k = 0; k < 7; ++k) b.update(0.001f);
  const SimulationState sa = a.getState(), sb = b.getState();
  CHECK(sa.hasLowParts() && sb.hasLowParts());
  bool any_low = false;
  for (float v : sa.pos_lo_x) any_low = any_low || v != 0.0f;
  CHECK(any_low);
  const size_t bytes = 4 * sa.particle_count;
  CHECK(std::memcmp(sa.pos_x.data(), sb.pos_x.data(), bytes) == 0);
  CHECK(std::memcmp(sa.pos_lo_x.data(), sb.pos_lo_x.data(), bytes) == 0);
  CHECK(std::memcmp(sa.pos_lo_z.data(), sb.pos_lo_z.data(), bytes) == 0);
  CHECK(std::memcmp(sa.vel_y.data(), sb.vel_y.data(), bytes) == 0);
  std::remove(path.c_str());
}

TEST_GPU(precision_mode_rules) {
  ParticleSystem ps;
  SimulationConfig c;
  c.particle_count = 500;
  ps.initialize(c);
  const float e_plain = ps.computeTotalEnergy();
  ps.setPrecisionMode(true);
  CHECK(std::fabs(ps.computeTotalEnergy() - e_plain) <= 1e-5f * std::fabs(e_plain));
  CHECK_THROWS(ps.setForceMethod(ForceMethod::BARNES_HUT), ValidationException);
  for (int k = 0; k < 3; ++k) ps.update(0.001f);
  const SimulationState ds = ps.getState();
  ps.setPrecisionMode(false);  // hi + lo rounded to FP32
  const SimulationState plain = ps.getState();
  CHECK(!plain.hasLowParts());
  for (size_t i = 0; i < 500; ++i)
    CHECK(plain.pos_y[i] == (float)((double)ds.pos_y[i] + (double)ds.pos_lo_y[i]));
  ps.setForceMethod(ForceMethod::BARNES_HUT);
  CHECK_THROWS(ps.setPrecisionMode(true), ValidationException);
  SimulationState bh_ds = ds;
  bh_ds.force_method = ForceMethod::BARNES_HUT;
  CHECK_THROWS(ps.setState(bh_ds), ValidationException);
}

}  // namespace

int main(int argc, char** argv) {
  const bool want_gpu = argc > 1 && std::string(argv[1]) == "--gpu";
  int failed = 0, ran = 0;
  for (const Case& c : registry()) {
    if (c.gpu != want_gpu) continue;
    ++ran;
    try {
      c.fn();
      std::cout << "PASS " << c.name << std::endl;
    } catch (const std::exception& e) {
      ++failed;
      std::cout << "FAIL " << c.name << ": " << e.what() << std::endl;
    }
  }
  std::cout << ran - failed << "/" << ran << " passed" << std::endl;
  return failed ? 1 : 0;
}
