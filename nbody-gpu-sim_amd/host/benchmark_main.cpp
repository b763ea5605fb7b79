// nbody_benchmarks: the reference's benchmark suite (benchmarks/benchmark_main.cpp of
// build-workbench/nbody-gpu-sim) on the GPU engine. Same command line
// (--benchmark <name|all>, --particle-count N, --iterations N, --output path,
// --list), benchmark names, parameters and JSON records:
//   serialization.round_trip     .nbody v1 save + load through a memory stream
//   force.direct_n2              DirectForceCalculator::computeForces
//   force.barnes_hut             BarnesHutCalculator::computeForces (tree build + walk)
//   force.spatial_hash           SpatialHashCalculator::computeForces (grid build + forces)
//   integration.velocity_verlet  ParticleSystem::update (one Velocity Verlet step)
// GPU benchmarks run 2 untimed warm-up calls, then time each call to the end of
// its GPU work (wall clock around the call and a device synchronisation).
#include <chrono>
#include <cstdlib>
#include <exception>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "nbody/barnes_hut_tree.hpp"
#include "nbody/error_handling.hpp"
#include "nbody/force_calculator.hpp"
#include "nbody/particle_system.hpp"
#include "nbody/performance_observability.hpp"
#include "nbody/serialization.hpp"
#include "nbx.h"

using namespace nbody;

namespace {

struct BenchmarkOptions {
  std::string benchmark_name = "all";
  size_t particle_count = 4096;
  size_t iterations = 5;
  std::string output_path;
};

struct BenchmarkDefinition {
  std::string name;
  std::string description;
  bool requires_gpu;
  BenchmarkRunRecord (*run)(const BenchmarkOptions&);
};

void synchronizeDevice() { NBODY_CHECK(nbx_synchronize()); }

template <typename F>
double timedCalls(F&& call, size_t iterations) {
  for (int w = 0; w < 2; ++w) call();  // first calls allocate operator state
  synchronizeDevice();
  double total = 0.0;
  for (size_t it = 0; it < iterations; ++it) {
    const auto t0 = std::chrono::steady_clock::now();
    call();
    synchronizeDevice();
    total += Milliseconds(std::chrono::steady_clock::now() - t0).count();
  }
  return total / (double)iterations;
}

// Deterministic state of the serialization benchmark (benchmark_main.cpp:32-58).
SimulationState makeState(size_t n) {
  SimulationState s;
  s.particle_count = n;
  s.simulation_time = 1.0f;
  s.dt = 0.001f;
  s.G = 1.0f;
  s.softening = 0.1f;
  s.force_method = ForceMethod::DIRECT_N2;
  s.pos_x.resize(n);
  s.pos_y.resize(n);
  s.pos_z.resize(n);
  s.vel_x.resize(n);
  s.vel_y.resize(n);
  s.vel_z.resize(n);
  s.mass.assign(n, 1.0f);
  for (size_t i = 0; i < n; ++i) {
    s.pos_x[i] = (float)(i % 97) * 0.01f;
    s.pos_y[i] = (float)((i * 3) % 89) * 0.02f;
    s.pos_z[i] = (float)((i * 7) % 83) * 0.03f;
    s.vel_x[i] = 0.001f * (float)(i % 11);
    s.vel_y[i] = 0.002f * (float)(i % 13);
    s.vel_z[i] = 0.003f * (float)(i % 17);
  }
  return s;
}

BenchmarkRunRecord baseRecord(const BenchmarkOptions& o, const char* name, ForceMethod m) {
  BenchmarkRunRecord r;
  r.benchmark_name = name;
  r.force_method = m;
  r.particle_count = o.particle_count;
  r.iterations = o.iterations;
  r.parameters["particle_count"] = (double)o.particle_count;
  return r;
}

BenchmarkRunRecord runSerialization(const BenchmarkOptions& o) {
  BenchmarkRunRecord r = baseRecord(o, "serialization.round_trip", ForceMethod::DIRECT_N2);
  const SimulationState state = makeState(o.particle_count);
  consumeGlobalPhaseSnapshot();
  double total_ms = 0.0;
  size_t total_bytes = 0;
  for (size_t it = 0; it < o.iterations; ++it) {
    std::stringstream stream(std::ios::in | std::ios::out | std::ios::binary);
    const auto t0 = std::chrono::steady_clock::now();
    Serializer::save(stream, state);
    total_bytes += (size_t)stream.tellp();
    stream.seekg(0);
    const SimulationState loaded = Serializer::load(stream);
    total_ms += Milliseconds(std::chrono::steady_clock::now() - t0).count();
    if (loaded.particle_count != state.particle_count)
      throw std::runtime_error("Serialization benchmark round-trip lost particle data");
  }
  r.metrics["wall_time_ms"] = total_ms / (double)o.iterations;
  r.metrics["bytes_per_iteration"] = (double)total_bytes / (double)o.iterations;
  r.phase_timings = consumeGlobalPhaseSnapshot();
  return r;
}

// runForceBenchmark and runBarnesHutBenchmark (benchmark_main.cpp:114-156, :199-244): spherical initial state, the
// factory's calculator, computeForces timed per call.
BenchmarkRunRecord runForce(const BenchmarkOptions& o, ForceMethod method, const char* name) {
  SimulationConfig config;
  config.particle_count = o.particle_count;
  config.force_method = method;
  config.init_distribution = InitDistribution::SPHERICAL;
  ParticleSystem system;
  system.initialize(config);
  BenchmarkRunRecord r = baseRecord(o, name, method);
  if (method == ForceMethod::BARNES_HUT) {  // the Barnes-Hut record has no block size (:210-214)
    r.parameters["theta"] = config.barnes_hut_theta;
  } else {
    r.parameters["cuda_block_size"] = (double)config.cuda_block_size;
    if (method == ForceMethod::SPATIAL_HASH) {
      r.parameters["cell_size"] = config.spatial_hash_cell_size;
      r.parameters["cutoff_radius"] = config.spatial_hash_cutoff;
    }
  }
  auto calc = createForceCalculator(method, config);
  consumeGlobalPhaseSnapshot();
  r.metrics["wall_time_ms"] =
      timedCalls([&] { calc->computeForces(system.getDeviceData()); }, o.iterations);
  r.phase_timings = consumeGlobalPhaseSnapshot();
  if (method == ForceMethod::BARNES_HUT) {  // :236-241, plus the GPU phases of one more call
    for (const auto& p : r.phase_timings)
      if (p.samples > 0) r.metrics[p.name + "_ms"] = p.total_duration.count() / (double)p.samples;
    auto* bh = dynamic_cast<BarnesHutCalculator*>(calc.get());
    if (bh && bh->getTree()) {
      nbx_bh* t = bh->getTree()->handle();
      NBODY_CHECK(nbx_bh_set_timing(t, 1));
      calc->computeForces(system.getDeviceData());
      synchronizeDevice();
      double ms[8];
      NBODY_CHECK(nbx_bh_phase_ms(t, ms));
      NBODY_CHECK(nbx_bh_set_timing(t, 0));
      const char* names[8] = {"bbox", "morton", "sort", "paths", "tree", "com", "export", "forces"};
      for (int k = 0; k < 8; ++k) r.metrics[std::string("gpu.") + names[k] + "_ms"] = ms[k];
      consumeGlobalPhaseSnapshot();
    }
  }
  return r;
}

BenchmarkRunRecord runDirect(const BenchmarkOptions& o) {
  return runForce(o, ForceMethod::DIRECT_N2, "force.direct_n2");
}
BenchmarkRunRecord runBarnesHut(const BenchmarkOptions& o) {
  return runForce(o, ForceMethod::BARNES_HUT, "force.barnes_hut");
}
BenchmarkRunRecord runSpatialHash(const BenchmarkOptions& o) {
  return runForce(o, ForceMethod::SPATIAL_HASH, "force.spatial_hash");
}

// runIntegrationBenchmark (:158-193): one ParticleSystem::update per call.
BenchmarkRunRecord runIntegration(const BenchmarkOptions& o) {
  SimulationConfig config;
  config.particle_count = o.particle_count;
  config.force_method = ForceMethod::DIRECT_N2;
  config.init_distribution = InitDistribution::SPHERICAL;
  ParticleSystem system;
  system.initialize(config);
  BenchmarkRunRecord r = baseRecord(o, "integration.velocity_verlet", config.force_method);
  r.parameters["dt"] = config.dt;
  r.parameters["cuda_block_size"] = (double)config.cuda_block_size;
  consumeGlobalPhaseSnapshot();
  r.metrics["wall_time_ms"] = timedCalls([&] { system.update(system.getTimeStep()); }, o.iterations);
  r.phase_timings = consumeGlobalPhaseSnapshot();
  return r;
}

std::vector<BenchmarkDefinition> makeBenchmarks() {
  return {
      {"serialization.round_trip", "Binary checkpoint serialization and load round-trip", false,
       runSerialization},
      {"force.direct_n2", "Direct N^2 force calculation", true, runDirect},
      {"force.barnes_hut", "Barnes-Hut force calculation", true, runBarnesHut},
      {"force.spatial_hash", "Spatial hash force calculation", true, runSpatialHash},
      {"integration.velocity_verlet", "Velocity Verlet integration step", true, runIntegration},
  };
}

void printUsage(const std::vector<BenchmarkDefinition>& benchmarks) {
  std::cout << "Usage: nbody_benchmarks [--benchmark <name|all>] [--particle-count N]\n"
               "                        [--iterations N] [--output path] [--list]\n\n"
               "Available benchmarks:\n";
  for (const auto& b : benchmarks)
    std::cout << "  - " << b.name << (b.requires_gpu ? " (GPU)" : "") << ": " << b.description
              << "\n";
}

BenchmarkOptions parseOptions(int argc, char* argv[],
                              const std::vector<BenchmarkDefinition>& benchmarks) {
  BenchmarkOptions o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--benchmark" && i + 1 < argc) {
      o.benchmark_name = argv[++i];
    } else if (a == "--particle-count" && i + 1 < argc) {
      o.particle_count = (size_t)std::stoull(argv[++i]);
    } else if (a == "--iterations" && i + 1 < argc) {
      o.iterations = (size_t)std::stoull(argv[++i]);
    } else if (a == "--output" && i + 1 < argc) {
      o.output_path = argv[++i];
    } else if (a == "--list" || a == "--help" || a == "-h") {
      printUsage(benchmarks);
      std::exit(0);
    } else {
      throw ValidationException("Unknown benchmark argument: " + a);
    }
  }
  if (o.iterations == 0) throw ValidationException("Benchmark iterations must be greater than zero");
  return o;
}

}  // namespace

int main(int argc, char* argv[]) {
  try {
    const auto benchmarks = makeBenchmarks();
    const BenchmarkOptions o = parseOptions(argc, argv, benchmarks);
    std::vector<BenchmarkRunRecord> results;
    for (const auto& b : benchmarks)
      if (o.benchmark_name == "all" || o.benchmark_name == b.name) results.push_back(b.run(o));
    if (results.empty()) throw ValidationException("Requested benchmark was not found");
    std::cout << serializeBenchmarkRunRecords(results) << "\n";
    if (!o.output_path.empty()) writeBenchmarkRunRecords(o.output_path, results);
  } catch (const std::exception& e) {
    std::cerr << "Benchmark error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
