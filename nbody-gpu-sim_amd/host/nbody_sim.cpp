// Headless front-end: the reference's --benchmark mode (src/main.cpp:347-401)
// without a window: initialise (or --import a checkpoint), run the Velocity
// Verlet steps on the GPU, optionally --export, and print the benchmark JSON
// with per-phase GPU times.
#include <chrono>
#include <iostream>

#include "nbody/app_cli.hpp"
#include "nbody/error_handling.hpp"
#include "nbody/particle_system.hpp"
#include "nbody/performance_observability.hpp"
#include "nbody/serialization.hpp"
#include "nbx.h"

using namespace nbody;

int main(int argc, char* argv[]) {
  try {
    const AppCliOptions opt = parseAppCliOptions(argc, argv);
    if (opt.show_help) {
      std::cout << appCliUsage();
      return 0;
    }
    SimulationConfig config;
    config.particle_count = opt.particle_count;
    config.force_method = opt.force_method;
    config.dt = opt.dt;
    config.G = opt.G;
    config.softening = opt.softening;
    config.barnes_hut_theta = opt.barnes_hut_theta;
    config.spatial_hash_cell_size = opt.spatial_hash_cell_size;
    config.spatial_hash_cutoff = opt.spatial_hash_cutoff;
    ParticleSystem system;
    system.initialize(config);
    if (!opt.import_path.empty()) {
      std::cout << "Importing state from: " << opt.import_path << std::endl;
      SimulationState st = Serializer::load(opt.import_path);
      system.setState(st);
      std::cout << "Imported " << st.particle_count << " particles" << std::endl;
    }
    const size_t steps = opt.benchmark_mode ? opt.benchmark_steps : 1;
    NBODY_CHECK(nbx_synchronize());
    const auto t0 = std::chrono::steady_clock::now();
    for (size_t s = 0; s < steps; ++s) system.update(system.getTimeStep());
    NBODY_CHECK(nbx_synchronize());
    const auto t1 = std::chrono::steady_clock::now();
    if (!opt.export_path.empty()) {
      const SimulationState st = system.getState();
      std::cout << "Exporting checkpoint to: " << opt.export_path << std::endl;
      Serializer::save(opt.export_path, st);
      std::cout << "Exported " << st.particle_count << " particles" << std::endl;
    }
    BenchmarkRunRecord r;
    r.benchmark_name = "application.benchmark_mode";
    r.force_method = system.getForceMethod();
    r.particle_count = system.getParticleCount();
    r.iterations = steps;
    r.metrics["wall_time_ms"] =
        std::chrono::duration<double, std::milli>(t1 - t0).count() / (double)steps;
    r.parameters["dt"] = system.getTimeStep();
    r.parameters["gravity"] = system.getGravitationalConstant();
    r.parameters["softening"] = system.getSofteningParameter();
    r.parameters["particle_count"] = (double)system.getParticleCount();
    if (system.getForceMethod() == ForceMethod::BARNES_HUT) {
      r.parameters["theta"] = config.barnes_hut_theta;
    } else if (system.getForceMethod() == ForceMethod::SPATIAL_HASH) {
      r.parameters["cell_size"] = config.spatial_hash_cell_size;
      r.parameters["cutoff_radius"] = config.spatial_hash_cutoff;
    }
    const std::vector<BenchmarkRunRecord> recs{r};
    std::cout << serializeBenchmarkRunRecords(recs) << std::endl;
    if (!opt.benchmark_output_path.empty()) writeBenchmarkRunRecords(opt.benchmark_output_path, recs);
    return 0;
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 1;
  }
}
