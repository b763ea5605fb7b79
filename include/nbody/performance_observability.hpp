// Benchmark records of the headless front-end, with the JSON schema of the
// reference's performance_observability (build-workbench/nbody-gpu-sim):
// {"benchmarks":[{benchmark_name, force_method, particle_count, iterations,
//   metrics{}, parameters{}, phase_timings[{name,total_duration_ms,samples}]}]}
#pragma once

#include <cstddef>
#include <map>
#include <string>
#include <vector>

#include "nbody/types.hpp"

namespace nbody {

struct PhaseTimingSnapshot {
  std::string name;
  double total_duration_ms = 0.0;
  size_t samples = 0;
};

struct BenchmarkRunRecord {
  std::string benchmark_name;
  ForceMethod force_method = ForceMethod::DIRECT_N2;
  size_t particle_count = 0;
  size_t iterations = 0;
  std::map<std::string, double> metrics;
  std::map<std::string, double> parameters;
  std::vector<PhaseTimingSnapshot> phase_timings;
};

std::string serializeBenchmarkRunRecords(const std::vector<BenchmarkRunRecord>& records);
void writeBenchmarkRunRecords(const std::string& path,
                              const std::vector<BenchmarkRunRecord>& records);

}  // namespace nbody
