// Phase profiling and benchmark records of the engine's C++ API, mirroring the
// reference's include/nbody/performance_observability.hpp
// (build-workbench/nbody-gpu-sim). JSON schema:
// {"benchmarks":[{benchmark_name, force_method, particle_count, iterations,
//   metrics{}, parameters{}, phase_timings[{name,total_duration_ms,samples}]}]}
// NBODY_PROFILE_SCOPE records host wall time of a scope into the global
// profiler when the library is built with NBODY_ENABLE_PROFILING=1 (off when
// undefined, as in the reference; this package's Makefile defines it to 1).
// GPU work inside a scope is asynchronous unless the scope waits for it.
#pragma once

#include <chrono>
#include <cstddef>
#include <map>
#include <mutex>
#include <string>
#include <string_view>
#include <vector>

#include "nbody/types.hpp"

namespace nbody {

using Milliseconds = std::chrono::duration<double, std::milli>;

struct PhaseTiming {
  std::string name;
  Milliseconds total_duration{0.0};
  size_t samples = 0;
};

// Thread-safe accumulation of named phase durations.
class PhaseProfiler {
 public:
  void record(std::string_view name, Milliseconds duration);
  std::vector<PhaseTiming> snapshot() const;
  void reset();
  // Snapshot and clear in one step (records between a separate snapshot() and
  // reset() would be lost).
  std::vector<PhaseTiming> consume();

 private:
  mutable std::mutex mutex_;
  std::vector<PhaseTiming> phases_;
};

class ScopedPhaseProfile {
 public:
  ScopedPhaseProfile(PhaseProfiler& profiler, std::string name);
  ~ScopedPhaseProfile();
  ScopedPhaseProfile(const ScopedPhaseProfile&) = delete;
  ScopedPhaseProfile& operator=(const ScopedPhaseProfile&) = delete;

 private:
  PhaseProfiler& profiler_;
  std::string name_;
  std::chrono::steady_clock::time_point start_time_;
};

struct BenchmarkRunRecord {
  std::string benchmark_name;
  ForceMethod force_method = ForceMethod::DIRECT_N2;
  size_t particle_count = 0;
  size_t iterations = 0;
  std::map<std::string, double> metrics;
  std::map<std::string, double> parameters;
  std::vector<PhaseTiming> phase_timings;
};

std::string serializeBenchmarkRunRecord(const BenchmarkRunRecord& record);
std::string serializeBenchmarkRunRecords(const std::vector<BenchmarkRunRecord>& records);
void writeBenchmarkRunRecords(const std::string& path,
                              const std::vector<BenchmarkRunRecord>& records);

PhaseProfiler& globalPhaseProfiler();
std::vector<PhaseTiming> consumeGlobalPhaseSnapshot();

#if defined(NBODY_ENABLE_PROFILING) && NBODY_ENABLE_PROFILING
#define NBODY_PROFILE_CAT2(a, b) a##b
#define NBODY_PROFILE_CAT(a, b) NBODY_PROFILE_CAT2(a, b)
#define NBODY_PROFILE_SCOPE(name)                                                  \
  ::nbody::ScopedPhaseProfile NBODY_PROFILE_CAT(nbody_scoped_phase_, __LINE__)( \
      ::nbody::globalPhaseProfiler(), (name))
#else
#define NBODY_PROFILE_SCOPE(name) \
  do {                            \
  } while (0)
#endif

}  // namespace nbody
