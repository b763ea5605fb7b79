// Phase profiler and benchmark JSON records (behaviour and schema of the
// reference's src/utils/performance_observability.cpp:50-160).
#include <cstdio>
#include <fstream>
#include <sstream>

#include "nbody/error_handling.hpp"
#include "nbody/performance_observability.hpp"

namespace nbody {

namespace {

std::string escapeJson(const std::string& s) {
  std::string o;
  for (char c : s) {
    switch (c) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\t': o += "\\t"; break;
      default:
        if ((unsigned char)c < 0x20) {
          char buf[8];
          snprintf(buf, sizeof buf, "\\u%04x", c);
          o += buf;
        } else {
          o += c;
        }
    }
  }
  return o;
}

void numberMap(std::ostringstream& out, const std::map<std::string, double>& m) {
  out << "{";
  bool first = true;
  for (const auto& kv : m) {
    if (!first) out << ",";
    first = false;
    out << "\"" << escapeJson(kv.first) << "\":" << kv.second;
  }
  out << "}";
}

PhaseProfiler g_profiler;

}  // namespace

void PhaseProfiler::record(std::string_view name, Milliseconds duration) {
  std::lock_guard<std::mutex> lock(mutex_);
  for (auto& p : phases_)
    if (p.name == name) {
      p.total_duration += duration;
      p.samples += 1;
      return;
    }
  PhaseTiming p;
  p.name = std::string(name);
  p.total_duration = duration;
  p.samples = 1;
  phases_.push_back(std::move(p));
}

std::vector<PhaseTiming> PhaseProfiler::snapshot() const {
  std::lock_guard<std::mutex> lock(mutex_);
  return phases_;
}

void PhaseProfiler::reset() {
  std::lock_guard<std::mutex> lock(mutex_);
  phases_.clear();
}

std::vector<PhaseTiming> PhaseProfiler::consume() {
  std::lock_guard<std::mutex> lock(mutex_);
  std::vector<PhaseTiming> out;
  out.swap(phases_);
  return out;
}

ScopedPhaseProfile::ScopedPhaseProfile(PhaseProfiler& profiler, std::string name)
    : profiler_(profiler), name_(std::move(name)), start_time_(std::chrono::steady_clock::now()) {}

ScopedPhaseProfile::~ScopedPhaseProfile() {
  profiler_.record(name_, std::chrono::duration_cast<Milliseconds>(
                              std::chrono::steady_clock::now() - start_time_));
}

PhaseProfiler& globalPhaseProfiler() { return g_profiler; }

std::vector<PhaseTiming> consumeGlobalPhaseSnapshot() { return g_profiler.consume(); }

std::string serializeBenchmarkRunRecord(const BenchmarkRunRecord& r) {
  std::ostringstream out;
  out.precision(17);
  out << "{\"benchmark_name\":\"" << escapeJson(r.benchmark_name) << "\",";
  out << "\"force_method\":\"" << forceMethodToString(r.force_method) << "\",";
  out << "\"particle_count\":" << r.particle_count << ",";
  out << "\"iterations\":" << r.iterations << ",";
  out << "\"metrics\":";
  numberMap(out, r.metrics);
  out << ",\"parameters\":";
  numberMap(out, r.parameters);
  out << ",\"phase_timings\":[";
  for (size_t i = 0; i < r.phase_timings.size(); ++i) {
    const auto& p = r.phase_timings[i];
    if (i) out << ",";
    out << "{\"name\":\"" << escapeJson(p.name) << "\",\"total_duration_ms\":"
        << p.total_duration.count() << ",\"samples\":" << p.samples << "}";
  }
  out << "]}";
  return out.str();
}

std::string serializeBenchmarkRunRecords(const std::vector<BenchmarkRunRecord>& records) {
  std::ostringstream out;
  out << "{\"benchmarks\":[";
  for (size_t i = 0; i < records.size(); ++i) {
    if (i) out << ",";
    out << serializeBenchmarkRunRecord(records[i]);
  }
  out << "]}";
  return out.str();
}

void writeBenchmarkRunRecords(const std::string& path,
                              const std::vector<BenchmarkRunRecord>& records) {
  std::ofstream f(path);
  if (!f) throw IOException("Failed to open benchmark output file: " + path);
  f << serializeBenchmarkRunRecords(records) << "\n";
}

}  // namespace nbody
