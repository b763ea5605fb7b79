// Command-line parsing of the headless front-end (reference: src/core/app_cli.cpp:11-161).
// `direct` is accepted as an alias of direct-n2.
#include <sstream>
#include <stdexcept>

#include "nbody/app_cli.hpp"
#include "nbody/error_handling.hpp"

namespace nbody {

namespace {

ForceMethod parseForceMethod(const std::string& v) {
  if (v == "direct" || v == "direct-n2" || v == "direct_n2") return ForceMethod::DIRECT_N2;
  if (v == "barnes-hut" || v == "barnes_hut") return ForceMethod::BARNES_HUT;
  if (v == "spatial-hash" || v == "spatial_hash") return ForceMethod::SPATIAL_HASH;
  throw ValidationException("Unsupported force method: " + v);
}

std::string requireValue(int argc, const char* const argv[], int& i, const std::string& flag) {
  if (i + 1 >= argc) throw ValidationException("Missing value for " + flag);
  return argv[++i];
}

size_t parseSize(const std::string& v, const std::string& flag) {
  try {
    if (!v.empty() && v[0] == '-') throw std::invalid_argument("negative");
    size_t pos = 0;
    const unsigned long long r = std::stoull(v, &pos);
    if (pos != v.size()) throw std::invalid_argument("trailing");
    return static_cast<size_t>(r);
  } catch (const std::exception&) {
    throw ValidationException("Invalid numeric value for " + flag + ": " + v);
  }
}

float parseFloat(const std::string& v, const std::string& flag) {
  try {
    size_t pos = 0;
    const float r = std::stof(v, &pos);
    if (pos != v.size()) throw std::invalid_argument("trailing");
    return r;
  } catch (const std::exception&) {
    throw ValidationException("Invalid numeric value for " + flag + ": " + v);
  }
}

}  // namespace

AppCliOptions parseAppCliOptions(int argc, const char* const argv[]) {
  AppCliOptions o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--help" || a == "-h") {
      o.show_help = true;
    } else if (a == "--particles") {
      o.particle_count = parseSize(requireValue(argc, argv, i, a), a);
    } else if (a == "--method") {
      o.force_method = parseForceMethod(requireValue(argc, argv, i, a));
    } else if (a == "--dt") {
      o.dt = parseFloat(requireValue(argc, argv, i, a), a);
    } else if (a == "--gravity") {
      o.G = parseFloat(requireValue(argc, argv, i, a), a);
    } else if (a == "--softening") {
      o.softening = parseFloat(requireValue(argc, argv, i, a), a);
    } else if (a == "--theta") {
      o.barnes_hut_theta = parseFloat(requireValue(argc, argv, i, a), a);
    } else if (a == "--cell-size") {
      o.spatial_hash_cell_size = parseFloat(requireValue(argc, argv, i, a), a);
    } else if (a == "--cutoff") {
      o.spatial_hash_cutoff = parseFloat(requireValue(argc, argv, i, a), a);
    } else if (a == "--benchmark") {
      o.benchmark_mode = true;
    } else if (a == "--benchmark-steps") {
      o.benchmark_steps = parseSize(requireValue(argc, argv, i, a), a);
      o.benchmark_mode = true;
    } else if (a == "--benchmark-output") {
      o.benchmark_output_path = requireValue(argc, argv, i, a);
      o.benchmark_mode = true;
    } else if (a == "--export") {
      o.export_path = requireValue(argc, argv, i, a);
    } else if (a == "--import") {
      o.import_path = requireValue(argc, argv, i, a);
    } else if (!a.empty() && a.front() == '-') {
      throw ValidationException("Unknown argument: " + a);
    } else {
      o.particle_count = parseSize(a, "particle count");
    }
  }
  validateParticleCountRange(o.particle_count);
  validateTimeStep(o.dt);
  validateSoftening(o.softening);
  validateTheta(o.barnes_hut_theta);
  if (o.G <= 0.0f) throw ValidationException("Gravitational constant must be positive");
  if (o.spatial_hash_cell_size <= 0.0f)
    throw ValidationException("Spatial hash cell size must be positive");
  if (o.spatial_hash_cutoff <= 0.0f) throw ValidationException("Spatial hash cutoff must be positive");
  if (o.benchmark_steps == 0) throw ValidationException("Benchmark steps must be greater than zero");
  return o;
}

std::string appCliUsage() {
  std::ostringstream u;
  u << "Usage: nbody_sim [particle_count] [options]\n\n"
    << "Simulation options:\n"
    << "  --particles N          Set particle count\n"
    << "  --method NAME          direct | direct-n2 | barnes-hut\n"
    << "  --dt VALUE             Set integration time step\n"
    << "  --gravity VALUE        Set gravitational constant\n"
    << "  --softening VALUE      Set softening parameter\n"
    << "  --theta VALUE          Set Barnes-Hut theta\n"
    << "  --benchmark            Run a non-interactive benchmark and exit\n"
    << "  --benchmark-steps N    Set benchmark update steps\n"
    << "  --benchmark-output P   Write benchmark JSON to path P\n"
    << "\nData export/import:\n"
    << "  --export PATH          Export particle state to checkpoint file\n"
    << "  --import PATH          Import particle state from checkpoint file\n"
    << "\n  --help                 Show this message\n";
  return u.str();
}

}  // namespace nbody
