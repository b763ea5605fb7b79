// Exception types and input validators of the engine's C++ API, mirroring the
// reference's include/nbody/error_handling.hpp (build-workbench/nbody-gpu-sim).
// GPU runtime failures raise DeviceException (CudaException is an alias kept
// for code that catches the reference's name).
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>

namespace nbody {

class DeviceException : public std::runtime_error {
 public:
  DeviceException(const std::string& msg, const char* file, int line)
      : std::runtime_error("GPU Error: " + msg + " at " + file + ":" + std::to_string(line)),
        file_(file),
        line_(line) {}
  explicit DeviceException(const std::string& msg) : std::runtime_error("GPU Error: " + msg) {}
  const char* file() const noexcept { return file_; }
  int line() const noexcept { return line_; }

 private:
  const char* file_ = "";
  int line_ = 0;
};
using CudaException = DeviceException;

class ResourceException : public std::runtime_error {
 public:
  ResourceException(const std::string& msg, size_t requested, size_t available)
      : std::runtime_error("Resource Error: " + msg + " (requested: " + std::to_string(requested) +
                           " bytes, available: " + std::to_string(available) + " bytes)"),
        requested_(requested),
        available_(available) {}
  size_t getRequestedBytes() const noexcept { return requested_; }
  size_t getAvailableBytes() const noexcept { return available_; }

 private:
  size_t requested_;
  size_t available_;
};

class ValidationException : public std::runtime_error {
 public:
  explicit ValidationException(const std::string& msg)
      : std::runtime_error("Validation Error: " + msg) {}
};

class IOException : public std::runtime_error {
 public:
  explicit IOException(const std::string& msg) : std::runtime_error("I/O Error: " + msg) {}
};

// Turns a non-zero nbx status (include/nbx.h) into the matching exception.
void throwOnStatus(int status, const char* file, int line);
#define NBODY_CHECK(call) ::nbody::throwOnStatus((call), __FILE__, __LINE__)

struct SimulationConfig;
void validateResourceRequirements(size_t particle_count);
void validateSimulationConfig(const SimulationConfig& config);
void validateParticleCount(size_t count);
void validateParticleCountRange(size_t count);
void validateTimeStep(float dt);
void validateSoftening(float eps);
void validateTheta(float theta);

}  // namespace nbody
