// Force-method operators of the engine's C++ API, mirroring the reference's
// include/nbody/force_calculator.hpp (build-workbench/nbody-gpu-sim). Each
// computeForces call runs HIP kernels through the nbx C ABI on the calling
// thread's null stream, asynchronously.
#pragma once

#include <memory>

#include "nbody/types.hpp"

struct nbx_direct;  // workspace handle of the C ABI (nbx.h)

namespace nbody {

class ForceCalculator {
 public:
  virtual ~ForceCalculator() = default;
  virtual void computeForces(ParticleData* d_particles) = 0;
  virtual ForceMethod getMethod() const = 0;

  void setSofteningParameter(float eps) {
    softening_eps_ = eps;
    softening_eps2_ = eps * eps;
  }
  void setGravitationalConstant(float G) { G_ = G; }
  float getSofteningParameter() const noexcept { return softening_eps_; }
  float getGravitationalConstant() const noexcept { return G_; }

 protected:
  float softening_eps_ = 0.1f;
  float softening_eps2_ = 0.01f;
  float G_ = 1.0f;
};

// Owns the device workspace of the pair-symmetric kernel. It is allocated by
// reserve(), or by the first computeForces call (and again by a call on more
// bodies than it holds); other calls allocate nothing.
class DirectForceCalculator : public ForceCalculator {
 public:
  explicit DirectForceCalculator(int block_size = 256);
  ~DirectForceCalculator() override;
  DirectForceCalculator(const DirectForceCalculator&) = delete;
  DirectForceCalculator& operator=(const DirectForceCalculator&) = delete;
  void computeForces(ParticleData* d_particles) override;
  // computeForces with the Verlet kick v += (a_old + a) dt/2 fused (Integrator::integrate)
  void computeForcesKick(ParticleData* d_particles, float dt);
  void reserve(size_t max_particles);
  ForceMethod getMethod() const noexcept override { return ForceMethod::DIRECT_N2; }
  void setBlockSize(int size) { block_size_ = size; }
  int getBlockSize() const noexcept { return block_size_; }

 private:
  int block_size_;
  nbx_direct* workspace_ = nullptr;
  size_t capacity_ = 0;
};

class BarnesHutCalculator : public ForceCalculator {
 public:
  explicit BarnesHutCalculator(float theta = 0.5f);
  ~BarnesHutCalculator();
  void computeForces(ParticleData* d_particles) override;
  ForceMethod getMethod() const noexcept override { return ForceMethod::BARNES_HUT; }
  void setTheta(float theta) { theta_ = theta; }
  float getTheta() const noexcept { return theta_; }
  BarnesHutTree* getTree() noexcept { return tree_.get(); }

 private:
  std::unique_ptr<BarnesHutTree> tree_;
  float theta_;
};

// Short-range forces over a uniform cell grid. cell_size must be >= cutoff_radius
// for the 3x3x3 neighbour search to be complete; computeForces uses
// max(cell_size, cutoff_radius) as the grid's cell either way.
class SpatialHashCalculator : public ForceCalculator {
 public:
  SpatialHashCalculator(float cell_size = 2.0f, float cutoff_radius = 2.0f);
  ~SpatialHashCalculator();
  void computeForces(ParticleData* d_particles) override;
  ForceMethod getMethod() const noexcept override { return ForceMethod::SPATIAL_HASH; }
  void setCellSize(float size) { cell_size_ = size; }
  void setCutoffRadius(float radius) { cutoff_radius_ = radius; }
  float getCellSize() const noexcept { return cell_size_; }
  float getCutoffRadius() const noexcept { return cutoff_radius_; }
  SpatialHashGrid* getGrid() noexcept { return grid_.get(); }

 private:
  std::unique_ptr<SpatialHashGrid> grid_;
  float cell_size_;
  float cutoff_radius_;
};

std::unique_ptr<ForceCalculator> createForceCalculator(ForceMethod method,
                                                       const SimulationConfig& config);

// Host pair function: the acceleration on a particle at p1 due to mass m2 at p2
// (Plummer-softened), evaluated in IEEE FP32 like the reference's CPU reference.
Vec3 computeGravitationalAccelerationCPU(const Vec3& p1, const Vec3& p2, float m2, float G,
                                         float eps);

}  // namespace nbody
