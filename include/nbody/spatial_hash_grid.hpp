// Spatial-hash grid of the engine's C++ API, mirroring the reference's
// include/nbody/spatial_hash_grid.hpp (build-workbench/nbody-gpu-sim). build()
// and computeForces() run on the GPU through the nbx_grid_* entry points of
// nbx.h; cell numbering cx + cy*dims.x + cz*dims.x*dims.y.
#pragma once

#include <cstddef>
#include <vector>

#include "nbody/types.hpp"

struct nbx_grid;

namespace nbody {

class SpatialHashGrid {
 public:
  SpatialHashGrid(size_t max_particles, float cell_size = 1.0f);
  ~SpatialHashGrid();
  SpatialHashGrid(const SpatialHashGrid&) = delete;
  SpatialHashGrid& operator=(const SpatialHashGrid&) = delete;
  SpatialHashGrid(SpatialHashGrid&&) = delete;
  SpatialHashGrid& operator=(SpatialHashGrid&&) = delete;

  // Bounding box, cell of every particle and the particles grouped by cell
  // (ascending index within a cell). Throws std::runtime_error above 1e8 cells.
  void build(const ParticleData* d_particles);
  // Short-range forces: j != i in the 27 cells around i with r^2 < cutoff^2.
  void computeForces(ParticleData* d_particles, float cutoff, float G, float eps);

  int3 getGridDims() const;
  float getCellSize() const { return cell_size_; }
  int getTotalCells() const;
  size_t getMaxParticles() const noexcept { return max_particles_; }

  void copyCellDataToHost(std::vector<int>& cell_start, std::vector<int>& cell_end,
                          std::vector<int>& particle_cells, std::vector<int>& sorted_indices);
  bool verifyCellAssignment(const ParticleData* h_particles) const;

  static int3 getCellIndex(float x, float y, float z, float cell_size);
  static int hashCell(int3 cell, int3 grid_dims);

  nbx_grid* handle() noexcept { return handle_; }

 private:
  nbx_grid* handle_ = nullptr;
  size_t max_particles_;
  float cell_size_;
};

}  // namespace nbody
