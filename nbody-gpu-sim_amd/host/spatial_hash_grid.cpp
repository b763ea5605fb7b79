// SpatialHashGrid over the C ABI (reference: src/cuda/force_spatial_hash.cu:15-26,
// 160-368).
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "nbody/error_handling.hpp"
#include "nbody/spatial_hash_grid.hpp"
#include "nbx.h"

namespace nbody {

SpatialHashGrid::SpatialHashGrid(size_t max_particles, float cell_size)
    : max_particles_(max_particles), cell_size_(cell_size) {
  NBODY_CHECK(nbx_grid_create(max_particles, cell_size, &handle_));
}

SpatialHashGrid::~SpatialHashGrid() { nbx_grid_destroy(handle_); }

void SpatialHashGrid::build(const ParticleData* d_particles) {
  const int st =
      nbx_grid_build(handle_, reinterpret_cast<const nbx_particles*>(d_particles), nullptr);
  // The oversized-grid refusal is a plain runtime_error in the reference (:257-260);
  // a particle count above max_particles stays a ResourceException.
  if (st == NBX_ERR_CAPACITY && d_particles && d_particles->count <= max_particles_)
    throw std::runtime_error(nbx_last_error());
  NBODY_CHECK(st);
}

void SpatialHashGrid::computeForces(ParticleData* d_particles, float cutoff, float G, float eps) {
  NBODY_CHECK(nbx_grid_forces(handle_, reinterpret_cast<nbx_particles*>(d_particles), cutoff, G,
                              eps, nullptr));
}

int3 SpatialHashGrid::getGridDims() const {
  int32_t d[3] = {0, 0, 0};
  NBODY_CHECK(nbx_grid_info(handle_, d, nullptr, nullptr, nullptr));
  return make_int3(d[0], d[1], d[2]);
}

int SpatialHashGrid::getTotalCells() const {
  int32_t t = 0;
  NBODY_CHECK(nbx_grid_info(handle_, nullptr, nullptr, &t, nullptr));
  return t;
}

void SpatialHashGrid::copyCellDataToHost(std::vector<int>& cell_start, std::vector<int>& cell_end,
                                         std::vector<int>& particle_cells,
                                         std::vector<int>& sorted_indices) {
  const size_t total = (size_t)getTotalCells();
  cell_start.resize(total);
  cell_end.resize(total);
  particle_cells.resize(max_particles_);
  sorted_indices.resize(max_particles_);
  NBODY_CHECK(nbx_grid_copy_cells(handle_, cell_start.data(), cell_end.data(),
                                  particle_cells.data(), sorted_indices.data()));
}

// :342-368: every particle lies inside the (clamped) cell its position maps to.
bool SpatialHashGrid::verifyCellAssignment(const ParticleData* h_particles) const {
  int32_t dims[3] = {0, 0, 0};
  float box[6];
  NBODY_CHECK(nbx_grid_info(handle_, dims, nullptr, nullptr, box));
  for (size_t i = 0; i < h_particles->count; i++) {
    const float p[3] = {h_particles->pos_x[i], h_particles->pos_y[i], h_particles->pos_z[i]};
    const int3 c = getCellIndex(p[0] - box[0], p[1] - box[1], p[2] - box[2], cell_size_);
    const int cc[3] = {c.x, c.y, c.z};
    for (int a = 0; a < 3; ++a) {
      const int k = std::max(0, std::min(cc[a], dims[a] - 1));
      const float lo = box[a] + k * cell_size_;
      if (p[a] < lo || p[a] > lo + cell_size_) return false;
    }
  }
  return true;
}

int3 SpatialHashGrid::getCellIndex(float x, float y, float z, float cell_size) {
  return make_int3(static_cast<int>(std::floor(x / cell_size)),
                   static_cast<int>(std::floor(y / cell_size)),
                   static_cast<int>(std::floor(z / cell_size)));
}

int SpatialHashGrid::hashCell(int3 cell, int3 grid_dims) {
  const int cx = ((cell.x % grid_dims.x) + grid_dims.x) % grid_dims.x;
  const int cy = ((cell.y % grid_dims.y) + grid_dims.y) % grid_dims.y;
  const int cz = ((cell.z % grid_dims.z) + grid_dims.z) % grid_dims.z;
  return cx + cy * grid_dims.x + cz * grid_dims.x * grid_dims.y;
}

}  // namespace nbody
