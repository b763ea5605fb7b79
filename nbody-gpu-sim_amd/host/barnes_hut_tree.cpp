// BarnesHutTree over the C ABI (reference: src/cuda/force_barnes_hut.cu:252-570).
#include <cmath>

#include "nbody/barnes_hut_tree.hpp"
#include "nbody/error_handling.hpp"
#include "nbody/performance_observability.hpp"
#include "nbx.h"

namespace nbody {

BarnesHutTree::BarnesHutTree(size_t max_particles) : max_particles_(max_particles) {
  NBODY_CHECK(nbx_bh_create(max_particles, &handle_));
}

BarnesHutTree::~BarnesHutTree() { nbx_bh_destroy(handle_); }

void BarnesHutTree::build(const ParticleData* d_particles) {
  NBODY_PROFILE_SCOPE("barnes_hut.build");
  NBODY_CHECK(nbx_bh_build(handle_, reinterpret_cast<const nbx_particles*>(d_particles), nullptr));
  nodes_stale_ = true;
}

void BarnesHutTree::computeForces(ParticleData* d_particles, float theta, float G, float eps) {
  NBODY_CHECK(nbx_bh_forces(handle_, reinterpret_cast<nbx_particles*>(d_particles), theta, G, eps,
                            nullptr));
}

int BarnesHutTree::getNodeCount() const {
  int32_t c = 0, d = 0;
  NBODY_CHECK(nbx_bh_node_count(handle_, &c, &d));
  return c;
}

int BarnesHutTree::getMaxDepth() const {
  int32_t c = 0, d = 0;
  NBODY_CHECK(nbx_bh_node_count(handle_, &c, &d));
  return d;
}

void BarnesHutTree::refreshNodes() const {
  const int c = getNodeCount();
  h_nodes_.assign((size_t)c, OctreeNode{});
  if (c) NBODY_CHECK(nbx_bh_copy_nodes(handle_, h_nodes_.data(), h_nodes_.size()));
  nodes_stale_ = false;
}

void BarnesHutTree::copyNodesToHost() { refreshNodes(); }

const OctreeNode* BarnesHutTree::getNodes() const {
  if (nodes_stale_) refreshNodes();
  return h_nodes_.data();
}

bool BarnesHutTree::verifyMassConservation(const ParticleData* h_particles) const {
  float total = 0.0f;
  for (size_t i = 0; i < h_particles->count; i++) total += h_particles->mass[i];
  if (nodes_stale_) refreshNodes();
  if (h_nodes_.empty()) return false;
  return std::abs(total - h_nodes_[0].total_mass) < 0.001f * total;
}

unsigned int expandBits(unsigned int v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

unsigned int computeMortonCode(float x, float y, float z) {
  const unsigned ix = (unsigned)(std::fmin(std::fmax(x, 0.0f), 1.0f) * 1023.0f);
  const unsigned iy = (unsigned)(std::fmin(std::fmax(y, 0.0f), 1.0f) * 1023.0f);
  const unsigned iz = (unsigned)(std::fmin(std::fmax(z, 0.0f), 1.0f) * 1023.0f);
  return (expandBits(ix) << 2) | (expandBits(iy) << 1) | expandBits(iz);
}

}  // namespace nbody
