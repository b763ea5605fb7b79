// Barnes-Hut tree of the engine's C++ API, mirroring the reference's
// include/nbody/barnes_hut_tree.hpp (build-workbench/nbody-gpu-sim). The build
// runs entirely on the GPU (nbx_bh_build). getNodes() returns the nodes of the
// last build in the reference's 76-byte layout and numbering: the first call
// after a build copies them to the host (synchronously), as copyNodesToHost() does.
#pragma once

#include <cstdint>
#include <vector>

#include "nbody/types.hpp"

struct nbx_bh;

namespace nbody {

struct OctreeNode {
  Vec3 center;
  float half_size;
  Vec3 center_of_mass;
  float total_mass;
  int children[8];
  int particle_index;
  bool is_leaf;
  int particle_count;

  OctreeNode()
      : center(), half_size(0), center_of_mass(), total_mass(0), particle_index(-1),
        is_leaf(true), particle_count(0) {
    for (int i = 0; i < 8; i++) children[i] = -1;
  }
};
static_assert(sizeof(OctreeNode) == 76, "OctreeNode must keep the 76-byte layout");

class BarnesHutTree {
 public:
  explicit BarnesHutTree(size_t max_particles);
  ~BarnesHutTree();
  BarnesHutTree(const BarnesHutTree&) = delete;
  BarnesHutTree& operator=(const BarnesHutTree&) = delete;
  BarnesHutTree(BarnesHutTree&&) = delete;
  BarnesHutTree& operator=(BarnesHutTree&&) = delete;

  void build(const ParticleData* d_particles);
  void computeForces(ParticleData* d_particles, float theta, float G, float eps);

  int getNodeCount() const;
  int getMaxDepth() const;
  const OctreeNode* getNodes() const;
  void copyNodesToHost();
  bool verifyMassConservation(const ParticleData* h_particles) const;

  nbx_bh* handle() noexcept { return handle_; }

 private:
  nbx_bh* handle_ = nullptr;
  size_t max_particles_;
  mutable std::vector<OctreeNode> h_nodes_;
  mutable bool nodes_stale_ = false;  // a build ran since the last host copy
  void refreshNodes() const;
};

// Morton code utilities (force_barnes_hut.cu:34-50 of the reference).
unsigned int expandBits(unsigned int v);
unsigned int computeMortonCode(float x, float y, float z);

}  // namespace nbody
