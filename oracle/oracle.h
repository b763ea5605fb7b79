/*
 * CPU oracle for the N-body force engine — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's arithmetic for the hot path
 * (direct N^2 gravity, Velocity Verlet, Barnes-Hut bbox / Morton / stable
 * sort / octree build / centre of mass / traversal) plus an FP64 direct sum.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline. The
 * product path (libnbx) never links or calls it.
 *
 * Built with gcc -O2 -ffp-contract=off and no fast-math, so every float
 * expression rounds exactly as written (the reference host code is built for
 * baseline x86-64, which has no FMA instructions).
 *
 * Pinning (what checks each function besides this restatement):
 *   orc_pair_accel          closed forms of tests/test_force_calculation.cpp:17-64,181-253
 *   orc_direct_f32          N=1 zero, N=2 antisymmetry, N=32 grid KAT, FP64 numpy sum
 *   orc_direct_f64          independent numpy float64 sum (tests/golden)
 *   orc_drift / orc_kick    closed forms of tests/test_integrator.cpp:14-74
 *   orc_energy_f64          closed form KE=4.5 (test_integrator.cpp:76-102), numpy float64
 *   orc_expand_bits         exhaustive naive bit interleave over 0..1023
 *   orc_morton_codes        naive Python interleave on the same float normalisation
 *   orc_bbox                numpy min/max (exact)
 *   orc_stable_sort         numpy argsort(kind="stable")
 *   orc_build_tree          independent pure-Python insertion restatement + PR-octree
 *                           invariants + mass conservation (test_barnes_hut.cpp)
 *   orc_bh_forces           theta->0 limit vs FP64 direct sum, 2% KAT of
 *                           test_spatial_hash.cpp:171-227, convergence KAT
 *   orc_grid_cell_index     KAT of test_spatial_hash.cpp:34-47
 *   orc_grid_hash_cell      Python modulo on every offset of a small box
 *   orc_grid_dims           numpy float32 restatement of the sizing formulas
 *   orc_grid_build          numpy bincount/cumsum and a numpy stable argsort of the
 *                           (cell, sub-cell) key;
 *                           one-cell-per-particle property (test_spatial_hash.cpp:133-167);
 *                           every particle inside its cell's bounds (verifyCellAssignment)
 *   orc_grid_forces_f32/f64 brute-force cutoff sum without a grid (same pair set,
 *                           FP64), cutoff-boundary KAT of test_spatial_hash.cpp:81-127,
 *                           cutoff >= all distances reduces to the FP64 direct sum
 */
#ifndef NBODY_ORACLE_H
#define NBODY_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Octree node, byte layout of nbody::OctreeNode (include/nbody/barnes_hut_tree.hpp:9-30). */
typedef struct orc_node {
  float cx, cy, cz;
  float half_size;
  float comx, comy, comz;
  float total_mass;
  int32_t children[8];
  int32_t particle_index;
  uint8_t is_leaf;
  uint8_t pad_[3];
  int32_t particle_count;
} orc_node;

/* force_direct.cu:109-117 computeGravitationalAccelerationCPU. out = {ax, ay, az}. */
void orc_pair_accel(const float p1[3], const float p2[3], float m2, float G, float eps, float out[3]);

/* Sum of the CPU pair function over j != i in ascending j, as
 * tests/test_force_calculation.cpp:131-140 forms its reference.
 * targets == NULL means all i (then ax/ay/az have n entries), else ax[k]
 * belongs to targets[k]. nthreads <= 0 uses the OpenMP default. */
void orc_direct_f32(size_t n, const float* x, const float* y, const float* z, const float* m,
                    float G, float eps, const int64_t* targets, size_t ntargets,
                    float* ax, float* ay, float* az, int nthreads);

/* FP64 direct sum from FP32 inputs: a_i = G sum_{j!=i} m_j r_ij / (|r_ij|^2 + eps^2)^{3/2}.
 * eps2 is given as the float the kernels use (eps*eps rounded to float). */
void orc_direct_f64(size_t n, const float* x, const float* y, const float* z, const float* m,
                    double G, double eps2, const int64_t* targets, size_t ntargets,
                    double* ax, double* ay, double* az, int nthreads);

/* Same, with positions given as double (hi+lo pairs already summed by the caller). */
void orc_direct_f64_pos64(size_t n, const double* x, const double* y, const double* z,
                          const float* m, double G, double eps2, const int64_t* targets,
                          size_t ntargets, double* ax, double* ay, double* az, int nthreads);

/* integrator.cu:40-49 (acc_old <- acc) and :12-22 (x += v*dt + a*(0.5f*dt*dt)). */
void orc_store_old(size_t n, const float* ax, const float* ay, const float* az,
                   float* aox, float* aoy, float* aoz);
void orc_drift(size_t n, float* x, float* y, float* z, const float* vx, const float* vy,
               const float* vz, const float* ax, const float* ay, const float* az, float dt);
/* integrator.cu:25-37: v += (a_old + a_new)*(0.5f*dt). */
void orc_kick(size_t n, float* vx, float* vy, float* vz, const float* aox, const float* aoy,
              const float* aoz, const float* ax, const float* ay, const float* az, float dt);

/* Full reference-semantics run (particle_system.cpp:45-61,122-136 with
 * integrator.cu:175-189): a(0) by orc_direct_f32, then `steps` Verlet steps.
 * Arrays are updated in place; ao* receive the last acc_old. */
void orc_run_direct(size_t n, float* x, float* y, float* z, float* vx, float* vy, float* vz,
                    float* ax, float* ay, float* az, float* aox, float* aoy, float* aoz,
                    const float* m, float G, float eps, float dt, int steps, int nthreads);

/* FP64 energies and momentum: KE = sum 1/2 m v^2; PE = -G sum_{i<j} m_i m_j / sqrt(r^2+eps^2)
 * (the softened definition of integrator.cu:82-104, eps2 the float eps*eps the kernels use);
 * P = sum m v. */
void orc_energy_f64(size_t n, const float* x, const float* y, const float* z, const float* vx,
                    const float* vy, const float* vz, const float* m, double G, double eps2,
                    double* ke, double* pe, double mom[3], int nthreads);

/* ---- Barnes-Hut pieces (force_barnes_hut.cu) ---- */
uint32_t orc_expand_bits(uint32_t v);                       /* :34-40 */
uint32_t orc_morton_code(float x, float y, float z);        /* :42-50 */
/* :78-122 + :266-306: exact per-axis min/max starting from +-1e30. */
void orc_bbox(size_t n, const float* x, const float* y, const float* z, float bmin[3],
              float bmax[3]);
/* :308-322 + :125-139: size = (max-min)+0.001f, inv = 1.0f/size, code of (p-min)*inv. */
void orc_morton_codes(size_t n, const float* x, const float* y, const float* z,
                      const float bmin[3], const float bmax[3], uint32_t* codes, int32_t* idx);
/* :324-328 thrust::sort_by_key semantics: ascending, ties keep ascending index. */
void orc_stable_sort(size_t n, uint32_t* codes, int32_t* idx);
/* :391-543 sequential insertion build + :512-538 COM. nodes must hold max_nodes
 * entries. Returns 0, or -1 if the input would overrun max_nodes. */
int orc_build_tree(size_t n, const float* x, const float* y, const float* z, const float* m,
                   const int32_t* sorted_idx, const float bmin[3], const float bmax[3],
                   size_t max_nodes, orc_node* nodes, int32_t* node_count, int32_t* max_depth,
                   int32_t* budget_merges, int32_t* depth_merges);
/* :142-249 traversal with 1.0f/sqrtf for rsqrtf; theta and eps as the host passes them
 * (:552-553 squares both). targets as in orc_direct_f32. */
void orc_bh_forces(size_t n, const orc_node* nodes, int32_t node_count, const float* x,
                   const float* y, const float* z, const float* m, float theta, float G, float eps,
                   const int64_t* targets, size_t ntargets, float* ax, float* ay, float* az,
                   int nthreads);
/* One full BH force call (:577-583): bbox, Morton, stable sort, build, traversal.
 * nodes scratch of max_nodes entries; codes/idx scratch of n entries. */
int orc_bh_full(size_t n, const float* x, const float* y, const float* z, const float* m,
                float theta, float G, float eps, size_t max_nodes, orc_node* nodes,
                uint32_t* codes, int32_t* idx, float* ax, float* ay, float* az, int nthreads);

/* ---- Spatial hash pieces (force_spatial_hash.cu) ---- */
/* SpatialHashGrid::getCellIndex (:15-18): (int)floorf(p / cell) per axis. */
void orc_grid_cell_index(float x, float y, float z, float cell, int32_t out[3]);
/* SpatialHashGrid::hashCell (:20-26): each axis wrapped into [0, dims), linearised. */
int32_t orc_grid_hash_cell(const int32_t cell[3], const int32_t dims[3]);
/* Grid sizing of SpatialHashGrid::build (:190-260): exact bbox padded by 0.001f
 * (:232-238), dims = (int)ceilf((max-min)/cell) + 1 (:250-252), total = product
 * in 64 bits (:255). bbox = {min x,y,z, max x,y,z} after padding. Returns 0, or
 * -1 when total exceeds 1e8 cells (:257-260) or a ratio is not finite. */
int orc_grid_dims(size_t n, const float* x, const float* y, const float* z, float cell,
                  float bbox[6], int32_t dims[3], int64_t* total);
/* Cell contents of SpatialHashGrid::build (:278-308): particle_cell (n) by
 * :36-49, cell_start / cell_end (total) as the exclusive scan of the counts and
 * start + count (:291-300), sorted (n) the indices grouped by cell. Within a cell
 * (an order the reference's atomic scatter, :72-81, leaves open): ascending Morton
 * index of the 4 x 4 x 4 sub-cell, sub = floor(4u) - 4c per axis with
 * u = (p - min) / cell, then ascending particle index; grids above 2^26 cells
 * order by particle index alone. */
void orc_grid_build(size_t n, const float* x, const float* y, const float* z, float cell,
                    const float bbox[6], const int32_t dims[3], int32_t* particle_cell,
                    int32_t* cell_start, int32_t* cell_end, int32_t* sorted);
/* spatialHashForceKernel (:84-158) in IEEE FP32 with 1.0f/sqrtf for rsqrtf:
 * target cell from its position with the clamp (:103-108), neighbour cells in
 * dz, dy, dx order, bodies in `sorted` order, j == i skipped, pair kept when
 * r^2 < cutoff*cutoff with r^2 = fmaf(dz, dz, fmaf(dx, dx, dy*dy)) in FP32 (the
 * contracted form of the reference's dx*dx + dy*dy + dz*dz). targets as in
 * orc_direct_f32. */
void orc_grid_forces_f32(size_t n, const float* x, const float* y, const float* z,
                         const float* m, const float bbox[6], const int32_t dims[3], float cell,
                         const int32_t* cell_start, const int32_t* cell_end,
                         const int32_t* sorted, float cutoff, float G, float eps,
                         const int64_t* targets, size_t ntargets, float* ax, float* ay,
                         float* az, int nthreads);
/* The same pair set summed in FP64: G m_j r_ij (r^2 + eps2)^{-3/2} in double.
 * mag (may be NULL) receives sum_j |G m_j| |r_ij| (r^2+eps2)^{-3/2} per target,
 * the scale for rounding-error bounds; npairs (may be NULL) the kept pairs. */
void orc_grid_forces_f64(size_t n, const float* x, const float* y, const float* z,
                         const float* m, const float bbox[6], const int32_t dims[3], float cell,
                         const int32_t* cell_start, const int32_t* cell_end,
                         const int32_t* sorted, float cutoff, double G, double eps2,
                         const int64_t* targets, size_t ntargets, double* ax, double* ay,
                         double* az, double* mag, int64_t* npairs, int nthreads);
/* Brute force without a grid: every j != i whose FP32 r^2 (the same fused form)
 * is below cutoff*cutoff, summed in FP64. Equal to the grid sum whenever the
 * 27-cell search is complete (cell >= cutoff). */
void orc_cutoff_forces_f64(size_t n, const float* x, const float* y, const float* z,
                           const float* m, float cutoff, double G, double eps2,
                           const int64_t* targets, size_t ntargets, double* ax, double* ay,
                           double* az, int64_t* npairs, int nthreads);

int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
