/*
 * nbx — C ABI of the MI355X (gfx950) N-body force engine.
 *
 * Plain pointers, sizes and status codes; no C++ or torch types. Every entry
 * point returns NBX_OK (0) or a non-zero status; nbx_last_error() returns the
 * calling thread's message for the last failure. Device pointers are HIP
 * device memory; a stream argument is a hipStream_t passed as void* (NULL is
 * the null stream). Calls are asynchronous on their stream unless the comment
 * says "synchronous".
 *
 * Each declaration names the reference interface it replaces
 * (build-workbench/nbody-gpu-sim, paths relative to its root).
 */
#ifndef NBX_H
#define NBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  NBX_OK = 0,
  NBX_ERR_INVALID = 1,  /* bad argument (maps to ValidationException) */
  NBX_ERR_HIP = 2,      /* HIP runtime failure (maps to the GPU exception) */
  NBX_ERR_NOMEM = 3,    /* device or host allocation failed (ResourceException) */
  NBX_ERR_CAPACITY = 4, /* input larger than the object was created for */
  NBX_ERR_STATE = 5     /* call order violated (e.g. forces before build) */
};

typedef void* nbx_stream; /* hipStream_t */

/* Layout-identical to nbody::ParticleData (include/nbody/types.hpp:96-131):
 * 13 FP32 SoA arrays + count. Pointers are device pointers unless a function
 * says host. */
typedef struct nbx_particles {
  float* pos_x;
  float* pos_y;
  float* pos_z;
  float* vel_x;
  float* vel_y;
  float* vel_z;
  float* acc_x;
  float* acc_y;
  float* acc_z;
  float* acc_old_x;
  float* acc_old_y;
  float* acc_old_z;
  float* mass;
  size_t count;
} nbx_particles;

/* ------------------------------------------------------------ runtime */
const char* nbx_status_string(int status);
const char* nbx_last_error(void);
/* Device 0 properties (replaces the cudaGetDeviceProperties probe of
 * src/utils/error_handling.cpp:26-45). Synchronous. */
int nbx_device_info(char* name, size_t name_len, int* cu_count, int* max_clock_khz,
                    size_t* total_mem);
int nbx_malloc(void** ptr, size_t bytes);
int nbx_free(void* ptr);
int nbx_memcpy_htod(void* dst, const void* src, size_t bytes);
int nbx_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int nbx_memcpy_dtod(void* dst, const void* src, size_t bytes);
int nbx_synchronize(void);
int nbx_stream_create(nbx_stream* out);
int nbx_stream_destroy(nbx_stream s);
int nbx_stream_synchronize(nbx_stream s);
/* HIP events for timing kernels on the stream they run on. */
int nbx_event_create(void** ev);
int nbx_event_destroy(void* ev);
int nbx_event_record(void* ev, nbx_stream s);
int nbx_event_elapsed_ms(void* start, void* stop, float* ms); /* synchronous */

/* ParticleDataManager (include/nbody/particle_data.hpp:9-28, src/cuda/particle_init.cu:11-160).
 * alloc: 13 device arrays of count floats, acc and acc_old zeroed (particle_init.cu:11-45).
 * free: releases non-null arrays and zeroes the struct (:47-75).
 * upload/download: pos, vel, acc, mass — not acc_old (:134-160). Synchronous. */
int nbx_particles_alloc(nbx_particles* p, size_t count);
int nbx_particles_free(nbx_particles* p);
int nbx_particles_upload(nbx_particles* dev, const nbx_particles* host);
int nbx_particles_download(nbx_particles* host, const nbx_particles* dev);

/* ------------------------------------------------------------ direct N^2 */
/* launchDirectForceKernel(ParticleData*, float G, float eps2, int block_size)
 * (src/cuda/force_direct.cu:88-98): acc_i = G sum_{j!=i} m_j r_ij (|r_ij|^2+eps2)^{-3/2}.
 * block_size is validated (1..1024) as the reference launch would. Like the
 * reference launcher it takes no workspace and allocates nothing: it runs the
 * one-sided tiled kernel (the tile shape from the autotuner; a forced
 * pair-symmetric shape falls back to the one-sided default). */
int nbx_direct_forces(nbx_particles* d, float G, float eps2, int block_size, nbx_stream s);
/* Forces for a contiguous block of targets [i0, i0+ni) against all j (sample
 * checks and the ensemble's standalone comparison). acc written at i0..i0+ni-1. */
int nbx_direct_forces_range(nbx_particles* d, size_t i0, size_t ni, float G, float eps2,
                            nbx_stream s);
/* One whole Velocity Verlet step with direct forces (Integrator::integrate,
 * src/cuda/integrator.cu:175-189): store+drift fused, then the force kernel
 * with the kick fused into its epilogue. Same arithmetic as the four separate
 * reference kernels. One-sided kernel, no workspace. */
int nbx_direct_step(nbx_particles* d, float G, float eps2, float dt, nbx_stream s);
/* Second half of nbx_direct_step: the force kernel with the kick
 * v += (a_old + a_new)*(0.5f*dt) in its epilogue (after nbx_verlet_drift). */
int nbx_direct_forces_kick(nbx_particles* d, float G, float eps2, float dt, nbx_stream s);

/* DirectForceCalculator (include/nbody/force_calculator.hpp:29-39,
 * src/cuda/force_direct.cu:101-106) with the workspace of the pair-symmetric
 * kernel, which evaluates each unordered pair once and applies it to both bodies.
 * The workspace is allocated once by nbx_direct_create for up to max_particles
 * bodies and released by nbx_direct_destroy; the compute calls allocate nothing.
 * scratch_bytes bounds the FP32 partial sums (0: the one-round need, at most
 * 1 GiB); a smaller workspace runs the tiles in rounds of tile distances and
 * gives the same accelerations bit for bit. Below the one-sided cut-over (or
 * when a one-sided shape is forced or tuned) the one-sided kernel runs.
 * Calls that share a workspace must be ordered (one stream, or synchronised):
 * overlapping calls race on its scratch. NBX_ERR_CAPACITY for more bodies than
 * the workspace was created for; NBX_ERR_STATE on another device. */
typedef struct nbx_direct nbx_direct;
int nbx_direct_create(size_t max_particles, size_t scratch_bytes, nbx_direct** out);
int nbx_direct_destroy(nbx_direct* op);
/* DirectForceCalculator::computeForces(ParticleData*) (force_direct.cu:101-106). */
int nbx_direct_compute(nbx_direct* op, nbx_particles* d, float G, float eps2, nbx_stream s);
/* The forces with the kick v += (a_old + a_new)*(0.5f*dt) fused (after nbx_verlet_drift). */
int nbx_direct_compute_kick(nbx_direct* op, nbx_particles* d, float G, float eps2, float dt,
                            nbx_stream s);
/* Integrator::integrate (src/cuda/integrator.cu:175-189) with these forces:
 * store+drift, forces, kick. */
int nbx_direct_compute_step(nbx_direct* op, nbx_particles* d, float G, float eps2, float dt,
                            nbx_stream s);
/* What a compute call on n bodies runs now: the tile variant, the number of
 * rounds and the bytes of partial sums each round uses (0 for a one-sided
 * variant). No reference counterpart (diagnostics). */
int nbx_direct_plan(const nbx_direct* op, size_t n, int* variant, int* rounds,
                    size_t* partial_bytes);

/* nbx_direct_compute without a workspace and without any allocation: no
 * reference counterpart. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The pair-symmetric partial sums and FP64 carry live in a fixed 64 MiB device
 * array inside the library, one per GPU, held until the process ends; the
 * tiles run in as many rounds as it needs, with accelerations bit-identical to
 * nbx_direct_compute. NBX_ERR_CAPACITY above about 1.1 million bodies for the
 * pair-symmetric shapes (one-sided shapes have no limit). Every call on a GPU
 * shares its array: calls that overlap on one GPU race. */
int nbx_direct_compute_noalloc(nbx_particles* d, float G, float eps2, nbx_stream s);

/* Tile variants of the direct kernel, named iplI_nbB_wsW: I targets per lane,
 * stage groups of B bodies, W waves splitting the j range of one tile. */
int nbx_direct_variant_count(void);
int nbx_direct_variant_name(int v, char* buf, size_t len);
/* -1 restores the autotuned choice. */
int nbx_direct_set_variant(int v);
/* The variant a call on n bodies launches now (forced, tuned or default). */
int nbx_direct_variant_for(size_t n);
/* Time every variant on a synthetic input of n bodies; remember the fastest for
 * n's size class (used by later calls). Synchronous. best_ms may be NULL. */
int nbx_direct_autotune(size_t n, int reps, int* best_variant, double* best_ms);

/* ------------------------------------------------------------ integrator */
/* storeAccelerationsKernel (integrator.cu:40-49). */
int nbx_store_old_accelerations(nbx_particles* d, nbx_stream s);
/* updatePositionsKernel (integrator.cu:12-22): x += v*dt + a*(0.5f*dt*dt). */
int nbx_update_positions(nbx_particles* d, float dt, nbx_stream s);
/* updateVelocitiesKernel (integrator.cu:25-37): v += (a_old + a_new)*(0.5f*dt). */
int nbx_update_velocities(nbx_particles* d, float dt, nbx_stream s);
/* storeAccelerations + updatePositions fused in one launch. */
int nbx_verlet_drift(nbx_particles* d, float dt, nbx_stream s);

/* Energies and momentum with FP64 reductions (replaces Integrator::computeKineticEnergy /
 * computePotentialEnergy, integrator.cu:203-240, whose sums are FP32). PE uses the
 * softened pair potential of integrator.cu:82-104 with eps2 = eps*eps. Synchronous.
 * Any output pointer may be NULL to skip that quantity (PE is O(N^2)). */
int nbx_energy(const nbx_particles* d, float G, float eps2, double* ke, double* pe,
               double momentum[3], nbx_stream s);

/* ------------------------------------------------------------ Barnes-Hut */
/* BarnesHutTree (include/nbody/barnes_hut_tree.hpp:33-86, src/cuda/force_barnes_hut.cu).
 * Node layout exported by nbx_bh_copy_nodes is the 76-byte OctreeNode
 * (barnes_hut_tree.hpp:9-30) in the reference's node numbering. */
typedef struct nbx_bh nbx_bh;
#define NBX_OCTREE_NODE_BYTES 76

/* BarnesHutTree(size_t max_particles) (:252-258): node budget 2*max_particles. */
int nbx_bh_create(size_t max_particles, nbx_bh** out);
int nbx_bh_destroy(nbx_bh* t);
/* BarnesHutTree::build (:330-336): bbox, 30-bit Morton codes, stable radix
 * sort, octree build and centre of mass, all on the GPU. */
int nbx_bh_build(nbx_bh* t, const nbx_particles* d, nbx_stream s);
/* BarnesHutTree::computeForces(ParticleData*, theta, G, eps) (:545-555). */
int nbx_bh_forces(nbx_bh* t, nbx_particles* d, float theta, float G, float eps, nbx_stream s);
/* getNodeCount / getMaxDepth (barnes_hut_tree.hpp:51-52). Synchronous. */
int nbx_bh_node_count(nbx_bh* t, int32_t* node_count, int32_t* max_depth);
/* copyNodesToHost (:557-560): node_count 76-byte nodes into host memory. Synchronous. */
int nbx_bh_copy_nodes(nbx_bh* t, void* host_nodes, size_t capacity_nodes);
/* Parity export (no reference counterpart): bbox {minx,miny,minz,maxx,maxy,maxz},
 * sorted Morton codes and the sorted particle indices of the last build.
 * Synchronous; any pointer may be NULL. */
int nbx_bh_copy_sort(nbx_bh* t, float bbox[6], uint32_t* codes, int32_t* indices, size_t n);
/* Build diagnostics: which builder ran (0 parallel, 1 sequential exact), the
 * number of multi-particle leaves at depth 32, and per-phase times of the last
 * build/forces in ms when timing is enabled (NULL to skip). Synchronous. */
int nbx_bh_build_info(nbx_bh* t, int32_t* builder, int32_t* depth32_leaves);
int nbx_bh_set_timing(nbx_bh* t, int enable);
/* Builder selection for cross-checks: 0 = automatic (parallel unless the node
 * budget could fire), 1 = always the exact sequential builder. */
int nbx_bh_set_builder(nbx_bh* t, int mode);
/* Task order of the traversal for cross-checks (no reference counterpart;
 * applies from the next build): 0 = 64-particle tasks in Hilbert order (the
 * default), 1 = in the tree's own order (path order; Morton order after the
 * sequential builder). Forces are the same bit for bit either way: only the
 * number of records the waves visit changes (nbx_bh_walk_census). */
int nbx_bh_set_task_order(nbx_bh* t, int mode);
/* Diagnostics (no reference counterpart): the walk nbx_bh_forces makes for
 * (theta, G, eps) on the last build, counted instead of summed. counts[0] the
 * records the waves visit (one per step of a 64-particle wave's walk),
 * [1] particle-node force evaluations, [2] particle-node openings, [3] active
 * lane-steps. The same tasks and decisions as the force walk; synchronous. */
int nbx_bh_walk_census(nbx_bh* t, const nbx_particles* d, float theta, float G, float eps,
                       int64_t counts[4]);
/* phases: bbox, morton, sort, paths, tree, com, export, forces (8 values, ms). */
int nbx_bh_phase_ms(nbx_bh* t, double ms[8]);

/* ------------------------------------------------------------ spatial hash */
/* SpatialHashGrid (include/nbody/spatial_hash_grid.hpp:9-62,
 * src/cuda/force_spatial_hash.cu:160-368): uniform cell grid for short-range
 * forces. Cell numbering cx + cy*dims[0] + cz*dims[0]*dims[1]. */
typedef struct nbx_grid nbx_grid;

/* SpatialHashGrid(size_t max_particles, float cell_size) (:161-174). */
int nbx_grid_create(size_t max_particles, float cell_size, nbx_grid** out);
int nbx_grid_destroy(nbx_grid* g);
/* SpatialHashGrid::build (:241-309): bbox padded by 0.001f, dims
 * (int)ceilf((max-min)/cell)+1, cell of every particle, particles grouped by
 * cell, cell_start/cell_end of every cell. Within a cell the order (which the
 * reference's atomic scatter leaves open) is the Morton index of the particle's
 * sub-cell (16 per axis up to 2^20 cells, 4 up to 2^26, none beyond), then the
 * particle index.
 * Reads the bbox back (24 bytes) to size the grid, so it waits for the stream.
 * NBX_ERR_CAPACITY "Spatial hash grid too large: ..." above 1e8 cells (:257-260)
 * or when count exceeds max_particles. */
int nbx_grid_build(nbx_grid* g, const nbx_particles* d, nbx_stream s);
/* SpatialHashGrid::computeForces(ParticleData*, cutoff, G, eps) (:311-322): acc of
 * every particle from the j != i in its 27 neighbour cells with r^2 < cutoff^2.
 * Needs a build on the same particle count (NBX_ERR_STATE otherwise). */
int nbx_grid_forces(nbx_grid* g, nbx_particles* d, float cutoff, float G, float eps,
                    nbx_stream s);
/* getGridDims / getCellSize / getTotalCells (spatial_hash_grid.hpp:27-29), plus the
 * padded bbox {min x,y,z, max x,y,z} of the last build. Any pointer may be NULL. */
int nbx_grid_info(nbx_grid* g, int32_t dims[3], float* cell_size, int32_t* total_cells,
                  float bbox[6]);
/* copyCellDataToHost (:324-340): cell_start/cell_end hold total_cells entries,
 * particle_cells/sorted_indices max_particles entries (entries past the built
 * count are -1). Synchronous; any pointer may be NULL. */
int nbx_grid_copy_cells(nbx_grid* g, int32_t* cell_start, int32_t* cell_end,
                        int32_t* particle_cells, int32_t* sorted_indices);
/* Force walk for cross-checks (no reference counterpart): 0 = automatic (shared
 * scalar walk where a wave's targets lie in one row of cells), 1 = per-lane walks
 * only. Results are identical either way. */
int nbx_grid_set_walk(nbx_grid* g, int mode);

/* ------------------------------------------------------------ precision form */
/* Double-single positions (hi + lo FP32 pairs), Kahan-compensated FP32 pair-force
 * sums, FP64 energies. No reference counterpart (extension of the direct path);
 * hi arrays are the nbx_particles pos arrays, lo arrays are extra. */
typedef struct nbx_ds_particles {
  nbx_particles p; /* pos_* hold the hi parts */
  float* pos_lo_x;
  float* pos_lo_y;
  float* pos_lo_z;
} nbx_ds_particles;
int nbx_ds_alloc(nbx_ds_particles* p, size_t count); /* lo parts zeroed */
int nbx_ds_free(nbx_ds_particles* p);
/* Forces; op is a direct workspace (nbx_direct_create) for the pair-symmetric
 * kernel, or NULL for the one-sided kernel (no scratch). */
int nbx_ds_forces(nbx_direct* op, nbx_ds_particles* d, float G, float eps2, nbx_stream s);
/* Kernel of nbx_ds_forces / nbx_ds_step: -1 pair-symmetric above 2^13 bodies when
 * a workspace is given (the default), 0 one-sided, 1 pair-symmetric (needs one). */
int nbx_ds_set_kernel(int kind);
/* store+drift in double-single: (hi,lo) += v*dt + a*(0.5*dt*dt) with renormalisation. */
int nbx_ds_drift(nbx_ds_particles* d, float dt, nbx_stream s);
/* drift, forces, kick. */
int nbx_ds_step(nbx_direct* op, nbx_ds_particles* d, float G, float eps2, float dt,
                nbx_stream s);
int nbx_ds_energy(const nbx_ds_particles* d, float G, float eps2, double* ke, double* pe,
                  double momentum[3], nbx_stream s);

/* ------------------------------------------------------------ batched ensemble */
/* Many independent systems of uneven N, concatenated in one SoA set. Forces of
 * every system in one persistent launch pulling (system, i-tile) items from a
 * device work queue (largest first); Verlet steps captured in hipGraphs.
 * No reference counterpart. */
typedef struct nbx_ensemble nbx_ensemble;
int nbx_ens_create(int nsys, const int64_t* counts, nbx_ensemble** out);
int nbx_ens_destroy(nbx_ensemble* e);
/* Concatenated particle arrays (device) and per-system offsets (host, nsys+1). */
int nbx_ens_particles(nbx_ensemble* e, nbx_particles* view, int64_t* offsets);
int nbx_ens_forces(nbx_ensemble* e, float G, float eps2, nbx_stream s);
/* `steps` Verlet steps; when graph_chunk > 0 the steps are replayed from a
 * hipGraph holding graph_chunk steps (steps must be a multiple). */
int nbx_ens_steps(nbx_ensemble* e, float G, float eps2, float dt, int steps, int graph_chunk,
                  nbx_stream s);
/* Baseline for the batched form: the same `steps`, advancing each system on
 * its own with the single-system entry point (nbx_direct_compute_step per
 * system, with a workspace the ensemble holds for its largest system). */
int nbx_ens_steps_sequential(nbx_ensemble* e, float G, float eps2, float dt, int steps,
                             nbx_stream s);
/* Tile shapes: the same (targets per lane, waves per tile) list and order as the
 * single-system variants (nbx_direct_variant_name). */
int nbx_ens_variant_count(void);
/* Choose the ensemble tile shape by timing (synchronous). */
int nbx_ens_autotune(nbx_ensemble* e, int reps, int* best_variant, double* best_ms);
int nbx_ens_set_variant(nbx_ensemble* e, int v);
/* Per-system FP64 energies (host arrays of nsys). Synchronous. */
int nbx_ens_energy(nbx_ensemble* e, float G, float eps2, double* ke, double* pe, nbx_stream s);

/* ------------------------------------------------------------ input generators */
/* Deterministic synthetic inputs written to HOST arrays (float), generated in
 * double from SplitMix64 streams and rounded once. Portable across compilers,
 * unlike std::uniform_real_distribution. Masses are M/n. */
/* Plummer sphere (G = 1): inverse-CDF radii truncated at enclosed-mass fraction
 * mass_cut (<= 0 means 0.999), von Neumann speeds on q^2 (1-q^2)^{7/2}
 * (Aarseth, Henon & Wielen 1974), isotropic directions, centre of mass and
 * mean velocity removed. */
int nbx_gen_plummer(size_t n, uint64_t seed, double a, double M, double mass_cut, float* x,
                    float* y, float* z, float* vx, float* vy, float* vz, float* m);
/* Uniform-density sphere of radius R, masses M/n, Gaussian velocities (sigma per axis). */
int nbx_gen_uniform_sphere(size_t n, uint64_t seed, double R, double M, double sigma, float* x,
                           float* y, float* z, float* vx, float* vy, float* vz, float* m);
/* Two equal Plummer spheres of n/2 each at x = -+sep/2 approaching with relative
 * speed vrel along x; each sphere has mass M_each. n must be even. */
int nbx_gen_two_plummer(size_t n, uint64_t seed, double a, double M_each, double sep,
                        double vrel, float* x, float* y, float* z, float* vx, float* vy,
                        float* vz, float* m);
/* nsys sizes floor(exp(U(ln nmin, ln nmax))), clamped to [nmin, nmax]. */
int nbx_gen_ensemble_sizes(int nsys, uint64_t seed, int64_t nmin, int64_t nmax, int64_t* counts);

#ifdef __cplusplus
}
#endif
#endif /* NBX_H */
