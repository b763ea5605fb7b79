// Batched ensemble of independent systems of uneven N (an extension; the
// reference advances one system). All systems share one concatenated SoA
// particle set. Per step:
//   drift  - one flat launch over every particle (store + drift, integrator.cu:12-49)
//   forces - one persistent launch: workgroups pull (system, i-tile) items from a
//            device queue ordered largest system first, and run the same
//            direct_tile() as the single-system kernel, so each system's
//            accelerations are bit-identical to a standalone call with the
//            same tile shape; the kick v += (a_old + a_new) dt/2 is fused
//   The last workgroup to finish resets the queue, so steps need no host work
//   and whole chunks of steps are replayed from a captured hipGraph.
#include <algorithm>
#include <map>
#include <mutex>
#include <new>
#include <numeric>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "direct_core.hpp"
#include "direct_sym.hpp"
#include "direct_sym_core.hpp"
#include "verlet.hpp"

namespace nbx {
namespace {

template <int IPL, int NB, int WS, bool CHECKED, bool KICK>
__global__ __launch_bounds__(64 * WS) void ens_forces_kernel(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    const float* __restrict__ m, const int32_t* __restrict__ sys_off,
    const int32_t* __restrict__ sys_n, const int2* __restrict__ items, int nitems, int* queue,
    int* done, double G, float eps2, float* ax, float* ay, float* az, KickArgs kick) {
  __shared__ double red[WS * 3 * 64 * IPL];
  __shared__ int item_sh;
  for (;;) {
    if (threadIdx.x == 0) item_sh = atomicAdd(queue, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(item_sh);
    __syncthreads();
    if (it >= nitems) break;
    const int2 w = items[it];
    const int off = sys_off[w.x];
    const int n = sys_n[w.x];
    KickArgs k = kick;
    if (KICK) {
      k.vx += off;
      k.vy += off;
      k.vz += off;
      k.aox += off;
      k.aoy += off;
      k.aoz += off;
    }
    direct_tile<IPL, NB, WS, CHECKED, KICK>(x + off, y + off, z + off, m + off, n,
                                            w.y * 64 * IPL, n, G, eps2, ax + off, ay + off,
                                            az + off, k, red);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(done, 1) == (int)gridDim.x - 1) {  // every workgroup is past the queue
      atomicExch(queue, 0);
      atomicExch(done, 0);
      __threadfence();
    }
  }
}

// Pair-symmetric batched forces: workgroups pull (system, tile) items, largest
// systems first, and run the tiles of the single-system symmetric kernel on
// that system's slice (sym_run_tile), writing its partial sums. The queue
// resets as in ens_forces_kernel.
template <int IPL, int W, bool CHECKED>
__global__ __launch_bounds__(64 * W, IPL <= 6 ? 4 : 2) void ens_sym_kernel(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    const float* __restrict__ m, const int32_t* __restrict__ sys_off,
    const int32_t* __restrict__ sys_n, const int64_t* __restrict__ part_off, float* part,
    const int2* __restrict__ items, int nitems, int* queue, int* done, float eps2) {
  constexpr int S = 64 * IPL * W;
  __shared__ float red[2][W][3][64];
  __shared__ int item_sh;
  for (;;) {
    if (threadIdx.x == 0) item_sh = atomicAdd(queue, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(item_sh);
    __syncthreads();
    if (it >= nitems) break;
    const int2 w = items[it];
    const int off = sys_off[w.x];
    const SymGeom g = sym_geom(sys_n[w.x], S, part + part_off[w.x]);
    int A, d;
    sym_tile_of(g, w.y, A, d);  // items hold only owned tiles
    sym_run_tile<IPL, W, CHECKED>(x + off, y + off, z + off, m + off, g, A, d, eps2, red);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(done, 1) == (int)gridDim.x - 1) {  // every workgroup is past the queue
      atomicExch(queue, 0);
      atomicExch(done, 0);
      __threadfence();
    }
  }
}

// Accelerations (and kick) from the partial sums, one workgroup per
// (system, block of 256 bodies).
template <bool KICK>
__global__ __launch_bounds__(256) void ens_sym_reduce(
    const int32_t* __restrict__ blk_sys, const int32_t* __restrict__ blk_first,
    const int32_t* __restrict__ sys_off, const int32_t* __restrict__ sys_n,
    const int64_t* __restrict__ part_off, float* part, int S, double G, float* ax, float* ay,
    float* az, KickArgs kick) {
  const int s = blk_sys[blockIdx.x];
  const int b = blk_first[blockIdx.x] + threadIdx.x;
  const int n = sys_n[s];
  if (b >= n) return;
  const int off = sys_off[s];
  const SymGeom g = sym_geom(n, S, part + part_off[s]);
  KickArgs k = kick;
  if (KICK) {
    k.vx += off;
    k.vy += off;
    k.vz += off;
    k.aox += off;
    k.aoy += off;
    k.aoz += off;
  }
  sym_reduce_body<KICK>(g, b, G, ax + off, ay + off, az + off, k, nullptr, true);
}

// Per-system FP64 kinetic energy and softened potential, one workgroup per
// (system, block of 256 targets); partials summed per system in fixed order.
__global__ __launch_bounds__(256) void ens_energy_kernel(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz,
    const float* __restrict__ m, const int32_t* __restrict__ sys_off,
    const int32_t* __restrict__ sys_n, const int32_t* __restrict__ blk_sys,
    const int32_t* __restrict__ blk_first, double eps2, double* partial) {
  const int b = blockIdx.x;
  const int s = blk_sys[b];
  const int off = sys_off[s], n = sys_n[s];
  const int li = blk_first[b] + threadIdx.x;
  double ke = 0.0, pe = 0.0;
  if (li < n) {
    const int i = off + li;
    const double mi = m[i];
    const double ux = vx[i], uy = vy[i], uz = vz[i];
    ke = 0.5 * mi * (ux * ux + uy * uy + uz * uz);
    const double xi = x[i], yi = y[i], zi = z[i];
    double phi = 0.0;
    for (int j = 0; j < n; ++j) {
      const double dx = (double)x[off + j] - xi, dy = (double)y[off + j] - yi,
                   dz = (double)z[off + j] - zi;
      const double r2 = dx * dx + dy * dy + dz * dz + eps2;
      double r = (double)__builtin_amdgcn_rsqf((float)r2);
      r = r * (1.5 - 0.5 * r2 * r * r);
      r = r * (1.5 - 0.5 * r2 * r * r);
      phi += (j == li) ? 0.0 : (double)m[off + j] * r;
    }
    pe = mi * phi;
  }
  __shared__ double sh[2][4];
  for (int o = 32; o > 0; o >>= 1) {
    ke += __shfl_xor(ke, o, 64);
    pe += __shfl_xor(pe, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = ke;
    sh[1][threadIdx.x >> 6] = pe;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double t = 0.0;
    for (int q = 0; q < 4; ++q) t += sh[threadIdx.x][q];
    partial[2 * b + threadIdx.x] = t;
  }
}

struct EnsVariant {
  const char* name;
  int ipl, ws;
  int sym_block;  // 0: one-sided tiles; else the symmetric super-block size
  void* k[2][2];  // [checked][kick]; symmetric shapes apply the kick in ens_sym_reduce
};

#define NBX_ENS_V(I, B, W)                                                        \
  {"ipl" #I "_nb" #B "_ws" #W,                                                    \
   I,                                                                             \
   W,                                                                             \
   0,                                                                             \
   {{(void*)ens_forces_kernel<I, B, W, false, false>,                             \
     (void*)ens_forces_kernel<I, B, W, false, true>},                             \
    {(void*)ens_forces_kernel<I, B, W, true, false>,                              \
     (void*)ens_forces_kernel<I, B, W, true, true>}}},
#define NBX_ENS_SYM_V(I, W)                                                       \
  {"sym_ipl" #I "_ws" #W,                                                         \
   I,                                                                             \
   W,                                                                             \
   64 * I * W,                                                                    \
   {{(void*)ens_sym_kernel<I, W, false>, (void*)ens_sym_kernel<I, W, false>},     \
    {(void*)ens_sym_kernel<I, W, true>, (void*)ens_sym_kernel<I, W, true>}}},
// Same tile list and order as the single-system variants (direct.hip).
static const EnsVariant kEnsVariants[] = {NBX_DIRECT_VARIANTS(NBX_ENS_V)
                                              NBX_SYM_VARIANTS(NBX_ENS_SYM_V)};
#undef NBX_ENS_V
#undef NBX_ENS_SYM_V
constexpr int kEnsNumVariants = (int)(sizeof(kEnsVariants) / sizeof(kEnsVariants[0]));

}  // namespace
}  // namespace nbx

struct nbx_ensemble {
  int nsys = 0;
  size_t total = 0;
  std::vector<int64_t> offsets;  // host, nsys + 1
  nbx_particles p{};
  int32_t* d_off = nullptr;
  int32_t* d_n = nullptr;
  int* d_queue = nullptr;  // [0] queue head, [1] finished workgroups
  int variant = 0;
  int2* d_items[nbx::kEnsNumVariants] = {};
  int nitems[nbx::kEnsNumVariants] = {};
  int64_t* d_part_off[nbx::kEnsNumVariants] = {};  // symmetric shapes: per-system partial offset
  float* d_part = nullptr;                          // symmetric partial sums (largest shape)
  int grid[nbx::kEnsNumVariants][2][2] = {};
  hipStream_t own_stream = nullptr;
  // workspace of the one-system-at-a-time baseline (sized for the largest system)
  nbx_direct* seq = nullptr;
  // captured step chunk
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  std::tuple<int, float, float, float, int, hipStream_t> graph_key{};
  // energy
  int32_t* d_blk_sys = nullptr;
  int32_t* d_blk_first = nullptr;
  double* d_epart = nullptr;
  int eblocks = 0;
};

namespace nbx {
namespace {

int ens_forces_launch(nbx_ensemble* e, float G, float eps2, const KickArgs* kick, hipStream_t s,
                      int variant) {
  if (!(eps2 >= 0.0f) || !(eps2 < 3.0e38f))
    return fail(NBX_ERR_INVALID, "softening eps2 must be finite and >= 0");
  const int v = variant >= 0 ? variant : e->variant;
  const EnsVariant& V = kEnsVariants[v];
  const int chk = eps2 < kUncheckedMinEps2 ? 1 : 0;
  const int kk = kick ? 1 : 0;
  KickArgs k{};
  if (kick) k = *kick;
  const nbx_particles& p = e->p;
  const double Gd = (double)G;
  int nitems = e->nitems[v];
  int* done = e->d_queue + 1;
  if (V.sym_block) {
    void* args[] = {(void*)&p.pos_x,       (void*)&p.pos_y,       (void*)&p.pos_z,
                    (void*)&p.mass,        (void*)&e->d_off,      (void*)&e->d_n,
                    (void*)&e->d_part_off[v], (void*)&e->d_part,  (void*)&e->d_items[v],
                    (void*)&nitems,        (void*)&e->d_queue,    (void*)&done,
                    (void*)&eps2};
    NBX_HIP(hipLaunchKernel(V.k[chk][kk], dim3(e->grid[v][chk][kk]), dim3(64 * V.ws), args, 0, s));
    const int S = V.sym_block;
    if (kick)
      hipLaunchKernelGGL(ens_sym_reduce<true>, dim3(e->eblocks), dim3(256), 0, s, e->d_blk_sys,
                         e->d_blk_first, e->d_off, e->d_n, e->d_part_off[v], e->d_part, S, Gd,
                         p.acc_x, p.acc_y, p.acc_z, k);
    else
      hipLaunchKernelGGL(ens_sym_reduce<false>, dim3(e->eblocks), dim3(256), 0, s, e->d_blk_sys,
                         e->d_blk_first, e->d_off, e->d_n, e->d_part_off[v], e->d_part, S, Gd,
                         p.acc_x, p.acc_y, p.acc_z, k);
    NBX_LAUNCHED();
    return NBX_OK;
  }
  void* args[] = {(void*)&p.pos_x,  (void*)&p.pos_y, (void*)&p.pos_z,   (void*)&p.mass,
                  (void*)&e->d_off, (void*)&e->d_n,  (void*)&e->d_items[v], (void*)&nitems,
                  (void*)&e->d_queue, (void*)&done,  (void*)&Gd,        (void*)&eps2,
                  (void*)&p.acc_x,  (void*)&p.acc_y, (void*)&p.acc_z,   (void*)&k};
  NBX_HIP(hipLaunchKernel(V.k[chk][kk], dim3(e->grid[v][chk][kk]), dim3(64 * V.ws), args, 0, s));
  return NBX_OK;
}

int ens_step_launch(nbx_ensemble* e, float G, float eps2, float dt, hipStream_t s, int variant) {
  NBX_TRY(verlet_drift_launch(&e->p, dt, true, s));
  KickArgs k{e->p.vel_x, e->p.vel_y, e->p.vel_z, e->p.acc_old_x, e->p.acc_old_y,
             e->p.acc_old_z, 0.5f * dt};
  return ens_forces_launch(e, G, eps2, &k, s, variant);
}

void ens_release(nbx_ensemble* e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  nbx_particles_free(&e->p);
  void* bufs[] = {e->d_off, e->d_n, e->d_queue, e->d_blk_sys, e->d_blk_first, e->d_epart};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (auto* it : e->d_items)
    if (it) (void)hipFree(it);
  for (auto* it : e->d_part_off)
    if (it) (void)hipFree(it);
  if (e->d_part) (void)hipFree(e->d_part);
  if (e->seq) (void)nbx_direct_destroy(e->seq);
  if (e->own_stream) (void)hipStreamDestroy(e->own_stream);
}

}  // namespace
}  // namespace nbx

using namespace nbx;

extern "C" {

int nbx_ens_create(int nsys, const int64_t* counts, nbx_ensemble** out) {
  if (!out) return fail(NBX_ERR_INVALID, "nbx_ens_create: null output");
  *out = nullptr;
  if (nsys < 1 || !counts) return fail(NBX_ERR_INVALID, "nbx_ens_create: need >= 1 system");
  nbx_ensemble* e = new (std::nothrow) nbx_ensemble();
  if (!e) return fail(NBX_ERR_NOMEM, "nbx_ens_create: host allocation");
  e->nsys = nsys;
  e->offsets.assign(nsys + 1, 0);
  for (int s = 0; s < nsys; ++s) {
    if (counts[s] < 1 || counts[s] > INT32_MAX) {
      delete e;
      return fail(NBX_ERR_INVALID, "system %d has %lld particles", s, (long long)counts[s]);
    }
    e->offsets[s + 1] = e->offsets[s] + counts[s];
  }
  e->total = (size_t)e->offsets[nsys];
  if (e->total > (size_t)INT32_MAX) {
    delete e;
    return fail(NBX_ERR_INVALID, "ensemble of %zu particles exceeds 2^31-1", e->total);
  }
  int rc = nbx_particles_alloc(&e->p, e->total);
  std::vector<int32_t> off(nsys), n(nsys);
  for (int s = 0; s < nsys; ++s) {
    off[s] = (int32_t)e->offsets[s];
    n[s] = (int32_t)counts[s];
  }
  auto dmal = [&](void** p, size_t bytes) {
    if (rc == NBX_OK && hipMalloc(p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      rc = fail(NBX_ERR_NOMEM, "ensemble buffers (%zu bytes)", bytes);
    }
  };
  dmal((void**)&e->d_off, 4 * nsys);
  dmal((void**)&e->d_n, 4 * nsys);
  dmal((void**)&e->d_queue, 2 * sizeof(int));
  if (rc == NBX_OK) {
    if (hipMemcpy(e->d_off, off.data(), 4 * nsys, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(e->d_n, n.data(), 4 * nsys, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(e->d_queue, 0, 2 * sizeof(int)) != hipSuccess)
      rc = fail(NBX_ERR_HIP, "ensemble set-up copies failed");
  }
  // work items per tile shape: every (system, tile), largest systems first
  std::vector<int> order(nsys);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
  hipDeviceProp_t prop;
  if (rc == NBX_OK && hipGetDeviceProperties(&prop, 0) != hipSuccess)
    rc = fail(NBX_ERR_HIP, "hipGetDeviceProperties failed");
  size_t part_floats = 0;
  for (int v = 0; v < kEnsNumVariants && rc == NBX_OK; ++v) {
    const EnsVariant& V = kEnsVariants[v];
    std::vector<int2> items;
    if (!V.sym_block) {
      const int tile = 64 * V.ipl;
      for (int s : order)
        for (int t = 0; t * tile < n[s]; ++t) items.push_back(make_int2(s, t));
    } else {
      // owned tiles of every system, diagonal tiles first; partial sums back to back
      std::vector<int64_t> poff(nsys);
      size_t at = 0;
      for (int s = 0; s < nsys; ++s) {
        poff[s] = (int64_t)at;
        at += sym_partial_floats((size_t)n[s], V.sym_block);
      }
      part_floats = std::max(part_floats, at);
      for (int s : order) {
        const int K = (n[s] + V.sym_block - 1) / V.sym_block, D = K / 2;
        for (int t = 0; t < K + K * D; ++t)
          if (t < K || (t - K) % D + 1 <= sym_rowlen((t - K) / D, K))
            items.push_back(make_int2(s, t));
      }
      dmal((void**)&e->d_part_off[v], sizeof(int64_t) * nsys);
      if (rc == NBX_OK && hipMemcpy(e->d_part_off[v], poff.data(), sizeof(int64_t) * nsys,
                                    hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(NBX_ERR_HIP, "partial offset upload failed");
    }
    e->nitems[v] = (int)items.size();
    dmal((void**)&e->d_items[v], sizeof(int2) * items.size());
    if (rc == NBX_OK &&
        hipMemcpy(e->d_items[v], items.data(), sizeof(int2) * items.size(),
                  hipMemcpyHostToDevice) != hipSuccess)
      rc = fail(NBX_ERR_HIP, "work item upload failed");
    for (int c = 0; c < 2 && rc == NBX_OK; ++c)
      for (int k = 0; k < 2 && rc == NBX_OK; ++k) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, V.k[c][k], 64 * V.ws, 0) !=
            hipSuccess)
          rc = fail(NBX_ERR_HIP, "occupancy query failed");
        e->grid[v][c][k] = std::max(1, per_cu) * prop.multiProcessorCount;
      }
  }
  if (part_floats) dmal((void**)&e->d_part, sizeof(float) * part_floats);
  // energy blocks: (system, first target) pairs
  std::vector<int32_t> bs, bf;
  for (int s = 0; s < nsys; ++s)
    for (int f = 0; f < n[s]; f += 256) {
      bs.push_back(s);
      bf.push_back(f);
    }
  e->eblocks = (int)bs.size();
  dmal((void**)&e->d_blk_sys, 4 * bs.size());
  dmal((void**)&e->d_blk_first, 4 * bf.size());
  dmal((void**)&e->d_epart, 2 * sizeof(double) * bs.size());
  if (rc == NBX_OK &&
      (hipMemcpy(e->d_blk_sys, bs.data(), 4 * bs.size(), hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(e->d_blk_first, bf.data(), 4 * bf.size(), hipMemcpyHostToDevice) != hipSuccess))
    rc = fail(NBX_ERR_HIP, "energy block table upload failed");
  if (rc == NBX_OK && hipStreamCreateWithFlags(&e->own_stream, hipStreamNonBlocking) != hipSuccess)
    rc = fail(NBX_ERR_HIP, "stream creation failed");
  if (rc == NBX_OK) rc = nbx_direct_create((size_t)*std::max_element(n.begin(), n.end()), 0, &e->seq);
  e->variant = 4;  // ipl1_nb8_ws16: smallest tile, best for systems of a few thousand bodies
  if (rc != NBX_OK) {
    ens_release(e);
    delete e;
    return rc;
  }
  *out = e;
  return NBX_OK;
}

int nbx_ens_destroy(nbx_ensemble* e) {
  if (!e) return NBX_OK;
  ens_release(e);
  delete e;
  return NBX_OK;
}

int nbx_ens_particles(nbx_ensemble* e, nbx_particles* view, int64_t* offsets) {
  if (!e) return fail(NBX_ERR_INVALID, "nbx_ens_particles: null ensemble");
  if (view) *view = e->p;
  if (offsets) std::copy(e->offsets.begin(), e->offsets.end(), offsets);
  return NBX_OK;
}

int nbx_ens_forces(nbx_ensemble* e, float G, float eps2, nbx_stream s) {
  if (!e) return fail(NBX_ERR_INVALID, "nbx_ens_forces: null ensemble");
  return ens_forces_launch(e, G, eps2, nullptr, as_stream(s), -1);
}

int nbx_ens_steps(nbx_ensemble* e, float G, float eps2, float dt, int steps, int graph_chunk,
                  nbx_stream stream) {
  if (!e) return fail(NBX_ERR_INVALID, "nbx_ens_steps: null ensemble");
  if (steps < 0) return fail(NBX_ERR_INVALID, "nbx_ens_steps: negative step count");
  hipStream_t s = stream ? as_stream(stream) : e->own_stream;
  if (graph_chunk <= 0) {
    for (int k = 0; k < steps; ++k) NBX_TRY(ens_step_launch(e, G, eps2, dt, s, -1));
    return NBX_OK;
  }
  if (steps % graph_chunk)
    return fail(NBX_ERR_INVALID, "steps (%d) must be a multiple of graph_chunk (%d)", steps,
                graph_chunk);
  const auto key = std::make_tuple(graph_chunk, G, eps2, dt, e->variant, s);
  if (!e->exec || e->graph_key != key) {
    if (e->exec) {
      (void)hipGraphExecDestroy(e->exec);
      e->exec = nullptr;
    }
    if (e->graph) {
      (void)hipGraphDestroy(e->graph);
      e->graph = nullptr;
    }
    NBX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = NBX_OK;
    for (int k = 0; k < graph_chunk && rc == NBX_OK; ++k) rc = ens_step_launch(e, G, eps2, dt, s, -1);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &g);
    if (rc != NBX_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (ec != hipSuccess) return fail(NBX_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ec));
    e->graph = g;
    NBX_HIP(hipGraphInstantiate(&e->exec, g, nullptr, nullptr, 0));
    e->graph_key = key;
  }
  for (int c = 0; c < steps / graph_chunk; ++c) NBX_HIP(hipGraphLaunch(e->exec, s));
  return NBX_OK;
}

int nbx_ens_steps_sequential(nbx_ensemble* e, float G, float eps2, float dt, int steps,
                             nbx_stream stream) {
  if (!e) return fail(NBX_ERR_INVALID, "nbx_ens_steps_sequential: null ensemble");
  hipStream_t s = stream ? as_stream(stream) : e->own_stream;
  for (int k = 0; k < steps; ++k)
    for (int q = 0; q < e->nsys; ++q) {
      nbx_particles v = e->p;
      const size_t off = (size_t)e->offsets[q];
      float** f[13] = {&v.pos_x, &v.pos_y, &v.pos_z, &v.vel_x, &v.vel_y, &v.vel_z, &v.acc_x,
                       &v.acc_y, &v.acc_z, &v.acc_old_x, &v.acc_old_y, &v.acc_old_z, &v.mass};
      for (float** a : f) *a += off;
      v.count = (size_t)(e->offsets[q + 1] - e->offsets[q]);
      NBX_TRY(nbx_direct_compute_step(e->seq, &v, G, eps2, dt, s));
    }
  return NBX_OK;
}

int nbx_ens_set_variant(nbx_ensemble* e, int v) {
  if (!e || v < 0 || v >= kEnsNumVariants)
    return fail(NBX_ERR_INVALID, "nbx_ens_set_variant: variant %d out of range", v);
  e->variant = v;
  return NBX_OK;
}

int nbx_ens_autotune(nbx_ensemble* e, int reps, int* best_variant, double* best_ms) {
  if (!e) return fail(NBX_ERR_INVALID, "nbx_ens_autotune: null ensemble");
  if (reps < 1) reps = 3;
  hipStream_t s = e->own_stream;
  hipEvent_t a, b;
  NBX_HIP(hipEventCreate(&a));
  NBX_HIP(hipEventCreate(&b));
  int best = -1;
  float best_t = 1e30f;
  int rc = NBX_OK;
  for (int v = 0; v < kEnsNumVariants && rc == NBX_OK; ++v) {
    rc = ens_forces_launch(e, 1.0f, 1e-4f, nullptr, s, v);
    (void)hipEventRecord(a, s);
    for (int r = 0; r < reps && rc == NBX_OK; ++r) rc = ens_forces_launch(e, 1.0f, 1e-4f, nullptr, s, v);
    (void)hipEventRecord(b, s);
    if (hipEventSynchronize(b) != hipSuccess) rc = fail(NBX_ERR_HIP, "autotune sync failed");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    ms /= (float)reps;
    if (rc == NBX_OK && ms < best_t) {
      best_t = ms;
      best = v;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) return rc;
  e->variant = best;
  if (best_variant) *best_variant = best;
  if (best_ms) *best_ms = best_t;
  return NBX_OK;
}

int nbx_ens_energy(nbx_ensemble* e, float G, float eps2, double* ke, double* pe, nbx_stream stream) {
  if (!e || !ke || !pe) return fail(NBX_ERR_INVALID, "nbx_ens_energy: null argument");
  hipStream_t s = stream ? as_stream(stream) : e->own_stream;
  const nbx_particles& p = e->p;
  hipLaunchKernelGGL(ens_energy_kernel, dim3(e->eblocks), dim3(256), 0, s, p.pos_x, p.pos_y,
                     p.pos_z, p.vel_x, p.vel_y, p.vel_z, p.mass, e->d_off, e->d_n, e->d_blk_sys,
                     e->d_blk_first, (double)eps2, e->d_epart);
  NBX_LAUNCHED();
  std::vector<double> part(2 * (size_t)e->eblocks);
  NBX_HIP(hipMemcpyAsync(part.data(), e->d_epart, sizeof(double) * part.size(),
                         hipMemcpyDeviceToHost, s));
  NBX_HIP(hipStreamSynchronize(s));
  for (int q = 0; q < e->nsys; ++q) ke[q] = pe[q] = 0.0;
  int b = 0;
  for (int q = 0; q < e->nsys; ++q) {
    const int64_t nq = e->offsets[q + 1] - e->offsets[q];
    for (int64_t f = 0; f < nq; f += 256, ++b) {
      ke[q] += part[2 * b];
      pe[q] += part[2 * b + 1];
    }
    pe[q] *= -0.5 * (double)G;
  }
  return NBX_OK;
}

int nbx_ens_variant_count(void) { return kEnsNumVariants; }

}  // extern "C"
