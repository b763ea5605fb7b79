// Direct N^2 force kernels, launch dispatch over tile variants, and the
// in-project autotuner. Replaces launchDirectForceKernel / DirectForceCalculator
// (src/cuda/force_direct.cu:88-106).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <vector>

#include "common.hpp"
#include "direct_core.hpp"
#include "direct_sym.hpp"
#include "verlet.hpp"

namespace nbx {

template <int IPL, int NB, int WS, bool CHECKED, bool KICK>
__global__ __launch_bounds__(64 * WS) void direct_forces_kernel(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    const float* __restrict__ m, int n, int i0, int iend, double G, float eps2, float* ax,
    float* ay, float* az, KickArgs kick) {
  __shared__ double red[WS * 3 * 64 * IPL];
  const int ibase = i0 + (int)blockIdx.x * 64 * IPL;
  direct_tile<IPL, NB, WS, CHECKED, KICK>(x, y, z, m, n, ibase, iend, G, eps2, ax, ay, az, kick,
                                          red);
}

struct DirectVariant {
  const char* name;
  int ipl;
  int nb;  // bodies per stage group (one-sided), 0 for the symmetric kernel
  int ws;  // waves per workgroup
  // Super-block size of the symmetric kernel, 0 for one-sided shapes. Symmetric
  // shapes only evaluate whole systems ([i0, iend) = [0, n)).
  int sym_block;
  // launches over targets [i0, iend) of the system d; symmetric shapes use w
  int (*launch)(const nbx_particles* d, int i0, int iend, double G, float eps2, bool checked,
                const KickArgs* kick, const SymWork* w, hipStream_t s);
};

template <int IPL, int NB, int WS>
static int launch_variant(const nbx_particles* d, int i0, int iend, double G, float eps2,
                           bool checked, const KickArgs* kick, const SymWork*, hipStream_t s) {
  const int n = (int)d->count;
  const int tiles = (int)div_up((size_t)(iend - i0), 64 * IPL);
  dim3 grid(tiles), block(64 * WS);
  KickArgs k{};
  if (kick) k = *kick;
  if (checked) {
    if (kick)
      hipLaunchKernelGGL((direct_forces_kernel<IPL, NB, WS, true, true>), grid, block, 0, s, d->pos_x,
                         d->pos_y, d->pos_z, d->mass, n, i0, iend, G, eps2, d->acc_x, d->acc_y,
                         d->acc_z, k);
    else
      hipLaunchKernelGGL((direct_forces_kernel<IPL, NB, WS, true, false>), grid, block, 0, s,
                         d->pos_x, d->pos_y, d->pos_z, d->mass, n, i0, iend, G, eps2, d->acc_x,
                         d->acc_y, d->acc_z, k);
  } else {
    if (kick)
      hipLaunchKernelGGL((direct_forces_kernel<IPL, NB, WS, false, true>), grid, block, 0, s,
                         d->pos_x, d->pos_y, d->pos_z, d->mass, n, i0, iend, G, eps2, d->acc_x,
                         d->acc_y, d->acc_z, k);
    else
      hipLaunchKernelGGL((direct_forces_kernel<IPL, NB, WS, false, false>), grid, block, 0, s,
                         d->pos_x, d->pos_y, d->pos_z, d->mass, n, i0, iend, G, eps2, d->acc_x,
                         d->acc_y, d->acc_z, k);
  }
  NBX_LAUNCHED();
  return NBX_OK;
}

template <int IPL, int W>
static int launch_sym(const nbx_particles* d, int, int, double G, float eps2, bool checked,
                      const KickArgs* kick, const SymWork* w, hipStream_t s) {
  return direct_sym_launch<IPL, W>(d, G, eps2, checked, kick, *w, s);
}

// One-sided shapes first, in the order the batched kernel shares, then the
// symmetric ones.
#define NBX_DIRECT_V(I, B, W) {"ipl" #I "_nb" #B "_ws" #W, I, B, W, 0, launch_variant<I, B, W>},
#define NBX_SYM_V(I, W) {"sym_ipl" #I "_ws" #W, I, 0, W, 64 * I * W, launch_sym<I, W>},
static const DirectVariant kVariants[] = {NBX_DIRECT_VARIANTS(NBX_DIRECT_V)
                                              NBX_SYM_VARIANTS(NBX_SYM_V)};
#undef NBX_DIRECT_V
#undef NBX_SYM_V
constexpr int kNumVariants = (int)(sizeof(kVariants) / sizeof(kVariants[0]));

// Autotuned choice per size class (floor(log2 n)); -1 = not tuned.
static std::mutex g_tune_mu;
static int g_forced = -1;
static int g_tuned[64];
static bool g_tuned_init = false;

static int size_class(size_t n) {
  int c = 0;
  while ((size_t(1) << (c + 1)) <= n && c < 62) ++c;
  return c;
}

// Defaults when nothing was tuned, from sweeps on MI355X (uniform-sphere
// bodies, equal masses): the one-sided kernel for small systems and for calls
// that bring no workspace, the symmetric ones above 2^14 bodies.
#define NBX_COUNT_V(I, B, W) +1
constexpr int kNumOneSided = 0 NBX_DIRECT_VARIANTS(NBX_COUNT_V);
#undef NBX_COUNT_V

static int one_sided_default(size_t n) {
  if (n >= (size_t(1) << 17)) return 1;  // ipl1_nb4_ws8
  if (n >= (size_t(1) << 15)) return 3;  // ipl1_nb8_ws8
  return 4;                              // ipl1_nb8_ws16
}

static int default_variant(size_t n) {
  // From forces-only timings of every shape (Plummer spheres of 4,096 to
  // 262,144 bodies, warm clocks):
  //   n >= 3 * 2^16: sym_ipl6_ws4, 1,536-body super-blocks, 4 waves per SIMD (6
  //     targets per lane take 1.5x the pairs of 4 per rotation step; 8 need
  //     163 VGPRs, 3 waves); 1.5-6 % faster than sym_ipl4_ws4 up to 2^20;
  //   2^17 <= n < 3 * 2^16: sym_ipl8_ws1 (512), 1-2.6 % faster than sym_ipl6_ws4;
  //   2^16 <= n < 2^17: sym_ipl6_ws1 (384), 4-11 % faster than sym_ipl6_ws4;
  //   2^14 < n < 2^16: sym_ipl4_ws1 (256), 1.7-1.9x the one-sided kernels;
  //   n <= 2^14: the one-sided kernels (a symmetric tile's partial sums and
  //     their reduction cost more than the halved pair work saves).
  if (n >= 3 * (size_t(1) << 16)) return kNumOneSided + 8;
  if (n >= (size_t(1) << 17)) return kNumOneSided + 11;
  if (n >= (size_t(1) << 16)) return kNumOneSided + 10;
  if (n > (size_t(1) << 14)) return kNumOneSided + 7;
  return one_sided_default(n);
}

int direct_pick_variant(size_t n) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (!g_tuned_init) {
    for (int& t : g_tuned) t = -1;
    g_tuned_init = true;
  }
  if (g_forced >= 0) return g_forced;
  const int c = size_class(n);
  if (g_tuned[c] >= 0) return g_tuned[c];
  return default_variant(n);
}

int direct_launch(const nbx_particles* d, size_t i0, size_t ni, float G, float eps2,
                  const KickArgs* kick, hipStream_t s, int variant, const SymWork* w) {
  if (ni == 0) return NBX_OK;
  if (!(eps2 >= 0.0f) || !std::isfinite(eps2))
    return fail(NBX_ERR_INVALID, "softening eps2 must be finite and >= 0 (got %g)", (double)eps2);
  int v = variant >= 0 ? variant : direct_pick_variant(d->count);
  if (kVariants[v].sym_block && (!w || i0 != 0 || ni != d->count)) v = one_sided_default(d->count);
  const bool checked = eps2 < kUncheckedMinEps2;
  return kVariants[v].launch(d, (int)i0, (int)(i0 + ni), (double)G, eps2, checked, kick, w, s);
}

int direct_sym_block_of(int variant) { return kVariants[variant].sym_block; }

// Workspace sizes for up to max_particles bodies: the FP64 carry covers the
// padded length of every symmetric shape (n rounded up to its super-block);
// the partials default to the largest one-round need over the symmetric
// shapes, capped at kDefaultPartialCap, and are at least the three slots of a
// one-distance round.
constexpr size_t kDefaultPartialCap = size_t(1) << 30;

static size_t padded(size_t n) {
  size_t p = n;
  for (const DirectVariant& V : kVariants)
    if (V.sym_block) p = std::max(p, (n + V.sym_block - 1) / V.sym_block * V.sym_block);
  return p;
}
static size_t min_partial_bytes(size_t n) { return 3 * 3 * padded(n) * sizeof(float); }

int direct_ws_create(size_t max_particles, size_t scratch_bytes, nbx_direct** out) {
  if (!out) return fail(NBX_ERR_INVALID, "nbx_direct_create: null output");
  *out = nullptr;
  if (max_particles == 0 || max_particles > (size_t)INT32_MAX)
    return fail(NBX_ERR_INVALID, "nbx_direct_create: max_particles %zu outside [1, 2^31-1]",
                max_particles);
  size_t one_round = 0;
  for (const DirectVariant& V : kVariants)
    if (V.sym_block)
      one_round = std::max(one_round, direct_sym_partial_bytes(max_particles, V.sym_block));
  const size_t part = std::min(one_round, scratch_bytes ? scratch_bytes : kDefaultPartialCap);
  if (part < one_round && part < min_partial_bytes(max_particles))
    return fail(NBX_ERR_INVALID,
                "nbx_direct_create: scratch of %zu bytes is below the %zu that %zu bodies need",
                scratch_bytes, min_partial_bytes(max_particles), max_particles);
  nbx_direct* op = new (std::nothrow) nbx_direct();
  if (!op) return fail(NBX_ERR_NOMEM, "nbx_direct_create: host allocation");
  op->max_particles = max_particles;
  if (hipGetDevice(&op->device) != hipSuccess) {
    delete op;
    return fail(NBX_ERR_HIP, "nbx_direct_create: hipGetDevice failed");
  }
  const size_t carry = part < one_round ? 3 * padded(max_particles) * sizeof(double) : 0;
  if (hipMalloc(reinterpret_cast<void**>(&op->work.part), part) != hipSuccess ||
      (carry && hipMalloc(reinterpret_cast<void**>(&op->work.carry), carry) != hipSuccess)) {
    (void)hipGetLastError();
    if (op->work.part) (void)hipFree(op->work.part);
    delete op;
    return fail(NBX_ERR_NOMEM, "nbx_direct_create: %zu + %zu bytes of device scratch", part,
                carry);
  }
  op->work.part_bytes = part;
  op->work.carry_bytes = carry;
  *out = op;
  return NBX_OK;
}

static int check_ws(const nbx_direct* op, const nbx_particles* d, const char* who) {
  if (!op) return fail(NBX_ERR_INVALID, "%s: null workspace", who);
  NBX_TRY(check_particles(d, who));
  if (d->count > op->max_particles)
    return fail(NBX_ERR_CAPACITY, "%s: %zu bodies exceed the workspace's %zu", who, d->count,
                op->max_particles);
  int dev = 0;
  NBX_HIP(hipGetDevice(&dev));
  if (dev != op->device)
    return fail(NBX_ERR_STATE, "%s: workspace of device %d used on device %d", who, op->device,
                dev);
  return NBX_OK;
}

// Scratch of the allocation-free entry point (nbx_direct_compute_noalloc): a
// fixed device array, one copy per GPU, held until the process ends. It takes
// the FP64 carry (3 doubles per padded body, 256-byte aligned) followed by the
// FP32 partial sums of one round; an evaluation whose partials fit whole takes
// it all as one round. Nothing in it needs clearing: a round's reduction reads
// only partials its own tiles wrote, and the carry is written by the first
// round before later ones read it. Calls that overlap on one GPU race on it.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's copy of `scratch`, looked up on the first call
// with the device current. Host code uses this address only: the array's own
// name is a host-side stand-in, not device memory.
static int noalloc_scratch(void** out) {
  int dev = 0;
  NBX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(NBX_ERR_STATE, "nbx_direct_compute_noalloc: device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(NBX_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return NBX_OK;
}

// The SymWork of an n-body evaluation with super-blocks of S inside the fixed
// scratch at base; NBX_ERR_CAPACITY beyond what it holds.
static int noalloc_work(void* base, size_t n, int S, SymWork* w) {
  unsigned char* b = static_cast<unsigned char*>(base);
  if (direct_sym_partial_bytes(n, S) <= kNoAllocScratchBytes) {
    *w = SymWork{reinterpret_cast<float*>(b), kNoAllocScratchBytes, nullptr, 0};
    return NBX_OK;
  }
  const size_t carry = (direct_sym_carry_bytes(n, S) + 255) / 256 * 256;
  if (carry >= kNoAllocScratchBytes ||
      direct_sym_round_len(n, S, kNoAllocScratchBytes - carry) < 1)
    return fail(NBX_ERR_CAPACITY,
                "nbx_direct_compute_noalloc: %zu bodies exceed its %zu-byte scratch with "
                "%d-body super-blocks (at most about %zu bodies)",
                n, kNoAllocScratchBytes, S, kNoAllocScratchBytes / 60);
  *w = SymWork{reinterpret_cast<float*>(b + carry), kNoAllocScratchBytes - carry,
               reinterpret_cast<double*>(b), carry};
  return NBX_OK;
}

}  // namespace nbx

using namespace nbx;

extern "C" {

int nbx_direct_forces(nbx_particles* d, float G, float eps2, int block_size, nbx_stream s) {
  NBX_TRY(check_particles(d, "nbx_direct_forces"));
  if (block_size < 1 || block_size > 1024)
    return fail(NBX_ERR_INVALID, "block size must be between 1 and 1024 (got %d)", block_size);
  return direct_launch(d, 0, d->count, G, eps2, nullptr, as_stream(s), -1, nullptr);
}

int nbx_direct_forces_range(nbx_particles* d, size_t i0, size_t ni, float G, float eps2,
                            nbx_stream s) {
  NBX_TRY(check_particles(d, "nbx_direct_forces_range"));
  if (i0 > d->count || ni > d->count - i0)
    return fail(NBX_ERR_INVALID, "target range [%zu, %zu) outside [0, %zu)", i0, i0 + ni,
                d->count);
  return direct_launch(d, i0, ni, G, eps2, nullptr, as_stream(s), -1, nullptr);
}

int nbx_direct_step(nbx_particles* d, float G, float eps2, float dt, nbx_stream s) {
  NBX_TRY(check_particles(d, "nbx_direct_step"));
  hipStream_t st = as_stream(s);
  NBX_TRY(verlet_drift_launch(d, dt, true, st));
  KickArgs k{d->vel_x, d->vel_y, d->vel_z, d->acc_old_x, d->acc_old_y, d->acc_old_z, 0.5f * dt};
  return direct_launch(d, 0, d->count, G, eps2, &k, st, -1, nullptr);
}

int nbx_direct_forces_kick(nbx_particles* d, float G, float eps2, float dt, nbx_stream s) {
  NBX_TRY(check_particles(d, "nbx_direct_forces_kick"));
  KickArgs k{d->vel_x, d->vel_y, d->vel_z, d->acc_old_x, d->acc_old_y, d->acc_old_z, 0.5f * dt};
  return direct_launch(d, 0, d->count, G, eps2, &k, as_stream(s), -1, nullptr);
}

int nbx_direct_create(size_t max_particles, size_t scratch_bytes, nbx_direct** out) {
  return direct_ws_create(max_particles, scratch_bytes, out);
}

int nbx_direct_destroy(nbx_direct* op) {
  if (!op) return NBX_OK;
  int rc = NBX_OK;
  if (op->work.part && hipFree(op->work.part) != hipSuccess)
    rc = fail(NBX_ERR_HIP, "nbx_direct_destroy: hipFree failed");
  if (op->work.carry && hipFree(op->work.carry) != hipSuccess)
    rc = fail(NBX_ERR_HIP, "nbx_direct_destroy: hipFree failed");
  delete op;
  return rc;
}

int nbx_direct_compute(nbx_direct* op, nbx_particles* d, float G, float eps2, nbx_stream s) {
  NBX_TRY(check_ws(op, d, "nbx_direct_compute"));
  return direct_launch(d, 0, d->count, G, eps2, nullptr, as_stream(s), -1, &op->work);
}

int nbx_direct_compute_kick(nbx_direct* op, nbx_particles* d, float G, float eps2, float dt,
                            nbx_stream s) {
  NBX_TRY(check_ws(op, d, "nbx_direct_compute_kick"));
  KickArgs k{d->vel_x, d->vel_y, d->vel_z, d->acc_old_x, d->acc_old_y, d->acc_old_z, 0.5f * dt};
  return direct_launch(d, 0, d->count, G, eps2, &k, as_stream(s), -1, &op->work);
}

int nbx_direct_compute_step(nbx_direct* op, nbx_particles* d, float G, float eps2, float dt,
                            nbx_stream s) {
  NBX_TRY(check_ws(op, d, "nbx_direct_compute_step"));
  hipStream_t st = as_stream(s);
  NBX_TRY(verlet_drift_launch(d, dt, true, st));
  KickArgs k{d->vel_x, d->vel_y, d->vel_z, d->acc_old_x, d->acc_old_y, d->acc_old_z, 0.5f * dt};
  return direct_launch(d, 0, d->count, G, eps2, &k, st, -1, &op->work);
}

int nbx_direct_compute_noalloc(nbx_particles* d, float G, float eps2, nbx_stream s) {
  NBX_TRY(check_particles(d, "nbx_direct_compute_noalloc"));
  const int v = direct_pick_variant(d->count);
  const int S = kVariants[v].sym_block;
  if (!S) return direct_launch(d, 0, d->count, G, eps2, nullptr, as_stream(s), v, nullptr);
  void* base = nullptr;
  NBX_TRY(noalloc_scratch(&base));
  SymWork w;
  NBX_TRY(noalloc_work(base, d->count, S, &w));
  return direct_launch(d, 0, d->count, G, eps2, nullptr, as_stream(s), v, &w);
}

int nbx_direct_plan(const nbx_direct* op, size_t n, int* variant, int* rounds,
                    size_t* partial_bytes) {
  if (!op) return fail(NBX_ERR_INVALID, "nbx_direct_plan: null workspace");
  if (n == 0 || n > op->max_particles)
    return fail(NBX_ERR_CAPACITY, "nbx_direct_plan: %zu bodies outside [1, %zu]", n,
                op->max_particles);
  const int v = direct_pick_variant(n);
  const int S = kVariants[v].sym_block;
  int r = 1;
  size_t bytes = 0;
  if (S) {
    const int R = direct_sym_round_len(n, S, op->work.part_bytes);
    if (R < 0) return fail(NBX_ERR_CAPACITY, "nbx_direct_plan: scratch too small for %zu", n);
    r = direct_sym_rounds(n, S, R);
    bytes = std::min(op->work.part_bytes, direct_sym_partial_bytes(n, S));
  }
  if (variant) *variant = v;
  if (rounds) *rounds = r;
  if (partial_bytes) *partial_bytes = bytes;
  return NBX_OK;
}

int nbx_direct_variant_count(void) { return kNumVariants; }

int nbx_direct_variant_name(int v, char* buf, size_t len) {
  if (v < 0 || v >= kNumVariants || !buf || !len)
    return fail(NBX_ERR_INVALID, "variant %d out of range", v);
  snprintf(buf, len, "%s", kVariants[v].name);
  return NBX_OK;
}

int nbx_direct_set_variant(int v) {
  if (v < -1 || v >= kNumVariants) return fail(NBX_ERR_INVALID, "variant %d out of range", v);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_forced = v;
  return NBX_OK;
}

int nbx_direct_variant_for(size_t n) { return direct_pick_variant(n); }

int nbx_direct_autotune(size_t n, int reps, int* best_variant, double* best_ms) {
  if (n == 0 || n > (size_t)INT32_MAX) return fail(NBX_ERR_INVALID, "autotune size %zu", n);
  if (reps < 1) reps = 3;
  // Synthetic bodies on a lattice-jittered cube; only timing matters here.
  nbx_particles d;
  NBX_TRY(nbx_particles_alloc(&d, n));
  std::vector<float> h(n);
  for (size_t i = 0; i < n; ++i) h[i] = (float)((i * 2654435761u) % 1000003u) * 1e-6f;
  int rc = NBX_OK;
  if ((rc = nbx_memcpy_htod(d.pos_x, h.data(), n * 4)) ||
      (rc = nbx_memcpy_htod(d.pos_y, h.data(), n * 4)) ||
      (rc = nbx_memcpy_htod(d.pos_z, h.data(), n * 4))) {
    nbx_particles_free(&d);
    return rc;
  }
  for (auto& v : h) v = 1.0f / (float)n;
  if ((rc = nbx_memcpy_htod(d.mass, h.data(), n * 4))) {
    nbx_particles_free(&d);
    return rc;
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    nbx_particles_free(&d);
    return fail(NBX_ERR_HIP, "hipEventCreate failed");
  }
  nbx_direct* ws = nullptr;
  if ((rc = direct_ws_create(n, 0, &ws))) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    nbx_particles_free(&d);
    return rc;
  }
  int best = -1;
  float best_t = 1e30f;
  for (int v = 0; v < kNumVariants && rc == NBX_OK; ++v) {
    rc = direct_launch(&d, 0, n, 1.0f, 1e-4f, nullptr, nullptr, v, &ws->work);  // warm-up
    if (rc) break;
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps && rc == NBX_OK; ++r)
      rc = direct_launch(&d, 0, n, 1.0f, 1e-4f, nullptr, nullptr, v, &ws->work);
    (void)hipEventRecord(e1, nullptr);
    if (hipEventSynchronize(e1) != hipSuccess) {
      rc = fail(NBX_ERR_HIP, "autotune synchronisation failed");
      break;
    }
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= (float)reps;
    if (ms < best_t) {
      best_t = ms;
      best = v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)nbx_direct_destroy(ws);
  nbx_particles_free(&d);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tuned_init) {
      for (int& t : g_tuned) t = -1;
      g_tuned_init = true;
    }
    g_tuned[size_class(n)] = best;
  }
  if (best_variant) *best_variant = best;
  if (best_ms) *best_ms = best_t;
  return NBX_OK;
}

}  // extern "C"
