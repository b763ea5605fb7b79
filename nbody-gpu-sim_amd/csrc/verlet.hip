// Velocity Verlet kernels (src/cuda/integrator.cu:11-49, order :175-189) and
// the FP64 energy / momentum reductions that replace the FP32 ones of
// integrator.cu:52-120,203-240.
//
// The elementwise kernels round exactly like the reference formulas evaluated
// in IEEE FP32 without contraction (add_rn / mul_rn), so they are
// bit-identical to the CPU restatement. They are HBM-bound: drift moves
// 48 B/particle (+24 B with the acc_old store), kick 48 B/particle.
#include <algorithm>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "verlet.hpp"

namespace nbx {

__global__ void drift_kernel(float* __restrict__ x, float* __restrict__ y, float* __restrict__ z,
                             const float* __restrict__ vx, const float* __restrict__ vy,
                             const float* __restrict__ vz, const float* __restrict__ ax,
                             const float* __restrict__ ay, const float* __restrict__ az,
                             float* __restrict__ aox, float* __restrict__ aoy,
                             float* __restrict__ aoz, int n, float dt, float dt2_half,
                             int store) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float a0 = ax[i], a1 = ay[i], a2 = az[i];
  if (store) {
    aox[i] = a0;
    aoy[i] = a1;
    aoz[i] = a2;
  }
  x[i] = add_rn(x[i], add_rn(mul_rn(vx[i], dt), mul_rn(a0, dt2_half)));
  y[i] = add_rn(y[i], add_rn(mul_rn(vy[i], dt), mul_rn(a1, dt2_half)));
  z[i] = add_rn(z[i], add_rn(mul_rn(vz[i], dt), mul_rn(a2, dt2_half)));
}

__global__ void store_kernel(const float* __restrict__ ax, const float* __restrict__ ay,
                             const float* __restrict__ az, float* __restrict__ aox,
                             float* __restrict__ aoy, float* __restrict__ aoz, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  aox[i] = ax[i];
  aoy[i] = ay[i];
  aoz[i] = az[i];
}

__global__ void kick_kernel(float* __restrict__ vx, float* __restrict__ vy,
                            float* __restrict__ vz, const float* __restrict__ aox,
                            const float* __restrict__ aoy, const float* __restrict__ aoz,
                            const float* __restrict__ ax, const float* __restrict__ ay,
                            const float* __restrict__ az, int n, float dt_half) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vx[i] = add_rn(vx[i], mul_rn(add_rn(aox[i], ax[i]), dt_half));
  vy[i] = add_rn(vy[i], mul_rn(add_rn(aoy[i], ay[i]), dt_half));
  vz[i] = add_rn(vz[i], mul_rn(add_rn(aoz[i], az[i]), dt_half));
}

constexpr int kElemBlock = 256;

int verlet_drift_launch(nbx_particles* d, float dt, bool store, hipStream_t s) {
  const int n = (int)d->count;
  const float dt2_half = (0.5f * dt) * dt;  // integrator.cu:17 evaluates 0.5f * dt * dt
  hipLaunchKernelGGL(drift_kernel, dim3(div_up(n, kElemBlock)), dim3(kElemBlock), 0, s, d->pos_x,
                     d->pos_y, d->pos_z, d->vel_x, d->vel_y, d->vel_z, d->acc_x, d->acc_y,
                     d->acc_z, d->acc_old_x, d->acc_old_y, d->acc_old_z, n, dt, dt2_half,
                     store ? 1 : 0);
  NBX_LAUNCHED();
  return NBX_OK;
}

int verlet_kick_launch(nbx_particles* d, float dt, hipStream_t s) {
  const int n = (int)d->count;
  hipLaunchKernelGGL(kick_kernel, dim3(div_up(n, kElemBlock)), dim3(kElemBlock), 0, s, d->vel_x,
                     d->vel_y, d->vel_z, d->acc_old_x, d->acc_old_y, d->acc_old_z, d->acc_x,
                     d->acc_y, d->acc_z, n, 0.5f * dt);
  NBX_LAUNCHED();
  return NBX_OK;
}

// ---------------------------------------------------------------- energies

constexpr int kRedBlock = 256;
constexpr int kRedMaxBlocks = 1024;

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum of 4 doubles into out[blockIdx.x*4 + k]; fixed order -> deterministic.
__device__ void block_sum4(double a, double b, double c, double e, double* out) {
  __shared__ double sh[4][kRedBlock / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  e = wave_sum(e);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sh[0][w] = a;
    sh[1][w] = b;
    sh[2][w] = c;
    sh[3][w] = e;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int k = 0; k < kRedBlock / 64; ++k) t += sh[threadIdx.x][k];
    out[blockIdx.x * 4 + threadIdx.x] = t;
  }
}

// KE = sum 1/2 m v^2 and P = sum m v, FP64, grid-stride over a fixed grid.
__global__ __launch_bounds__(kRedBlock) void kinetic_kernel(const float* __restrict__ vx,
                                                            const float* __restrict__ vy,
                                                            const float* __restrict__ vz,
                                                            const float* __restrict__ m, int n,
                                                            double* partial) {
  double ke = 0.0, px = 0.0, py = 0.0, pz = 0.0;
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock) {
    const double mi = m[i], ux = vx[i], uy = vy[i], uz = vz[i];
    ke += 0.5 * mi * (ux * ux + uy * uy + uz * uz);
    px += mi * ux;
    py += mi * uy;
    pz += mi * uz;
  }
  block_sum4(ke, px, py, pz, partial);
}

// phi_i = sum_{j != i} m_j / sqrt(|r_ij|^2 + eps2) with the pair term in FP64:
// v_rsq_f32 seed refined by one Newton step in FP64 (error ~1e-14). Bodies are
// read with wave-uniform addresses (scalar loads). partial gets sum_i m_i phi_i
// per block in slot 0 (slots 1..3 zero).
__global__ __launch_bounds__(kRedBlock) void potential_kernel(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    const float* __restrict__ m, const float* __restrict__ xlo, const float* __restrict__ ylo,
    const float* __restrict__ zlo, int n, double eps2, double* partial) {
  const int i = blockIdx.x * kRedBlock + threadIdx.x;
  const bool ok = i < n;
  const int ic = ok ? i : 0;
  double xi = x[ic], yi = y[ic], zi = z[ic];
  if (xlo) {
    xi += (double)xlo[ic];
    yi += (double)ylo[ic];
    zi += (double)zlo[ic];
  }
  double phi = 0.0;
  for (int j = 0; j < n; ++j) {
    double xj = x[j], yj = y[j], zj = z[j];
    if (xlo) {
      xj += (double)xlo[j];
      yj += (double)ylo[j];
      zj += (double)zlo[j];
    }
    const double dx = xj - xi, dy = yj - yi, dz = zj - zi;
    const double r2 = dx * dx + dy * dy + dz * dz + eps2;
    double r = (double)__builtin_amdgcn_rsqf((float)r2);
    r = r * (1.5 - 0.5 * r2 * r * r);
    r = r * (1.5 - 0.5 * r2 * r * r);
    phi += (j == i) ? 0.0 : (double)m[j] * r;
  }
  const double v = ok ? (double)m[i] * phi : 0.0;
  block_sum4(v, 0.0, 0.0, 0.0, partial);
}

__global__ void finish_kernel(const double* partial, int nblocks, double* out) {
  // one thread per slot, sequential over blocks: fixed order
  const int k = threadIdx.x;
  if (k >= 4) return;
  double t = 0.0;
  for (int b = 0; b < nblocks; ++b) t += partial[b * 4 + k];
  out[k] = t;
}

static Scratch g_energy_scratch;
static std::mutex g_energy_mu;

int energy_launch(const nbx_particles* d, const float* xlo, const float* ylo, const float* zlo,
                  float G, float eps2, double* ke, double* pe, double* mom, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_energy_mu);
  const int n = (int)d->count;
  const int pblocks = (int)div_up(n, kRedBlock);
  const int kblocks = (int)std::min<size_t>(div_up(n, kRedBlock), kRedMaxBlocks);
  const size_t need = (size_t)(std::max(pblocks, kblocks) * 4 + 8) * sizeof(double);
  NBX_TRY(g_energy_scratch.ensure(need));
  double* partial = static_cast<double*>(g_energy_scratch.ptr);
  double* result = partial + (size_t)std::max(pblocks, kblocks) * 4;
  double host[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (ke || mom) {
    hipLaunchKernelGGL(kinetic_kernel, dim3(kblocks), dim3(kRedBlock), 0, s, d->vel_x, d->vel_y,
                       d->vel_z, d->mass, n, partial);
    NBX_LAUNCHED();
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(64), 0, s, partial, kblocks, result);
    NBX_LAUNCHED();
  }
  if (pe) {
    hipLaunchKernelGGL(potential_kernel, dim3(pblocks), dim3(kRedBlock), 0, s, d->pos_x,
                       d->pos_y, d->pos_z, d->mass, xlo, ylo, zlo, n, (double)eps2, partial);
    NBX_LAUNCHED();
    hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(64), 0, s, partial, pblocks, result + 4);
    NBX_LAUNCHED();
  }
  NBX_HIP(hipMemcpyAsync(host, result, 8 * sizeof(double), hipMemcpyDeviceToHost, s));
  NBX_HIP(hipStreamSynchronize(s));
  if (ke) *ke = host[0];
  if (mom) {
    mom[0] = host[1];
    mom[1] = host[2];
    mom[2] = host[3];
  }
  if (pe) *pe = -0.5 * (double)G * host[4];
  return NBX_OK;
}

}  // namespace nbx

using namespace nbx;

extern "C" {

int nbx_store_old_accelerations(nbx_particles* d, nbx_stream s) {
  NBX_TRY(check_particles(d, "nbx_store_old_accelerations"));
  const int n = (int)d->count;
  hipLaunchKernelGGL(store_kernel, dim3(div_up(n, kElemBlock)), dim3(kElemBlock), 0,
                     as_stream(s), d->acc_x, d->acc_y, d->acc_z, d->acc_old_x, d->acc_old_y,
                     d->acc_old_z, n);
  NBX_LAUNCHED();
  return NBX_OK;
}

int nbx_update_positions(nbx_particles* d, float dt, nbx_stream s) {
  NBX_TRY(check_particles(d, "nbx_update_positions"));
  return verlet_drift_launch(d, dt, false, as_stream(s));
}

int nbx_update_velocities(nbx_particles* d, float dt, nbx_stream s) {
  NBX_TRY(check_particles(d, "nbx_update_velocities"));
  return verlet_kick_launch(d, dt, as_stream(s));
}

int nbx_verlet_drift(nbx_particles* d, float dt, nbx_stream s) {
  NBX_TRY(check_particles(d, "nbx_verlet_drift"));
  return verlet_drift_launch(d, dt, true, as_stream(s));
}

int nbx_energy(const nbx_particles* d, float G, float eps2, double* ke, double* pe,
               double momentum[3], nbx_stream s) {
  NBX_TRY(check_particles(d, "nbx_energy"));
  return energy_launch(d, nullptr, nullptr, nullptr, G, eps2, ke, pe, momentum, as_stream(s));
}

}  // extern "C"
