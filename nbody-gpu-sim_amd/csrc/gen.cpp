// Deterministic synthetic inputs (host). Every value is drawn in double from
// SplitMix64 streams (Steele, Lea & Flood 2014) and rounded to float once, so
// the same bytes reach the GPU path and the CPU oracle on any compiler.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "nbx.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
  double open_uniform() {                                          // (0, 1)
    double u;
    do u = uniform();
    while (u <= 0.0);
    return u;
  }
};

constexpr double kPi = 3.14159265358979323846;

void isotropic(SplitMix64& r, double len, double& ox, double& oy, double& oz) {
  const double cz = 2.0 * r.uniform() - 1.0;
  const double ph = 2.0 * kPi * r.uniform();
  const double sz = std::sqrt(std::max(0.0, 1.0 - cz * cz));
  ox = len * sz * std::cos(ph);
  oy = len * sz * std::sin(ph);
  oz = len * cz;
}

struct Body {
  double x, y, z, vx, vy, vz;
};

// Plummer sphere, G = 1, total mass M, scale a (Aarseth, Henon & Wielen 1974).
void plummer(size_t n, SplitMix64& r, double a, double M, double mass_cut, std::vector<Body>& out) {
  if (!(mass_cut > 0.0) || mass_cut > 1.0) mass_cut = 0.999;
  out.resize(n);
  for (size_t i = 0; i < n; ++i) {
    double u;
    do u = r.uniform() * mass_cut;
    while (u <= 0.0);
    const double rad = a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
    Body& b = out[i];
    isotropic(r, rad, b.x, b.y, b.z);
    // speed: q = v/v_esc with density q^2 (1-q^2)^{7/2}, max below 0.1
    double q, g;
    do {
      q = r.uniform();
      g = 0.1 * r.uniform();
    } while (g > q * q * std::pow(1.0 - q * q, 3.5));
    const double vesc = std::sqrt(2.0 * M / a) * std::pow(1.0 + (rad * rad) / (a * a), -0.25);
    isotropic(r, q * vesc, b.vx, b.vy, b.vz);
  }
}

void recentre(std::vector<Body>& b) {
  double c[6] = {0, 0, 0, 0, 0, 0};
  for (const Body& p : b) {
    c[0] += p.x;
    c[1] += p.y;
    c[2] += p.z;
    c[3] += p.vx;
    c[4] += p.vy;
    c[5] += p.vz;
  }
  const double inv = b.empty() ? 0.0 : 1.0 / (double)b.size();
  for (Body& p : b) {
    p.x -= c[0] * inv;
    p.y -= c[1] * inv;
    p.z -= c[2] * inv;
    p.vx -= c[3] * inv;
    p.vy -= c[4] * inv;
    p.vz -= c[5] * inv;
  }
}

void emit(const std::vector<Body>& b, size_t off, double mass, float* x, float* y, float* z,
          float* vx, float* vy, float* vz, float* m) {
  for (size_t i = 0; i < b.size(); ++i) {
    x[off + i] = (float)b[i].x;
    y[off + i] = (float)b[i].y;
    z[off + i] = (float)b[i].z;
    vx[off + i] = (float)b[i].vx;
    vy[off + i] = (float)b[i].vy;
    vz[off + i] = (float)b[i].vz;
    m[off + i] = (float)mass;
  }
}

bool outputs_ok(const float* x, const float* y, const float* z, const float* vx, const float* vy,
                const float* vz, const float* m) {
  return x && y && z && vx && vy && vz && m;
}

}  // namespace

extern "C" {

int nbx_gen_plummer(size_t n, uint64_t seed, double a, double M, double mass_cut, float* x,
                    float* y, float* z, float* vx, float* vy, float* vz, float* m) {
  if (!outputs_ok(x, y, z, vx, vy, vz, m) || !(a > 0.0) || !(M > 0.0)) return NBX_ERR_INVALID;
  SplitMix64 r(seed);
  std::vector<Body> b;
  plummer(n, r, a, M, mass_cut, b);
  recentre(b);
  emit(b, 0, n ? M / (double)n : 0.0, x, y, z, vx, vy, vz, m);
  return NBX_OK;
}

int nbx_gen_uniform_sphere(size_t n, uint64_t seed, double R, double M, double sigma, float* x,
                           float* y, float* z, float* vx, float* vy, float* vz, float* m) {
  if (!outputs_ok(x, y, z, vx, vy, vz, m) || !(R > 0.0) || !(M > 0.0) || sigma < 0.0)
    return NBX_ERR_INVALID;
  SplitMix64 r(seed);
  std::vector<Body> b(n);
  for (size_t i = 0; i < n; ++i) {
    const double rad = R * std::cbrt(r.uniform());
    isotropic(r, rad, b[i].x, b[i].y, b[i].z);
    double g[3];
    for (int k = 0; k < 3; k += 2) {  // Box-Muller pairs
      const double u1 = r.open_uniform(), u2 = r.uniform();
      const double rr = std::sqrt(-2.0 * std::log(u1));
      g[k] = rr * std::cos(2.0 * kPi * u2);
      if (k + 1 < 3) g[k + 1] = rr * std::sin(2.0 * kPi * u2);
    }
    b[i].vx = sigma * g[0];
    b[i].vy = sigma * g[1];
    b[i].vz = sigma * g[2];
  }
  emit(b, 0, n ? M / (double)n : 0.0, x, y, z, vx, vy, vz, m);
  return NBX_OK;
}

int nbx_gen_two_plummer(size_t n, uint64_t seed, double a, double M_each, double sep,
                        double vrel, float* x, float* y, float* z, float* vx, float* vy,
                        float* vz, float* m) {
  if (!outputs_ok(x, y, z, vx, vy, vz, m) || (n % 2) != 0 || !(a > 0.0) || !(M_each > 0.0))
    return NBX_ERR_INVALID;
  const size_t h = n / 2;
  for (int s = 0; s < 2; ++s) {
    SplitMix64 r(seed + 0x632BE59BD9B4E019ull * (uint64_t)(s + 1));
    std::vector<Body> b;
    plummer(h, r, a, M_each, 0.999, b);
    recentre(b);
    const double cx = (s == 0 ? -0.5 : 0.5) * sep;
    const double ux = (s == 0 ? 0.5 : -0.5) * vrel;
    for (Body& p : b) {
      p.x += cx;
      p.vx += ux;
    }
    emit(b, s * h, h ? M_each / (double)h : 0.0, x, y, z, vx, vy, vz, m);
  }
  return NBX_OK;
}

int nbx_gen_ensemble_sizes(int nsys, uint64_t seed, int64_t nmin, int64_t nmax,
                           int64_t* counts) {
  if (nsys < 0 || !counts || nmin < 1 || nmax < nmin) return NBX_ERR_INVALID;
  SplitMix64 r(seed);
  const double l0 = std::log((double)nmin), l1 = std::log((double)nmax);
  for (int s = 0; s < nsys; ++s) {
    int64_t c = (int64_t)std::floor(std::exp(l0 + (l1 - l0) * r.uniform()));
    if (c < nmin) c = nmin;
    if (c > nmax) c = nmax;
    counts[s] = c;
  }
  return NBX_OK;
}

}  // extern "C"
