/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle.h for the rules and for
 * what pins each function). Each function cites the reference file:line whose
 * arithmetic it restates. Compiled with -ffp-contract=off: float expressions
 * round exactly in the order written here.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static int pick_threads(int nthreads) {
  if (nthreads > 0) return nthreads;
  return orc_max_threads();
}

/* ------------------------------------------------------------------ direct */

/* force_direct.cu:109-117; Vec3 ops of types.hpp:55-94 (dot = (x*x + y*y) + z*z). */
void orc_pair_accel(const float p1[3], const float p2[3], float m2, float G, float eps,
                    float out[3]) {
  const float rx = p2[0] - p1[0];
  const float ry = p2[1] - p1[1];
  const float rz = p2[2] - p1[2];
  const float len2 = (rx * rx + ry * ry) + rz * rz;
  const float dist2 = len2 + eps * eps;
  const float inv = 1.0f / sqrtf(dist2);
  const float inv3 = (inv * inv) * inv;
  const float f = (G * m2) * inv3;
  out[0] = rx * f;
  out[1] = ry * f;
  out[2] = rz * f;
}

/* tests/test_force_calculation.cpp:131-140: ref += pair(pi, pj, m_j) for j != i, ascending j. */
void orc_direct_f32(size_t n, const float* x, const float* y, const float* z, const float* m,
                    float G, float eps, const int64_t* targets, size_t ntargets, float* ax,
                    float* ay, float* az, int nthreads) {
  const size_t nt = targets ? ntargets : n;
  int th = pick_threads(nthreads);
  (void)th;
#pragma omp parallel for schedule(dynamic, 16) num_threads(th)
  for (long long k = 0; k < (long long)nt; ++k) {
    const size_t i = targets ? (size_t)targets[k] : (size_t)k;
    const float pi[3] = {x[i], y[i], z[i]};
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (size_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const float pj[3] = {x[j], y[j], z[j]};
      float a[3];
      orc_pair_accel(pi, pj, m[j], G, eps, a);
      sx += a[0];
      sy += a[1];
      sz += a[2];
    }
    ax[k] = sx;
    ay[k] = sy;
    az[k] = sz;
  }
}

void orc_direct_f64_pos64(size_t n, const double* x, const double* y, const double* z,
                          const float* m, double G, double eps2, const int64_t* targets,
                          size_t ntargets, double* ax, double* ay, double* az, int nthreads) {
  const size_t nt = targets ? ntargets : n;
  int th = pick_threads(nthreads);
  (void)th;
#pragma omp parallel for schedule(dynamic, 16) num_threads(th)
  for (long long k = 0; k < (long long)nt; ++k) {
    const size_t i = targets ? (size_t)targets[k] : (size_t)k;
    const double xi = x[i], yi = y[i], zi = z[i];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (size_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const double dx = x[j] - xi, dy = y[j] - yi, dz = z[j] - zi;
      const double r2 = dx * dx + dy * dy + dz * dz + eps2;
      const double inv = 1.0 / sqrt(r2);
      const double f = G * (double)m[j] * inv * inv * inv;
      sx += f * dx;
      sy += f * dy;
      sz += f * dz;
    }
    ax[k] = sx;
    ay[k] = sy;
    az[k] = sz;
  }
}

void orc_direct_f64(size_t n, const float* x, const float* y, const float* z, const float* m,
                    double G, double eps2, const int64_t* targets, size_t ntargets, double* ax,
                    double* ay, double* az, int nthreads) {
  double* xd = (double*)malloc(n * sizeof(double));
  double* yd = (double*)malloc(n * sizeof(double));
  double* zd = (double*)malloc(n * sizeof(double));
  for (size_t i = 0; i < n; ++i) {
    xd[i] = x[i];
    yd[i] = y[i];
    zd[i] = z[i];
  }
  orc_direct_f64_pos64(n, xd, yd, zd, m, G, eps2, targets, ntargets, ax, ay, az, nthreads);
  free(xd);
  free(yd);
  free(zd);
}

/* -------------------------------------------------------------- integrator */

/* integrator.cu:40-49 */
void orc_store_old(size_t n, const float* ax, const float* ay, const float* az, float* aox,
                   float* aoy, float* aoz) {
  for (size_t i = 0; i < n; ++i) {
    aox[i] = ax[i];
    aoy[i] = ay[i];
    aoz[i] = az[i];
  }
}

/* integrator.cu:12-22: dt2_half = 0.5f*dt*dt; x += v*dt + a*dt2_half. */
void orc_drift(size_t n, float* x, float* y, float* z, const float* vx, const float* vy,
               const float* vz, const float* ax, const float* ay, const float* az, float dt) {
  const float dt2_half = (0.5f * dt) * dt;
  for (size_t i = 0; i < n; ++i) {
    x[i] = x[i] + (vx[i] * dt + ax[i] * dt2_half);
    y[i] = y[i] + (vy[i] * dt + ay[i] * dt2_half);
    z[i] = z[i] + (vz[i] * dt + az[i] * dt2_half);
  }
}

/* integrator.cu:25-37: dt_half = 0.5f*dt; v += (a_old + a_new)*dt_half. */
void orc_kick(size_t n, float* vx, float* vy, float* vz, const float* aox, const float* aoy,
              const float* aoz, const float* ax, const float* ay, const float* az, float dt) {
  const float dt_half = 0.5f * dt;
  for (size_t i = 0; i < n; ++i) {
    vx[i] = vx[i] + (aox[i] + ax[i]) * dt_half;
    vy[i] = vy[i] + (aoy[i] + ay[i]) * dt_half;
    vz[i] = vz[i] + (aoz[i] + az[i]) * dt_half;
  }
}

/* particle_system.cpp:45-61 (a(0)) then integrator.cu:175-189 per step. */
void orc_run_direct(size_t n, float* x, float* y, float* z, float* vx, float* vy, float* vz,
                    float* ax, float* ay, float* az, float* aox, float* aoy, float* aoz,
                    const float* m, float G, float eps, float dt, int steps, int nthreads) {
  orc_direct_f32(n, x, y, z, m, G, eps, NULL, 0, ax, ay, az, nthreads);
  for (int s = 0; s < steps; ++s) {
    orc_store_old(n, ax, ay, az, aox, aoy, aoz);
    orc_drift(n, x, y, z, vx, vy, vz, ax, ay, az, dt);
    orc_direct_f32(n, x, y, z, m, G, eps, NULL, 0, ax, ay, az, nthreads);
    orc_kick(n, vx, vy, vz, aox, aoy, aoz, ax, ay, az, dt);
  }
}

void orc_energy_f64(size_t n, const float* x, const float* y, const float* z, const float* vx,
                    const float* vy, const float* vz, const float* m, double G, double eps2,
                    double* ke, double* pe, double mom[3], int nthreads) {
  double k = 0.0, px = 0.0, py = 0.0, pz = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double mi = m[i], ux = vx[i], uy = vy[i], uz = vz[i];
    k += 0.5 * mi * (ux * ux + uy * uy + uz * uz);
    px += mi * ux;
    py += mi * uy;
    pz += mi * uz;
  }
  double p = 0.0;
  int th = pick_threads(nthreads);
  (void)th;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : p) num_threads(th)
  for (long long i = 0; i < (long long)n; ++i) {
    const double xi = x[i], yi = y[i], zi = z[i], mi = m[i];
    double row = 0.0;
    for (size_t j = (size_t)i + 1; j < n; ++j) {
      const double dx = (double)x[j] - xi, dy = (double)y[j] - yi, dz = (double)z[j] - zi;
      row += (double)m[j] / sqrt(dx * dx + dy * dy + dz * dz + eps2);
    }
    p += mi * row;
  }
  *ke = k;
  *pe = -G * p;
  mom[0] = px;
  mom[1] = py;
  mom[2] = pz;
}

/* -------------------------------------------------------------- Barnes-Hut */

/* force_barnes_hut.cu:34-40: spread the low 10 bits of v to every third bit. */
uint32_t orc_expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

static uint32_t quantise_axis(float t) {
  /* force_barnes_hut.cu:45-47: clamp to [0,1] in float, scale by 1023, truncate. */
  const float c = fminf(fmaxf(t, 0.0f), 1.0f);
  return (uint32_t)(c * 1023.0f);
}

/* force_barnes_hut.cu:42-50: x takes the top bit of each triple. */
uint32_t orc_morton_code(float x, float y, float z) {
  return (orc_expand_bits(quantise_axis(x)) << 2) | (orc_expand_bits(quantise_axis(y)) << 1) |
         orc_expand_bits(quantise_axis(z));
}

/* force_barnes_hut.cu:78-122 and :266-306 (initial values :18-19). */
void orc_bbox(size_t n, const float* x, const float* y, const float* z, float bmin[3],
              float bmax[3]) {
  float lo[3] = {1e30f, 1e30f, 1e30f};
  float hi[3] = {-1e30f, -1e30f, -1e30f};
  for (size_t i = 0; i < n; ++i) {
    lo[0] = fminf(lo[0], x[i]);
    lo[1] = fminf(lo[1], y[i]);
    lo[2] = fminf(lo[2], z[i]);
    hi[0] = fmaxf(hi[0], x[i]);
    hi[1] = fmaxf(hi[1], y[i]);
    hi[2] = fmaxf(hi[2], z[i]);
  }
  for (int a = 0; a < 3; ++a) {
    bmin[a] = lo[a];
    bmax[a] = hi[a];
  }
}

/* force_barnes_hut.cu:308-322 (host sizes and reciprocals) and :125-139 (kernel). */
void orc_morton_codes(size_t n, const float* x, const float* y, const float* z,
                      const float bmin[3], const float bmax[3], uint32_t* codes, int32_t* idx) {
  const float sx = (bmax[0] - bmin[0]) + 0.001f;
  const float sy = (bmax[1] - bmin[1]) + 0.001f;
  const float sz = (bmax[2] - bmin[2]) + 0.001f;
  const float ix = 1.0f / sx, iy = 1.0f / sy, iz = 1.0f / sz;
  for (size_t i = 0; i < n; ++i) {
    const float nx = (x[i] - bmin[0]) * ix;
    const float ny = (y[i] - bmin[1]) * iy;
    const float nz = (z[i] - bmin[2]) * iz;
    codes[i] = orc_morton_code(nx, ny, nz);
    idx[i] = (int32_t)i;
  }
}

/* force_barnes_hut.cu:324-328: thrust::sort_by_key is a stable ascending sort.
 * Restated as a byte-wise LSD counting sort (stable by construction). */
void orc_stable_sort(size_t n, uint32_t* codes, int32_t* idx) {
  if (n < 2) return;
  uint32_t* k2 = (uint32_t*)malloc(n * sizeof(uint32_t));
  int32_t* v2 = (int32_t*)malloc(n * sizeof(int32_t));
  uint32_t* ks = codes;
  int32_t* vs = idx;
  uint32_t* kd = k2;
  int32_t* vd = v2;
  for (int shift = 0; shift < 32; shift += 8) {
    size_t cnt[257];
    memset(cnt, 0, sizeof(cnt));
    for (size_t i = 0; i < n; ++i) cnt[((ks[i] >> shift) & 0xFFu) + 1]++;
    for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
    for (size_t i = 0; i < n; ++i) {
      const size_t p = cnt[(ks[i] >> shift) & 0xFFu]++;
      kd[p] = ks[i];
      vd[p] = vs[i];
    }
    uint32_t* tk = ks;
    ks = kd;
    kd = tk;
    int32_t* tv = vs;
    vs = vd;
    vd = tv;
  }
  /* four passes: the result is back in the caller's arrays */
  free(k2);
  free(v2);
}

static void node_default(orc_node* nd) {
  memset(nd, 0, sizeof(*nd));
  for (int c = 0; c < 8; ++c) nd->children[c] = -1;
  nd->particle_index = -1;
  nd->is_leaf = 1;
  nd->particle_count = 0;
}

/* force_barnes_hut.cu:339-348: bit0 = x, bit1 = y, bit2 = z, ">=" goes high. */
static int octant_of(float px, float py, float pz, const orc_node* nd) {
  int o = 0;
  if (px >= nd->cx) o |= 1;
  if (py >= nd->cy) o |= 2;
  if (pz >= nd->cz) o |= 4;
  return o;
}

/* force_barnes_hut.cu:355-365 */
static void merge_into_leaf(orc_node* nd, int32_t pidx, float px, float py, float pz, float pm) {
  const float total = nd->total_mass + pm;
  const float cx = (nd->comx * nd->total_mass + px * pm) / total;
  const float cy = (nd->comy * nd->total_mass + py * pm) / total;
  const float cz = (nd->comz * nd->total_mass + pz * pm) / total;
  nd->comx = cx;
  nd->comy = cy;
  nd->comz = cz;
  nd->total_mass = total;
  nd->particle_count += 1;
  if (nd->particle_index < 0) nd->particle_index = pidx;
}

/* force_barnes_hut.cu:368-389. Returns the new node index or -1 if out of room. */
static int32_t new_leaf(orc_node* nodes, int32_t* count, size_t max_nodes, int32_t parent,
                        int oct, int32_t pidx, float px, float py, float pz, float pm) {
  if ((size_t)*count >= max_nodes) return -1;
  const int32_t id = (*count)++;
  orc_node* p = &nodes[parent];
  p->children[oct] = id;
  p->is_leaf = 0;
  const float hs = p->half_size * 0.5f;
  orc_node* c = &nodes[id];
  c->cx = p->cx + ((oct & 1) ? hs : -hs);
  c->cy = p->cy + ((oct & 2) ? hs : -hs);
  c->cz = p->cz + ((oct & 4) ? hs : -hs);
  c->half_size = hs;
  c->is_leaf = 1;
  c->particle_index = pidx;
  c->particle_count = 1;
  c->total_mass = pm;
  c->comx = px;
  c->comy = py;
  c->comz = pz;
  for (int k = 0; k < 8; ++k) c->children[k] = -1;
  return id;
}

#define ORC_MAX_DEPTH 32 /* force_barnes_hut.cu:24 */

/* force_barnes_hut.cu:391-543 */
int orc_build_tree(size_t n, const float* x, const float* y, const float* z, const float* m,
                   const int32_t* sorted_idx, const float bmin[3], const float bmax[3],
                   size_t max_nodes, orc_node* nodes, int32_t* node_count, int32_t* max_depth,
                   int32_t* budget_merges, int32_t* depth_merges) {
  if (max_nodes < 1) return -1;
  for (size_t k = 0; k < max_nodes; ++k) node_default(&nodes[k]);
  int32_t nb = 0, nd = 0;

  orc_node* root = &nodes[0];
  root->cx = (bmin[0] + bmax[0]) * 0.5f;
  root->cy = (bmin[1] + bmax[1]) * 0.5f;
  root->cz = (bmin[2] + bmax[2]) * 0.5f;
  const float ex = bmax[0] - bmin[0], ey = bmax[1] - bmin[1], ez = bmax[2] - bmin[2];
  float big = ex;
  if (ey > big) big = ey;
  if (ez > big) big = ez;
  root->half_size = big * 0.5f + 0.001f;
  root->is_leaf = 0;
  root->particle_index = -1;
  root->particle_count = (int32_t)n;

  int32_t count = 1;
  int32_t deepest = 0;
  for (size_t r = 0; r < n; ++r) {
    const int32_t pid = sorted_idx[r];
    const float px = x[pid], py = y[pid], pz = z[pid], pm = m[pid];
    int32_t cur = 0;
    int depth = 0;
    int merged = 0;
    while (depth < ORC_MAX_DEPTH) {
      orc_node* nd_ = &nodes[cur];
      if (nd_->is_leaf) {
        if ((size_t)count + ORC_MAX_DEPTH + 1 >= max_nodes) {
          merge_into_leaf(nd_, pid, px, py, pz, pm);
          merged = 1;
          nb++;
          break;
        }
        const int32_t old_id = nd_->particle_index;
        const float ox = nd_->comx, oy = nd_->comy, oz = nd_->comz, om = nd_->total_mass;
        nd_->is_leaf = 0;
        nd_->particle_index = -1;
        nd_->particle_count = 0;
        nd_->total_mass = 0.0f;
        nd_->comx = nd_->comy = nd_->comz = 0.0f;
        const int oo = octant_of(ox, oy, oz, nd_);
        if (nd_->children[oo] < 0) {
          if (new_leaf(nodes, &count, max_nodes, cur, oo, old_id, ox, oy, oz, om) < 0) return -1;
        }
      }
      const int o = octant_of(px, py, pz, &nodes[cur]);
      const int32_t ch = nodes[cur].children[o];
      if (ch < 0) {
        if (new_leaf(nodes, &count, max_nodes, cur, o, pid, px, py, pz, pm) < 0) return -1;
        break;
      }
      cur = ch;
      depth++;
    }
    if (!merged && depth >= ORC_MAX_DEPTH) {
      int32_t leaf = cur;
      while (!nodes[leaf].is_leaf) {
        for (int c = 0; c < 8; ++c) {
          if (nodes[leaf].children[c] >= 0) {
            leaf = nodes[leaf].children[c];
            break;
          }
        }
      }
      merge_into_leaf(&nodes[leaf], pid, px, py, pz, pm);
      nd++;
    }
    if (depth > deepest) deepest = depth;
  }

  /* force_barnes_hut.cu:512-538: reverse index order, children 0..7. */
  for (int32_t i = count - 1; i >= 0; --i) {
    orc_node* p = &nodes[i];
    if (p->is_leaf) continue;
    float tm = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
    int32_t cnt = 0;
    for (int c = 0; c < 8; ++c) {
      const int32_t ci = p->children[c];
      if (ci < 0) continue;
      const orc_node* q = &nodes[ci];
      tm += q->total_mass;
      sx += q->comx * q->total_mass;
      sy += q->comy * q->total_mass;
      sz += q->comz * q->total_mass;
      cnt += q->particle_count;
    }
    p->total_mass = tm;
    p->particle_count = cnt;
    if (tm > 0.0f) {
      p->comx = sx / tm;
      p->comy = sy / tm;
      p->comz = sz / tm;
    }
  }
  *node_count = count;
  *max_depth = deepest;
  if (budget_merges) *budget_merges = nb;
  if (depth_merges) *depth_merges = nd;
  return 0;
}

#define ORC_STACK 256 /* force_barnes_hut.cu:28 */

/* force_barnes_hut.cu:142-249 with 1.0f/sqrtf in place of rsqrtf. */
void orc_bh_forces(size_t n, const orc_node* nodes, int32_t node_count, const float* x,
                   const float* y, const float* z, const float* m, float theta, float G, float eps,
                   const int64_t* targets, size_t ntargets, float* ax, float* ay, float* az,
                   int nthreads) {
  const size_t nt = targets ? ntargets : n;
  const float theta2 = theta * theta;
  const float eps2 = eps * eps;
  int th = pick_threads(nthreads);
  (void)th;
#pragma omp parallel for schedule(dynamic, 64) num_threads(th)
  for (long long k = 0; k < (long long)nt; ++k) {
    const size_t i = targets ? (size_t)targets[k] : (size_t)k;
    const float xi = x[i], yi = y[i], zi = z[i], mi = m[i];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    int32_t stack[ORC_STACK];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
      const int32_t ni = stack[--sp];
      if (ni < 0 || ni >= node_count) continue;
      const orc_node* nd = &nodes[ni];
      if (nd->total_mass == 0.0f) continue;
      const int contains = fabsf(xi - nd->cx) <= nd->half_size &&
                           fabsf(yi - nd->cy) <= nd->half_size &&
                           fabsf(zi - nd->cz) <= nd->half_size;
      if (nd->is_leaf) {
        float me = nd->total_mass;
        float cx = nd->comx, cy = nd->comy, cz = nd->comz;
        if (contains) {
          me -= mi;
          if (me <= 0.0f) continue;
          cx = (cx * nd->total_mass - xi * mi) / me;
          cy = (cy * nd->total_mass - yi * mi) / me;
          cz = (cz * nd->total_mass - zi * mi) / me;
        }
        const float dx = cx - xi, dy = cy - yi, dz = cz - zi;
        const float d2 = ((dx * dx + dy * dy) + dz * dz) + eps2;
        const float inv = 1.0f / sqrtf(d2);
        const float f = (G * me) * ((inv * inv) * inv);
        sx += f * dx;
        sy += f * dy;
        sz += f * dz;
        continue;
      }
      const float dx = nd->comx - xi, dy = nd->comy - yi, dz = nd->comz - zi;
      const float d2 = ((dx * dx + dy * dy) + dz * dz) + eps2;
      const float size2 = (4.0f * nd->half_size) * nd->half_size;
      if (!contains && size2 / d2 < theta2) {
        const float inv = 1.0f / sqrtf(d2);
        const float f = (G * nd->total_mass) * ((inv * inv) * inv);
        sx += f * dx;
        sy += f * dy;
        sz += f * dz;
      } else if (sp + 8 <= ORC_STACK) {
        for (int c = 0; c < 8; ++c)
          if (nd->children[c] >= 0) stack[sp++] = nd->children[c];
      }
    }
    ax[k] = sx;
    ay[k] = sy;
    az[k] = sz;
  }
}

/* force_barnes_hut.cu:577-583 + :330-336 */
int orc_bh_full(size_t n, const float* x, const float* y, const float* z, const float* m,
                float theta, float G, float eps, size_t max_nodes, orc_node* nodes,
                uint32_t* codes, int32_t* idx, float* ax, float* ay, float* az, int nthreads) {
  float bmin[3], bmax[3];
  orc_bbox(n, x, y, z, bmin, bmax);
  orc_morton_codes(n, x, y, z, bmin, bmax, codes, idx);
  orc_stable_sort(n, codes, idx);
  int32_t cnt = 0, depth = 0;
  if (orc_build_tree(n, x, y, z, m, idx, bmin, bmax, max_nodes, nodes, &cnt, &depth, NULL, NULL))
    return -1;
  orc_bh_forces(n, nodes, cnt, x, y, z, m, theta, G, eps, NULL, 0, ax, ay, az, nthreads);
  return 0;
}

/* ------------------------------------------------------------ spatial hash */

/* force_spatial_hash.cu:15-18. */
void orc_grid_cell_index(float x, float y, float z, float cell, int32_t out[3]) {
  out[0] = (int32_t)floorf(x / cell);
  out[1] = (int32_t)floorf(y / cell);
  out[2] = (int32_t)floorf(z / cell);
}

/* force_spatial_hash.cu:20-26: wrap negative indices, then x fastest. */
int32_t orc_grid_hash_cell(const int32_t c[3], const int32_t d[3]) {
  const int32_t cx = ((c[0] % d[0]) + d[0]) % d[0];
  const int32_t cy = ((c[1] % d[1]) + d[1]) % d[1];
  const int32_t cz = ((c[2] % d[2]) + d[2]) % d[2];
  return cx + cy * d[0] + cz * d[0] * d[1];
}

/* force_spatial_hash.cu:190-260: the Barnes-Hut bbox reduction (:185-188), the host
 * padding (:232-238), dims (:250-252) and the 1e8-cell limit (:254-260). */
int orc_grid_dims(size_t n, const float* x, const float* y, const float* z, float cell,
                  float bbox[6], int32_t dims[3], int64_t* total) {
  float lo[3], hi[3];
  orc_bbox(n, x, y, z, lo, hi);
  for (int a = 0; a < 3; ++a) {
    lo[a] -= 0.001f;
    hi[a] += 0.001f;
    bbox[a] = lo[a];
    bbox[3 + a] = hi[a];
  }
  int64_t t = 1;
  int too_large = 0;
  for (int a = 0; a < 3; ++a) {
    const float r = (hi[a] - lo[a]) / cell;
    if (!(r <= 1e8f)) { /* the int cast below would overflow; the product is > 1e8 anyway */
      dims[a] = 0;
      too_large = 1;
      continue;
    }
    dims[a] = (int32_t)ceilf(r) + 1;
    if (!too_large) {
      t *= dims[a];
      if (t > 100000000) too_large = 1;
    }
  }
  *total = too_large ? -1 : t;
  return too_large ? -1 : 0;
}

static int32_t clamp_i32(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Cell coordinates of a position: :36-49 (assignCellsKernel) and :103-108 (the
 * force kernel repeats the same arithmetic and clamp). */
static void grid_coords(float px, float py, float pz, const float bbox[6], const int32_t dims[3],
                        float cell, int32_t c[3]) {
  c[0] = clamp_i32((int32_t)floorf((px - bbox[0]) / cell), 0, dims[0] - 1);
  c[1] = clamp_i32((int32_t)floorf((py - bbox[1]) / cell), 0, dims[1] - 1);
  c[2] = clamp_i32((int32_t)floorf((pz - bbox[2]) / cell), 0, dims[2] - 1);
}

/* r^2 of a pair as the reference writes it, dx*dx + dy*dy + dz*dz (:137), in the
 * fused form an LLVM-based CUDA compiler gives it under its default contraction:
 * fma(dz, dz, fma(dx, dx, dy*dy)). The spatial-hash kernel evaluates the same form,
 * so both keep exactly the same pairs. */
static float grid_d2(float rx, float ry, float rz) {
  return fmaf(rz, rz, fmaf(rx, rx, ry * ry));
}

/* Sub-cell coordinate of a position in its (clamped) cell c on one axis, ns sub-cells
 * per axis (16 or 4): floor(ns u) - ns c clamped to 0..ns-1 with u = (p - lo) / cell
 * as for the cell (ns u is exact). */
static int32_t grid_sub(float p, float lo, float cell, int32_t c, int32_t ns) {
  const float u = (p - lo) / cell;
  const int32_t s = (int32_t)floorf(u * (float)ns) - ns * c;
  return s < 0 ? 0 : (s > ns - 1 ? ns - 1 : s);
}

/* bits 3..0 of v to bits 9, 6, 3, 0 */
static uint32_t spread4(int32_t v) {
  return (uint32_t)((v & 1) | ((v & 2) << 2) | ((v & 4) << 4) | ((v & 8) << 6));
}

/* force_spatial_hash.cu:278-308. Cell ranges: the exclusive scan of the counts.
 * Order within a cell (open in the reference, whose scatter is atomic): the Morton
 * index of the position's sub-cell, then the particle index. Sub-cells are
 * 16 x 16 x 16 per cell for grids of at most 2^20 cells, 4 x 4 x 4 up to 2^26;
 * larger grids keep the particle index alone. */
void orc_grid_build(size_t n, const float* x, const float* y, const float* z, float cell,
                    const float bbox[6], const int32_t dims[3], int32_t* particle_cell,
                    int32_t* cell_start, int32_t* cell_end, int32_t* sorted) {
  const size_t total = (size_t)dims[0] * (size_t)dims[1] * (size_t)dims[2];
  const int32_t ns = total <= ((size_t)1 << 20) ? 16 : (total <= ((size_t)1 << 26) ? 4 : 0);
  const int shift = ns == 16 ? 12 : 6;
  int32_t* count = (int32_t*)calloc(total ? total : 1, sizeof(int32_t));
  uint32_t* key = (uint32_t*)malloc((n ? n : 1) * sizeof(uint32_t));
  for (size_t i = 0; i < n; ++i) {
    int32_t c[3];
    grid_coords(x[i], y[i], z[i], bbox, dims, cell, c);
    const int32_t id = c[0] + c[1] * dims[0] + c[2] * dims[0] * dims[1];
    particle_cell[i] = id;
    count[id] += 1;
    uint32_t k = (uint32_t)id;
    if (ns) {
      const uint32_t sub = (spread4(grid_sub(x[i], bbox[0], cell, c[0], ns)) << 2) |
                           (spread4(grid_sub(y[i], bbox[1], cell, c[1], ns)) << 1) |
                           spread4(grid_sub(z[i], bbox[2], cell, c[2], ns));
      k = (k << shift) | sub;
    }
    key[i] = k;
    sorted[i] = (int32_t)i;
  }
  int32_t run = 0;
  for (size_t c = 0; c < total; ++c) {
    cell_start[c] = run;
    run += count[c];
    cell_end[c] = run;
  }
  orc_stable_sort(n, key, sorted);
  free(key);
  free(count);
}

void orc_grid_forces_f32(size_t n, const float* x, const float* y, const float* z,
                         const float* m, const float bbox[6], const int32_t dims[3], float cell,
                         const int32_t* cell_start, const int32_t* cell_end,
                         const int32_t* sorted, float cutoff, float G, float eps,
                         const int64_t* targets, size_t ntargets, float* ax, float* ay,
                         float* az, int nthreads) {
  const size_t nt = targets ? ntargets : n;
  const float cutoff2 = cutoff * cutoff; /* :320 */
  const float eps2 = eps * eps;
  int th = pick_threads(nthreads);
  (void)th;
#pragma omp parallel for schedule(dynamic, 64) num_threads(th)
  for (long long t = 0; t < (long long)nt; ++t) {
    const int32_t i = targets ? (int32_t)targets[t] : (int32_t)t;
    const float xi = x[i], yi = y[i], zi = z[i];
    int32_t c[3];
    grid_coords(xi, yi, zi, bbox, dims, cell, c);
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int32_t nx = c[0] + dx, ny = c[1] + dy, nz = c[2] + dz;
          if (nx < 0 || nx >= dims[0] || ny < 0 || ny >= dims[1] || nz < 0 || nz >= dims[2])
            continue;
          const int32_t id = nx + ny * dims[0] + nz * dims[0] * dims[1];
          for (int32_t k = cell_start[id]; k < cell_end[id]; ++k) {
            const int32_t j = sorted[k];
            if (j == i) continue;
            const float rx = x[j] - xi, ry = y[j] - yi, rz = z[j] - zi;
            float d2 = grid_d2(rx, ry, rz);
            if (d2 < cutoff2) {
              d2 += eps2;
              const float inv = 1.0f / sqrtf(d2);
              const float inv3 = inv * inv * inv;
              const float f = G * m[j] * inv3;
              sx += f * rx;
              sy += f * ry;
              sz += f * rz;
            }
          }
        }
    ax[t] = sx;
    ay[t] = sy;
    az[t] = sz;
  }
}

/* FP64 contribution of one kept pair (shared by the grid and brute-force sums). */
static void add_f64(double G, double eps2, float mj, double rx, double ry, double rz, double* sx,
                    double* sy, double* sz, double* smag) {
  const double r2 = rx * rx + ry * ry + rz * rz;
  const double d2 = r2 + eps2;
  const double inv3 = 1.0 / (d2 * sqrt(d2));
  const double gm = G * (double)mj;
  *sx += gm * inv3 * rx;
  *sy += gm * inv3 * ry;
  *sz += gm * inv3 * rz;
  *smag += fabs(gm) * inv3 * sqrt(r2);
}

static int keep_pair(float xi, float yi, float zi, float xj, float yj, float zj, float cutoff2) {
  return grid_d2(xj - xi, yj - yi, zj - zi) < cutoff2;
}

void orc_grid_forces_f64(size_t n, const float* x, const float* y, const float* z,
                         const float* m, const float bbox[6], const int32_t dims[3], float cell,
                         const int32_t* cell_start, const int32_t* cell_end,
                         const int32_t* sorted, float cutoff, double G, double eps2,
                         const int64_t* targets, size_t ntargets, double* ax, double* ay,
                         double* az, double* mag, int64_t* npairs, int nthreads) {
  const size_t nt = targets ? ntargets : n;
  const float cutoff2 = cutoff * cutoff;
  int th = pick_threads(nthreads);
  (void)th;
#pragma omp parallel for schedule(dynamic, 64) num_threads(th)
  for (long long t = 0; t < (long long)nt; ++t) {
    const int32_t i = targets ? (int32_t)targets[t] : (int32_t)t;
    const float xi = x[i], yi = y[i], zi = z[i];
    int32_t c[3];
    grid_coords(xi, yi, zi, bbox, dims, cell, c);
    double sx = 0.0, sy = 0.0, sz = 0.0, smag = 0.0;
    int64_t np = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int32_t nx = c[0] + dx, ny = c[1] + dy, nz = c[2] + dz;
          if (nx < 0 || nx >= dims[0] || ny < 0 || ny >= dims[1] || nz < 0 || nz >= dims[2])
            continue;
          const int32_t id = nx + ny * dims[0] + nz * dims[0] * dims[1];
          for (int32_t k = cell_start[id]; k < cell_end[id]; ++k) {
            const int32_t j = sorted[k];
            if (j == i || !keep_pair(xi, yi, zi, x[j], y[j], z[j], cutoff2)) continue;
            add_f64(G, eps2, m[j], (double)x[j] - xi, (double)y[j] - yi, (double)z[j] - zi, &sx,
                    &sy, &sz, &smag);
            ++np;
          }
        }
    ax[t] = sx;
    ay[t] = sy;
    az[t] = sz;
    if (mag) mag[t] = smag;
    if (npairs) npairs[t] = np;
  }
}

void orc_cutoff_forces_f64(size_t n, const float* x, const float* y, const float* z,
                           const float* m, float cutoff, double G, double eps2,
                           const int64_t* targets, size_t ntargets, double* ax, double* ay,
                           double* az, int64_t* npairs, int nthreads) {
  const size_t nt = targets ? ntargets : n;
  const float cutoff2 = cutoff * cutoff;
  int th = pick_threads(nthreads);
  (void)th;
#pragma omp parallel for schedule(dynamic, 16) num_threads(th)
  for (long long t = 0; t < (long long)nt; ++t) {
    const size_t i = targets ? (size_t)targets[t] : (size_t)t;
    const float xi = x[i], yi = y[i], zi = z[i];
    double sx = 0.0, sy = 0.0, sz = 0.0, smag = 0.0;
    int64_t np = 0;
    for (size_t j = 0; j < n; ++j) {
      if (j == i || !keep_pair(xi, yi, zi, x[j], y[j], z[j], cutoff2)) continue;
      add_f64(G, eps2, m[j], (double)x[j] - xi, (double)y[j] - yi, (double)z[j] - zi, &sx, &sy,
              &sz, &smag);
      ++np;
    }
    ax[t] = sx;
    ay[t] = sy;
    az[t] = sz;
    if (npairs) npairs[t] = np;
  }
}
