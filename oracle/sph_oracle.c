/*
 * FP64 CPU oracle for the position-based-fluids substep -- TEST INFRASTRUCTURE.
 * Only tests/, the smoke entry point and bench.py's cpu_baseline leg use it.
 *
 * Restates AdamSimpson/SPHPC Simulation/src/particles.c, neighbors.c, simulation.c
 * and obstacle_cuda.cu in plain C.  Expressions keep the reference's operand
 * order (left-to-right C evaluation) so that, compiled with -ffp-contract=off,
 * results are bit-identical to the reference's own C code (checked against
 * oracle/_ref by tests/test_oracle_vs_ref.py).  OpenMP replaces each
 * `#pragma acc parallel loop`.
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include "sph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------------- */
/* Parameters: SetParameters (simulation.c:152-177) and AllocInitNeighbors  */
/* (neighbors.c:23-28).                                                     */
/* ---------------------------------------------------------------------- */
void orc_params_derive(orc_params *p) {
  const double h = p->h;
  p->W_norm = 315.0 / (64.0 * M_PI * pow(h, 9.0));      /* simulation.c:175 */
  p->DelW_poly_norm = -3.0 * 2.0 * p->W_norm;            /* simulation.c:176 */
  p->DelW_norm = -45.0 / (M_PI * pow(h, 6.0));           /* simulation.c:177 */
  p->nx = (int)(ceil((p->box_max[0] - p->box_min[0]) / h) + 1); /* neighbors.c:23 */
  p->ny = (int)(ceil((p->box_max[1] - p->box_min[1]) / h) + 1); /* neighbors.c:25 */
  p->nz = (int)(ceil((p->box_max[2] - p->box_min[2]) / h) + 1); /* neighbors.c:27 */
  if (p->max_neighbors <= 0) p->max_neighbors = 60;      /* neighbors.h:6 */
}

/* ---------------------------------------------------------------------- */
/* Smoothing kernels, particles.c:18-65                                      */
/* ---------------------------------------------------------------------- */
double orc_W(double r, double h, double norm) {          /* particles.c:20-26 */
  if (r > h) return 0.0;
  return norm * (h * h - r * r) * (h * h - r * r) * (h * h - r * r);
}

double orc_DelWPoly(double r, double h, double norm) {   /* particles.c:29-35 */
  if (r > h) return 0.0;
  return norm * (h * h - r * r) * (h * h - r * r) * r;
}

double orc_DelW(double r, double h, double norm) {       /* particles.c:40-46 */
  if (r > h) return 0.0;
  return norm * (h - r) * (h - r);
}

double orc_C(double r, double h, double norm) {          /* particles.c:54-65 */
  double c;
  if (r > h)
    c = 0.0;
  else if (r <= h / 2.0)
    c = (2.0 * (h - r) * (h - r) * (h - r) * r * r * r) - (h * h * h * h * h * h / 64.0);
  else
    c = (h - r) * (h - r) * (h - r) * r * r * r;
  return norm * c;
}

/* ---------------------------------------------------------------------- */
/* Neighbour search                                                          */
/* ---------------------------------------------------------------------- */
uint32_t orc_hash_val(const orc_params *p, double x, double y, double z) {
  /* HashVal, neighbors.c:79-98: absolute coordinates, spacing = h. */
  const double spacing = p->h;
  const unsigned int gx = (unsigned int)floor(x / spacing);
  const unsigned int gy = (unsigned int)floor(y / spacing);
  const unsigned int gz = (unsigned int)floor(z / spacing);
  const unsigned int nx = (unsigned int)p->nx;
  const unsigned int ny = (unsigned int)p->ny;
  return (nx * ny * gz) + (gy * nx + gx);
}

void orc_hash_particles(int n, const double *xs, const double *ys, const double *zs,
                        const orc_params *p, uint32_t *keys, uint32_t *ids) {
  /* HashParticles, neighbors.c:101-125 */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    keys[i] = orc_hash_val(p, xs[i], ys[i], zs[i]);
    ids[i] = (uint32_t)i;
  }
}

/* Stable LSD radix sort of (key, value) by key: the semantics of Thrust's
 * sort_by_key on unsigned keys (thrust_c.cu:86-91), restated as a textbook
 * counting-sort-per-byte.  Equal keys keep their input order. */
void orc_sort_by_key(int n, uint32_t *keys, uint32_t *vals) {
  if (n <= 1) return;
  uint32_t *tk = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
  uint32_t *tv = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
  uint32_t *ka = keys, *va = vals, *kb = tk, *vb = tv;
  for (int shift = 0; shift < 32; shift += 8) {
    size_t count[257];
    memset(count, 0, sizeof(count));
    for (int i = 0; i < n; ++i) count[((ka[i] >> shift) & 255u) + 1]++;
    for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
    for (int i = 0; i < n; ++i) {
      const size_t dst = count[(ka[i] >> shift) & 255u]++;
      kb[dst] = ka[i];
      vb[dst] = va[i];
    }
    uint32_t *t = ka; ka = kb; kb = t;
    t = va; va = vb; vb = t;
  }
  /* four passes: result is back in keys/vals */
  free(tk);
  free(tv);
}

/* lower_bound / upper_bound of every value 0..cells-1 (thrust_c.cu:93-120). */
void orc_find_lower_bounds(const uint32_t *keys, int n, int cells, uint32_t *out) {
#pragma omp parallel for schedule(static)
  for (int c = 0; c < cells; ++c) {
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if (keys[mid] < (uint32_t)c) lo = mid + 1; else hi = mid;
    }
    out[c] = (uint32_t)lo;
  }
}

void orc_find_upper_bounds(const uint32_t *keys, int n, int cells, uint32_t *out) {
#pragma omp parallel for schedule(static)
  for (int c = 0; c < cells; ++c) {
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if (keys[mid] <= (uint32_t)c) lo = mid + 1; else hi = mid;
    }
    out[c] = (uint32_t)lo;
  }
}

long orc_fill_neighbors(int n, const double *xs, const double *ys, const double *zs,
                        const orc_params *p, const uint32_t *start, const uint32_t *end,
                        const uint32_t *ids, uint32_t *nbr, int *cnt) {
  /* FillNeighbors, neighbors.c:217-298 */
  const double h2 = p->h * p->h;
  const double spacing = p->h;
  const int maxn = p->max_neighbors;
  long dropped = 0;
#pragma omp parallel for schedule(static) reduction(+ : dropped)
  for (int i = 0; i < n; ++i) {
    uint32_t *bucket = nbr + (size_t)i * (size_t)maxn;
    int count = 0;
    const double xp = xs[i], yp = ys[i], zp = zs[i];
    for (int dz = -1; dz <= 1; ++dz) {
      for (int dy = -1; dy <= 1; ++dy) {
        for (int dx = -1; dx <= 1; ++dx) {
          const double x = xp + dx * spacing;
          const double y = yp + dy * spacing;
          const double z = zp + dz * spacing;
          if (floor(x / spacing) > p->nx - 1 || x < 0.0 ||
              floor(y / spacing) > p->ny - 1 || y < 0.0 ||
              floor(z / spacing) > p->nz - 1 || z < 0.0)
            continue;
          const int cell = (int)orc_hash_val(p, x, y, z);
          const uint32_t s = start[cell], e = end[cell];
          if (s == e) continue;
          for (uint32_t j = s; j < e; ++j) {
            const int q = (int)ids[j];
            if (q == i) continue;
            const double xd = xp - xs[q];
            const double yd = yp - ys[q];
            const double zd = zp - zs[q];
            const double r2 = xd * xd + yd * yd + zd * zd;
            if (r2 < h2) {
              if (count < maxn) bucket[count++] = (uint32_t)q;
              else dropped++;
            }
          }
        }
      }
    }
    cnt[i] = count;
  }
  return dropped;
}

long orc_find_all_neighbors(int n, const double *xs, const double *ys, const double *zs,
                            const orc_params *p, uint32_t *nbr, int *cnt) {
  /* FindAllNeighbors, neighbors.c:68-75 */
  const int cells = p->nx * p->ny * p->nz;
  uint32_t *keys = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  uint32_t *ids = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
  uint32_t *start = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)cells);
  uint32_t *end = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)cells);
  orc_hash_particles(n, xs, ys, zs, p, keys, ids);
  orc_sort_by_key(n, keys, ids);
  orc_find_lower_bounds(keys, n, cells, start);
  orc_find_upper_bounds(keys, n, cells, end);
  const long dropped = orc_fill_neighbors(n, xs, ys, zs, p, start, end, ids, nbr, cnt);
  free(keys); free(ids); free(start); free(end);
  return dropped;
}

/* ---------------------------------------------------------------------- */
/* Elementwise stages                                                        */
/* ---------------------------------------------------------------------- */
void orc_apply_gravity(int n, double *vz, const orc_params *p) {
  /* ApplyGravity, particles.c:516-527 */
  const double dt = p->dt;
  const double g = -p->g;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) vz[i] += g * dt;
}

void orc_apply_boundary(double *x, double *y, double *z, const orc_params *p) {
  /* ApplyBoundaryConditions, particles.c:842-867 */
  const double eps = p->h * 0.0001;
  if (*x < p->box_min[0]) *x = p->box_min[0];
  else if (*x >= p->box_max[0]) *x = p->box_max[0] - eps;
  if (*y < p->box_min[1]) *y = p->box_min[1];
  else if (*y >= p->box_max[1]) *y = p->box_max[1] - eps;
  if (*z < p->box_min[2]) *z = p->box_min[2];
  else if (*z >= p->box_max[2]) *z = p->box_max[2] - eps;
}

void orc_predict_positions(int n, const double *x, const double *y, const double *z,
                           const double *vx, const double *vy, const double *vz,
                           double *xs, double *ys, double *zs, const orc_params *p) {
  /* PredictPositions, particles.c:748-784 */
  const double dt = p->dt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    xs[i] = x[i] + (vx[i] * dt);
    ys[i] = y[i] + (vy[i] * dt);
    zs[i] = z[i] + (vz[i] * dt);
    orc_apply_boundary(&xs[i], &ys[i], &zs[i], p);
  }
}

/* ---------------------------------------------------------------------- */
/* Jacobi constraint loop                                                    */
/* ---------------------------------------------------------------------- */
void orc_compute_densities(int n, const double *xs, const double *ys, const double *zs,
                           const uint32_t *nbr, const int *cnt, const orc_params *p,
                           double *density) {
  /* ComputeDensities, particles.c:461-514 (no r floor) */
  const double h = p->h;
  const double m = p->rest_mass;
  const double W_0 = orc_W(0.0, h, p->W_norm);
  const int maxn = p->max_neighbors;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const uint32_t *b = nbr + (size_t)i * (size_t)maxn;
    double rho = m * W_0;
    const double xp = xs[i], yp = ys[i], zp = zs[i];
    for (int j = 0; j < cnt[i]; ++j) {
      const int q = (int)b[j];
      const double xd = xp - xs[q];
      const double yd = yp - ys[q];
      const double zd = zp - zs[q];
      const double r = sqrt(xd * xd + yd * yd + zd * zd);
      rho += m * orc_W(r, h, p->W_norm);
    }
    density[i] = rho;
  }
}

void orc_compute_lambda(int n, const double *xs, const double *ys, const double *zs,
                        const double *density, const uint32_t *nbr, const int *cnt,
                        const orc_params *p, double *lambda) {
  /* ComputeLambda, particles.c:588-666 */
  const double rho0 = p->rest_density;
  const double m = p->rest_mass;
  const double h = p->h;
  const int maxn = p->max_neighbors;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const uint32_t *b = nbr + (size_t)i * (size_t)maxn;
    const double constraint = density[i] / rho0 - 1.0;
    const double Ci = (constraint < 0.0 ? 0.0 : constraint);
    double sum_C = 0.0, sgx = 0.0, sgy = 0.0, sgz = 0.0;
    const double xp = xs[i], yp = ys[i], zp = zs[i];
    for (int j = 0; j < cnt[i]; ++j) {
      const int q = (int)b[j];
      const double xd = xp - xs[q];
      const double yd = yp - ys[q];
      const double zd = zp - zs[q];
      double r = sqrt(xd * xd + yd * yd + zd * zd);
      if (r < h * 0.0001) r = h * 0.0001;
      const double grad = orc_DelW(r, h, p->DelW_norm);
      const double gx = m * grad * xd / r;
      const double gy = m * grad * yd / r;
      const double gz = m * grad * zd / r;
      sgx += gx;
      sgy += gy;
      sgz += gz;
      sum_C += (gx * gx + gy * gy + gz * gz);
    }
    sum_C += (sgx * sgx + sgy * sgy + sgz * sgz);
    sum_C /= (rho0 * rho0);
    lambda[i] = -Ci / (sum_C + p->lambda_eps);
  }
}

double orc_stiffness(const orc_params *p, int substep) {
  /* particles.c:693-694; substep 0 gives pow(0.05, inf) = 0, stiffness 1 */
  return 1.0 - pow((1.0 - p->k_stiff), 1.0 / (double)substep);
}

void orc_update_dps(int n, const double *xs, const double *ys, const double *zs,
                    const double *lambda, const uint32_t *nbr, const int *cnt,
                    const orc_params *p, int substep, double *dpx, double *dpy, double *dpz) {
  /* UpdateDPs, particles.c:668-746 (s_corr disabled, :730) */
  const double h = p->h;
  const double rho0 = p->rest_density;
  const double m = p->rest_mass;
  const double stiffness = orc_stiffness(p, substep);
  const int maxn = p->max_neighbors;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const uint32_t *b = nbr + (size_t)i * (size_t)maxn;
    double ax = 0.0, ay = 0.0, az = 0.0;
    const double xp = xs[i], yp = ys[i], zp = zs[i];
    for (int j = 0; j < cnt[i]; ++j) {
      const int q = (int)b[j];
      const double xd = xp - xs[q];
      const double yd = yp - ys[q];
      const double zd = zp - zs[q];
      double r = sqrt(xd * xd + yd * yd + zd * zd);
      if (r < h * 0.0001) r = h * 0.0001;
      const double s_corr = 0.0;
      const double delW = orc_DelW(r, h, p->DelW_norm);
      const double dp = m * (lambda[i] + lambda[q] + s_corr) * delW;
      ax += dp * xd / r;
      ay += dp * yd / r;
      az += dp * zd / r;
    }
    dpx[i] = stiffness * ax / rho0;
    dpy[i] = stiffness * ay / rho0;
    dpz[i] = stiffness * az / rho0;
  }
}

/* Trilinear sample with normalised coordinates, texel centres at (i+0.5)/dim
 * and clamp addressing: the documented filtering of tex3D with
 * cudaFilterModeLinear / cudaAddressModeClamp / normalizedCoords
 * (obstacle_cuda.cu:174-180), with exact weights. */
static float sdf_sample(const orc_obstacle *o, float u, float v, float w) {
  const int dims[3] = {o->dim[0], o->dim[1], o->dim[2]};
  const double c[3] = {(double)u * dims[0] - 0.5, (double)v * dims[1] - 0.5,
                       (double)w * dims[2] - 0.5};
  int i0[3], i1[3];
  double a[3];
  for (int k = 0; k < 3; ++k) {
    const double f = floor(c[k]);
    a[k] = c[k] - f;
    int lo = (int)f, hi = (int)f + 1;
    if (lo < 0) lo = 0;
    if (lo > dims[k] - 1) lo = dims[k] - 1;
    if (hi < 0) hi = 0;
    if (hi > dims[k] - 1) hi = dims[k] - 1;
    i0[k] = lo;
    i1[k] = hi;
  }
  const size_t sx = 1, sy = (size_t)dims[0], sz = (size_t)dims[0] * (size_t)dims[1];
#define T(ix, iy, iz) ((double)o->sdf[(size_t)(ix) * sx + (size_t)(iy) * sy + (size_t)(iz) * sz])
  const double c00 = T(i0[0], i0[1], i0[2]) * (1.0 - a[0]) + T(i1[0], i0[1], i0[2]) * a[0];
  const double c10 = T(i0[0], i1[1], i0[2]) * (1.0 - a[0]) + T(i1[0], i1[1], i0[2]) * a[0];
  const double c01 = T(i0[0], i0[1], i1[2]) * (1.0 - a[0]) + T(i1[0], i0[1], i1[2]) * a[0];
  const double c11 = T(i0[0], i1[1], i1[2]) * (1.0 - a[0]) + T(i1[0], i1[1], i1[2]) * a[0];
#undef T
  const double c0 = c00 * (1.0 - a[1]) + c10 * a[1];
  const double c1 = c01 * (1.0 - a[1]) + c11 * a[1];
  return (float)(c0 * (1.0 - a[2]) + c1 * a[2]);
}

void orc_collide(int n, double *xs, double *ys, double *zs, const orc_obstacle *o) {
  /* CalculateParticleCollisions_kernel, obstacle_cuda.cu:56-116 */
  if (!o || !o->sdf) return;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const double xw = xs[i], yw = ys[i], zw = zs[i];
    /* InsideObstacleBounds, obstacle_cuda.cu:20-31 (strict) */
    if (!(xw > o->wmin[0] && xw < o->wmax[0] && yw > o->wmin[1] && yw < o->wmax[1] &&
          zw > o->wmin[2] && zw < o->wmax[2]))
      continue;
    const float len_x = (float)(o->wmax[0] - o->wmin[0]);
    const float len_y = (float)(o->wmax[1] - o->wmin[1]);
    const float len_z = (float)(o->wmax[2] - o->wmin[2]);
    const float u = (float)((xw - o->wmin[0]) / len_x);
    const float v = (float)((yw - o->wmin[1]) / len_y);
    const float w = (float)((zw - o->wmin[2]) / len_z);
    const float d = sdf_sample(o, u, v, w);
    if (!(d < 0)) continue;
    const float sx = 1.0f / (float)o->dim[0];
    const float sy = 1.0f / (float)o->dim[1];
    const float sz = 1.0f / (float)o->dim[2];
    /* CalculateNormal, obstacle_cuda.cu:33-53 */
    const float gx = sdf_sample(o, u + sx, v, w) - sdf_sample(o, u - sx, v, w);
    const float gy = sdf_sample(o, u, v + sy, w) - sdf_sample(o, u, v - sy, w);
    const float gz = sdf_sample(o, u, v, w + sz) - sdf_sample(o, u, v, w - sz);
    const float dd = gx * gx + gy * gy + gz * gz;
    if (!(dd > 0.0f)) continue;  /* deviation: reference writes NaN */
    const float inv = 1.0f / sqrtf(dd);
    const float len_model = (float)(o->dim[0] * o->spacing);
    const float m2w = len_x / len_model;
    const float dist = fabsf(d) * m2w;
    ys[i] += dist * (gy * inv);
    xs[i] += dist * (gx * inv);
    zs[i] += dist * (gz * inv);
  }
}

void orc_update_position_stars(int n, double *xs, double *ys, double *zs,
                               const double *dpx, const double *dpy, const double *dpz,
                               const orc_params *p, const orc_obstacle *o) {
  /* UpdatePositionStars, particles.c:529-564 */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    xs[i] += dpx[i];
    ys[i] += dpy[i];
    zs[i] += dpz[i];
    orc_apply_boundary(&xs[i], &ys[i], &zs[i], p);
  }
  orc_collide(n, xs, ys, zs, o);
}

/* ---------------------------------------------------------------------- */
/* Velocity stages                                                           */
/* ---------------------------------------------------------------------- */
void orc_update_velocities(int n, const double *xs, const double *ys, const double *zs,
                           const double *x, const double *y, const double *z,
                           double *vx, double *vy, double *vz, const orc_params *p) {
  /* UpdateVelocities, particles.c:787-838 */
  const double dt = p->dt;
  const double vmax = p->v_max;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double ux = (xs[i] - x[i]) / dt;
    double uy = (ys[i] - y[i]) / dt;
    double uz = (zs[i] - z[i]) / dt;
    if (ux > vmax) ux = vmax; else if (ux < -vmax) ux = -vmax;
    if (uy > vmax) uy = vmax; else if (uy < -vmax) uy = -vmax;
    if (uz > vmax) uz = vmax; else if (uz < -vmax) uz = -vmax;
    vx[i] = ux;
    vy[i] = uy;
    vz[i] = uz;
  }
}

void orc_compute_color(int n, const double *xs, const double *ys, const double *zs,
                       const double *density, const uint32_t *nbr, const int *cnt,
                       const orc_params *p, double *cx, double *cy, double *cz) {
  /* ComputeSurfaceTension colour pass, particles.c:94-140 */
  const double h = p->h;
  const double m = p->rest_mass;
  const int maxn = p->max_neighbors;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const uint32_t *b = nbr + (size_t)i * (size_t)maxn;
    double ax = 0.0, ay = 0.0, az = 0.0;
    const double xp = xs[i], yp = ys[i], zp = zs[i];
    for (int j = 0; j < cnt[i]; ++j) {
      const int q = (int)b[j];
      const double xd = xp - xs[q];
      const double yd = yp - ys[q];
      const double zd = zp - zs[q];
      double r = sqrt(xd * xd + yd * yd + zd * zd);
      if (r < h * 0.0001) r = h * 0.0001;
      const double grad = orc_DelWPoly(r, h, p->DelW_poly_norm);
      ax += grad * m / density[q] * xd / r;
      ay += grad * m / density[q] * yd / r;
      az += grad * m / density[q] * zd / r;
    }
    cx[i] = h * ax;
    cy[i] = h * ay;
    cz[i] = h * az;
  }
}

void orc_apply_surface_force(int n, const double *xs, const double *ys, const double *zs,
                             const double *density, const double *cx, const double *cy,
                             const double *cz, const uint32_t *nbr, const int *cnt,
                             const orc_params *p, double *vx, double *vy, double *vz) {
  /* ComputeSurfaceTension force pass, particles.c:142-204 */
  const double rho0 = p->rest_density;
  const double m = p->rest_mass;
  const double h = p->h;
  const double dt = p->dt;
  const double C_norm = 32.0 / (M_PI * pow(h, 9.0));    /* particles.c:78 */
  const double gama = p->k;
  const int maxn = p->max_neighbors;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const uint32_t *b = nbr + (size_t)i * (size_t)maxn;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const double xp = xs[i], yp = ys[i], zp = zs[i];
    const double rho_p = density[i];
    const double cxp = cx[i], cyp = cy[i], czp = cz[i];
    const double mass_p = m;
    for (int j = 0; j < cnt[i]; ++j) {
      const int q = (int)b[j];
      const double xd = xp - xs[q];
      const double yd = yp - ys[q];
      const double zd = zp - zs[q];
      double r = sqrt(xd * xd + yd * yd + zd * zd);
      if (r < h * 0.0001) r = h * 0.0001;
      const double mass_q = m;
      const double c = orc_C(r, h, C_norm);
      const double fcx = -20.0 * gama * mass_p * mass_q * c * xd / r;
      const double fcy = -20.0 * gama * mass_p * mass_q * c * yd / r;
      const double fcz = -20.0 * gama * mass_p * mass_q * c * zd / r;
      const double fkx = -gama * mass_p * (cxp - cx[q]);
      const double fky = -gama * mass_p * (cyp - cy[q]);
      const double fkz = -gama * mass_p * (czp - cz[q]);
      const double K = 2.0 * rho0 / (rho_p + density[q]);
      fx += K * (fcx + fkx);
      fy += K * (fcy + fky);
      fz += K * (fcz + fkz);
    }
    vx[i] += dt * fx / rho_p;
    vy[i] += dt * fy / rho_p;
    vz[i] += dt * fz / rho_p;
  }
}

void orc_apply_viscosity(int n, const double *xs, const double *ys, const double *zs,
                         const double *density, const double *vx, const double *vy,
                         const double *vz, const uint32_t *nbr, const int *cnt,
                         const orc_params *p, double *ovx, double *ovy, double *ovz) {
  /* ApplyViscosity, particles.c:383-459, with the distance typo of :439 kept
   * (y_diff*y_diff * z_diff*z_diff).  Jacobi: reads v, writes ov. */
  const double c = p->c;
  const double h = p->h;
  const double m = p->rest_mass;
  const int maxn = p->max_neighbors;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const uint32_t *b = nbr + (size_t)i * (size_t)maxn;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    const double xp = xs[i], yp = ys[i], zp = zs[i];
    const double vxp = vx[i], vyp = vy[i], vzp = vz[i];
    for (int j = 0; j < cnt[i]; ++j) {
      const int q = (int)b[j];
      const double xd = xp - xs[q];
      const double yd = yp - ys[q];
      const double zd = zp - zs[q];
      const double dvx = vx[q] - vxp;
      const double dvy = vy[q] - vyp;
      const double dvz = vz[q] - vzp;
      const double r = sqrt(xd * xd + yd * yd * zd * zd);
      const double w = orc_W(r, h, p->W_norm);
      sx += dvx * w * m / density[q];
      sy += dvy * w * m / density[q];
      sz += dvz * w * m / density[q];
    }
    sx *= c;
    sy *= c;
    sz *= c;
    ovx[i] = vxp + sx;
    ovy[i] = vyp + sy;
    ovz[i] = vzp + sz;
  }
}

void orc_compute_vorticity(int n, const double *xs, const double *ys, const double *zs,
                           const double *density, const double *vx, const double *vy,
                           const double *vz, const uint32_t *nbr, const int *cnt,
                           const orc_params *p, double *wx, double *wy, double *wz) {
  /* ComputeVorticity, particles.c:207-283.  vort_model 1 weights each term by
   * m/rho_j (not in the reference). */
  const double h = p->h;
  const double m = p->rest_mass;
  const int maxn = p->max_neighbors;
  const int scaled = (p->vort_model == 1);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const uint32_t *b = nbr + (size_t)i * (size_t)maxn;
    double ox = 0.0, oy = 0.0, oz = 0.0;
    const double xp = xs[i], yp = ys[i], zp = zs[i];
    const double vxp = vx[i], vyp = vy[i], vzp = vz[i];
    for (int j = 0; j < cnt[i]; ++j) {
      const int q = (int)b[j];
      const double xd = xp - xs[q];
      const double yd = yp - ys[q];
      const double zd = zp - zs[q];
      const double dvx = vx[q] - vxp;
      const double dvy = vy[q] - vyp;
      const double dvz = vz[q] - vzp;
      double r = sqrt(xd * xd + yd * yd + zd * zd);
      if (r < h * 0.0001) r = h * 0.0001;
      double dw = orc_DelW(r, h, p->DelW_norm);
      if (scaled) dw = dw * (m / density[q]);
      const double dwx = dw * xd / r;
      const double dwy = dw * yd / r;
      const double dwz = dw * zd / r;
      ox += dvy * dwz - dvz * dwy;
      oy += dvz * dwx - dvx * dwz;
      oz += dvx * dwy - dvy * dwx;
    }
    wx[i] = ox;
    wy[i] = oy;
    wz[i] = oz;
  }
}

void orc_apply_vorticity_confinement(int n, const double *xs, const double *ys,
                                     const double *zs, const double *wx, const double *wy,
                                     const double *wz, const uint32_t *nbr, const int *cnt,
                                     const orc_params *p, double *vx, double *vy, double *vz) {
  /* ApplyVorticityConfinement, particles.c:285-381 (uses |omega_q|, :354) */
  const double dt = p->dt;
  const double eps = p->vort_eps;
  const double h = p->h;
  const double m = p->rest_mass;
  const int maxn = p->max_neighbors;
  const int scaled = (p->vort_model == 1);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const uint32_t *b = nbr + (size_t)i * (size_t)maxn;
    double ex = 0.0, ey = 0.0, ez = 0.0;
    const double xp = xs[i], yp = ys[i], zp = zs[i];
    for (int j = 0; j < cnt[i]; ++j) {
      const int q = (int)b[j];
      const double xd = xp - xs[q];
      const double yd = yp - ys[q];
      const double zd = zp - zs[q];
      double r = sqrt(xd * xd + yd * yd + zd * zd);
      if (r < h * 0.0001) r = h * 0.0001;
      const double dw = orc_DelW(r, h, p->DelW_norm);
      const double dwx = dw * xd / r;
      const double dwy = dw * yd / r;
      const double dwz = dw * zd / r;
      const double vm = sqrt(wx[q] * wx[q] + wy[q] * wy[q] + wz[q] * wz[q]);
      ex += vm * dwx;
      ey += vm * dwy;
      ez += vm * dwz;
    }
    double em = sqrt(ex * ex + ey * ey + ez * ez);
    if (em < h * 0.0001) em = h * 0.0001;
    const double Nx = ex / em, Ny = ey / em, Nz = ez / em;
    const double ox = wx[i], oy = wy[i], oz = wz[i];
    const double mass_p = scaled ? 1.0 : m;
    vx[i] += dt * eps * (Ny * oz - Nz * oy) / mass_p;
    vy[i] += dt * eps * (Nz * ox - Nx * oz) / mass_p;
    vz[i] += dt * eps * (Nx * oy - Ny * ox) / mass_p;
  }
}

void orc_update_positions(int n, const double *xs, const double *ys, const double *zs,
                          double *x, double *y, double *z) {
  /* UpdatePositions, particles.c:566-586 */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    x[i] = xs[i];
    y[i] = ys[i];
    z[i] = zs[i];
  }
}

double orc_max_density_error(int n, const double *density, double rest_density) {
  double mx = 0.0;
#pragma omp parallel for schedule(static) reduction(max : mx)
  for (int i = 0; i < n; ++i) {
    const double e = (density[i] - rest_density) / rest_density;
    if (e > mx) mx = e;
  }
  return mx;
}

/* ---------------------------------------------------------------------- */
/* Driver order, simulation.c:59-108                                         */
/* ---------------------------------------------------------------------- */
void orc_set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}

/* Rounds n values to the nearest float32 (the storage of an FP32 solver). */
static void round_f32(int n, double *a) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) a[i] = (double)(float)a[i];
}

/* The substep of simulation.c:59-108.  With f32 set, every array the stage
 * writes is rounded to float32 after the stage (velocity, x*, density,
 * lambda, delta p, the colour and vorticity fields): the same algorithm with
 * the storage of an FP32 solver, arithmetic inside each stage staying FP64. */
static int substep_impl(int n, orc_state *s, const orc_params *p, const orc_obstacle *o,
                        int iters, double tol, double *max_err_out, int f32) {
  const size_t N = (size_t)(n > 0 ? n : 1);
  double *dpx = (double *)malloc(sizeof(double) * N);
  double *dpy = (double *)malloc(sizeof(double) * N);
  double *dpz = (double *)malloc(sizeof(double) * N);
  double *t0 = (double *)malloc(sizeof(double) * N);
  double *t1 = (double *)malloc(sizeof(double) * N);
  double *t2 = (double *)malloc(sizeof(double) * N);
  double *u0 = (double *)malloc(sizeof(double) * N);
  double *u1 = (double *)malloc(sizeof(double) * N);
  double *u2 = (double *)malloc(sizeof(double) * N);

  orc_apply_gravity(n, s->vz, p);
  if (f32) round_f32(n, s->vz);
  orc_predict_positions(n, s->x, s->y, s->z, s->vx, s->vy, s->vz, s->xs, s->ys, s->zs, p);
  if (f32) { round_f32(n, s->xs); round_f32(n, s->ys); round_f32(n, s->zs); }
  orc_find_all_neighbors(n, s->xs, s->ys, s->zs, p, s->nbr, s->cnt);

  int done = 0;
  double err = 0.0;
  for (int it = 0; it <= iters; ++it) {
    if (it == iters && tol <= 0.0) break;
    orc_compute_densities(n, s->xs, s->ys, s->zs, s->nbr, s->cnt, p, s->density);
    if (f32) round_f32(n, s->density);
    if (tol > 0.0) {
      err = orc_max_density_error(n, s->density, p->rest_density);
      if (err <= tol || it == iters) break;
    }
    orc_compute_lambda(n, s->xs, s->ys, s->zs, s->density, s->nbr, s->cnt, p, s->lambda);
    if (f32) round_f32(n, s->lambda);
    orc_update_dps(n, s->xs, s->ys, s->zs, s->lambda, s->nbr, s->cnt, p, it, dpx, dpy, dpz);
    if (f32) { round_f32(n, dpx); round_f32(n, dpy); round_f32(n, dpz); }
    orc_update_position_stars(n, s->xs, s->ys, s->zs, dpx, dpy, dpz, p, o);
    if (f32) { round_f32(n, s->xs); round_f32(n, s->ys); round_f32(n, s->zs); }
    done++;
  }
  if (tol <= 0.0) err = orc_max_density_error(n, s->density, p->rest_density);
  if (max_err_out) *max_err_out = err;

  orc_update_velocities(n, s->xs, s->ys, s->zs, s->x, s->y, s->z, s->vx, s->vy, s->vz, p);
  if (f32) { round_f32(n, s->vx); round_f32(n, s->vy); round_f32(n, s->vz); }
  if (p->surface_tension_on) {
    orc_compute_color(n, s->xs, s->ys, s->zs, s->density, s->nbr, s->cnt, p, t0, t1, t2);
    if (f32) { round_f32(n, t0); round_f32(n, t1); round_f32(n, t2); }
    orc_apply_surface_force(n, s->xs, s->ys, s->zs, s->density, t0, t1, t2, s->nbr, s->cnt, p,
                            s->vx, s->vy, s->vz);
    if (f32) { round_f32(n, s->vx); round_f32(n, s->vy); round_f32(n, s->vz); }
  }
  if (p->viscosity_on) {
    orc_apply_viscosity(n, s->xs, s->ys, s->zs, s->density, s->vx, s->vy, s->vz, s->nbr,
                        s->cnt, p, u0, u1, u2);
    memcpy(s->vx, u0, sizeof(double) * (size_t)n);
    memcpy(s->vy, u1, sizeof(double) * (size_t)n);
    memcpy(s->vz, u2, sizeof(double) * (size_t)n);
    if (f32) { round_f32(n, s->vx); round_f32(n, s->vy); round_f32(n, s->vz); }
  }
  if (p->vorticity_on) {
    orc_compute_vorticity(n, s->xs, s->ys, s->zs, s->density, s->vx, s->vy, s->vz, s->nbr,
                          s->cnt, p, t0, t1, t2);
    if (f32) { round_f32(n, t0); round_f32(n, t1); round_f32(n, t2); }
    orc_apply_vorticity_confinement(n, s->xs, s->ys, s->zs, t0, t1, t2, s->nbr, s->cnt, p,
                                    s->vx, s->vy, s->vz);
    if (f32) { round_f32(n, s->vx); round_f32(n, s->vy); round_f32(n, s->vz); }
  }
  orc_update_positions(n, s->xs, s->ys, s->zs, s->x, s->y, s->z);

  free(dpx); free(dpy); free(dpz); free(t0); free(t1); free(t2);
  free(u0); free(u1); free(u2);
  return done;
}

int orc_substep(int n, orc_state *s, const orc_params *p, const orc_obstacle *o,
                int iters, double tol, double *max_err_out) {
  return substep_impl(n, s, p, o, iters, tol, max_err_out, 0);
}

int orc_substep_f32_storage(int n, orc_state *s, const orc_params *p, const orc_obstacle *o,
                            int iters, double tol, double *max_err_out) {
  return substep_impl(n, s, p, o, iters, tol, max_err_out, 1);
}
