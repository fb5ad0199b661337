/*
 * Scene and obstacle inputs (host side, C99).
 *
 *  - sph_scene_lattice: fluid blocks laid out like ConstructFluidVolume
 *    (geometry.c:42-55): x fastest, then y, then z; particle centres at
 *    origin + (i + 1/2) s.  An optional seeded jitter U(-j, j)*s per axis comes
 *    from a counter-based generator (splitmix64 of seed, particle, axis), so a
 *    scene is reproducible independent of thread count or block order.
 *  - sph_sdf_sample_shapes: union (min) of analytic signed distances sampled
 *    at the texel centres of a grid spanning a world box, the positions at
 *    which tex3D's linear filter reads texels (obstacle_cuda.cu:174-180).
 *  - SDFGen ASCII files (obstacle.c:13-66): "nx ny nz", "ox oy oz", "dx",
 *    then nx*ny*nz values, x fastest.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sphpc_amd.h"

static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

double sph_scene_uniform(uint64_t seed, uint64_t index, int axis) {
  const uint64_t z = splitmix64(seed ^ splitmix64(index * 3ull + (uint64_t)axis));
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

int sph_scene_lattice(const sph_block *blocks, int nblocks, double spacing, double jitter,
                      uint64_t seed, float *x, float *y, float *z, int cap, int *count) {
  if (!count || nblocks < 0 || (nblocks > 0 && !blocks) || !(spacing > 0.0) || jitter < 0.0)
    return SPH_ERR_ARG;
  long long total = 0;
  for (int b = 0; b < nblocks; ++b) {
    if (blocks[b].nx < 0 || blocks[b].ny < 0 || blocks[b].nz < 0) return SPH_ERR_ARG;
    total += (long long)blocks[b].nx * blocks[b].ny * blocks[b].nz;
  }
  if (total > 2147483647ll) return SPH_ERR_CAPACITY;
  *count = (int)total;
  if (!x || !y || !z) return SPH_OK; /* size query */
  if (total > cap) return SPH_ERR_CAPACITY;
  long long i = 0;
  for (int b = 0; b < nblocks; ++b) {
    const sph_block *B = &blocks[b];
    for (int kz = 0; kz < B->nz; ++kz) {
      const double pz = B->origin_z + kz * spacing + spacing / 2.0;
      for (int ky = 0; ky < B->ny; ++ky) {
        const double py = B->origin_y + ky * spacing + spacing / 2.0;
        for (int kx = 0; kx < B->nx; ++kx, ++i) {
          const double px = B->origin_x + kx * spacing + spacing / 2.0;
          double jx = 0.0, jy = 0.0, jz = 0.0;
          if (jitter > 0.0) {
            jx = (2.0 * sph_scene_uniform(seed, (uint64_t)i, 0) - 1.0) * jitter * spacing;
            jy = (2.0 * sph_scene_uniform(seed, (uint64_t)i, 1) - 1.0) * jitter * spacing;
            jz = (2.0 * sph_scene_uniform(seed, (uint64_t)i, 2) - 1.0) * jitter * spacing;
          }
          x[i] = (float)(px + jx);
          y[i] = (float)(py + jy);
          z[i] = (float)(pz + jz);
        }
      }
    }
  }
  return SPH_OK;
}

static double shape_distance(const sph_shape *s, double x, double y, double z) {
  const double px = x - s->center[0], py = y - s->center[1], pz = z - s->center[2];
  switch (s->kind) {
    case SPH_SHAPE_SPHERE:
      return sqrt(px * px + py * py + pz * pz) - s->radius;
    case SPH_SHAPE_TORUS: {
      /* ring of radius `radius` around `axis`, tube radius `radius2` */
      double a, b, c;
      if (s->axis == 0) { a = py; b = pz; c = px; }
      else if (s->axis == 1) { a = pz; b = px; c = py; }
      else { a = px; b = py; c = pz; }
      const double q = sqrt(a * a + b * b) - s->radius;
      return sqrt(q * q + c * c) - s->radius2;
    }
    case SPH_SHAPE_BOX: {
      const double qx = fabs(px) - s->half[0], qy = fabs(py) - s->half[1],
                   qz = fabs(pz) - s->half[2];
      const double ox = qx > 0 ? qx : 0, oy = qy > 0 ? qy : 0, oz = qz > 0 ? qz : 0;
      double inside = qx > qy ? qx : qy;
      inside = inside > qz ? inside : qz;
      if (inside > 0) inside = 0;
      return sqrt(ox * ox + oy * oy + oz * oz) + inside;
    }
    default:
      return HUGE_VAL;
  }
}

int sph_sdf_sample_shapes(const sph_shape *shapes, int nshapes, int dim_x, int dim_y, int dim_z,
                          const sph_aabb *bounds, float *out) {
  if (!shapes || nshapes < 1 || dim_x < 1 || dim_y < 1 || dim_z < 1 || !bounds || !out)
    return SPH_ERR_ARG;
  for (int s = 0; s < nshapes; ++s)
    if (shapes[s].kind < SPH_SHAPE_SPHERE || shapes[s].kind > SPH_SHAPE_BOX) return SPH_ERR_ARG;
  const double dx = (bounds->max_x - bounds->min_x) / dim_x;
  const double dy = (bounds->max_y - bounds->min_y) / dim_y;
  const double dz = (bounds->max_z - bounds->min_z) / dim_z;
  for (int k = 0; k < dim_z; ++k) {
    const double z = bounds->min_z + (k + 0.5) * dz;
    for (int j = 0; j < dim_y; ++j) {
      const double y = bounds->min_y + (j + 0.5) * dy;
      for (int i = 0; i < dim_x; ++i) {
        const double x = bounds->min_x + (i + 0.5) * dx;
        double d = HUGE_VAL;
        for (int s = 0; s < nshapes; ++s) {
          const double e = shape_distance(&shapes[s], x, y, z);
          if (e < d) d = e;
        }
        out[((size_t)k * dim_y + j) * dim_x + i] = (float)d;
      }
    }
  }
  return SPH_OK;
}

int sph_sdf_write_sdfgen(const char *path, int dim_x, int dim_y, int dim_z,
                         const double origin[3], double dx, const float *values) {
  if (!path || !origin || !values || dim_x < 1 || dim_y < 1 || dim_z < 1) return SPH_ERR_ARG;
  FILE *f = fopen(path, "w");
  if (!f) return SPH_ERR_ARG;
  fprintf(f, "%d %d %d\n%.9g %.9g %.9g\n%.9g\n", dim_x, dim_y, dim_z, origin[0], origin[1],
          origin[2], dx);
  const size_t n = (size_t)dim_x * dim_y * dim_z;
  for (size_t i = 0; i < n; ++i) fprintf(f, "%.9g\n", (double)values[i]);
  return fclose(f) == 0 ? SPH_OK : SPH_ERR_ARG;
}

/* AllocInitObstacle's parser (obstacle.c:13-66).  *values is malloc'd. */
int sph_sdf_read_sdfgen(const char *path, int dims[3], double origin[3], double *dx,
                        float **values) {
  if (!path || !dims || !origin || !dx || !values) return SPH_ERR_ARG;
  *values = NULL;
  FILE *f = fopen(path, "r");
  if (!f) return SPH_ERR_ARG;
  int rc = SPH_ERR_ARG;
  if (fscanf(f, "%d %d %d", &dims[0], &dims[1], &dims[2]) == 3 &&
      fscanf(f, "%lf %lf %lf", &origin[0], &origin[1], &origin[2]) == 3 &&
      fscanf(f, "%lf", dx) == 1 && dims[0] > 0 && dims[1] > 0 && dims[2] > 0) {
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    float *v = (float *)malloc(sizeof(float) * n);
    if (v) {
      size_t i = 0;
      for (; i < n; ++i)
        if (fscanf(f, "%f", &v[i]) != 1) break;
      if (i == n) {
        *values = v;
        rc = SPH_OK;
      } else {
        free(v);
      }
    }
  }
  fclose(f);
  return rc;
}

void sph_host_free(void *p) { free(p); }
