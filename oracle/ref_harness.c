/*
 * Test-only harness around the reference's own C code (TEST INFRASTRUCTURE).
 *
 * build_ref.sh compiles this file together with the unmodified reference
 * sources Simulation/src/particles.c and Simulation/src/neighbors.c, read in
 * place under the reference tree, into oracle/_ref/libsphref.so.  The functions
 * here only marshal flat arrays into the reference's structs
 * (particles.h:11-42, simulation.h:6-21, neighbors.h:14-29) and call the
 * reference's functions, so tests can compare the oracle with the reference
 * bit-for-bit.  Functions that need Thrust/CUDA/MPI (SortHash, FindCellBounds,
 * UpdatePositionStars's collision, AllocInit*) are not referenced and are
 * dropped at link time (--gc-sections).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "particles.h"
#include "simulation.h"
#include "neighbors.h"
#include "geometry.h"

typedef struct refh_params {
  double h, W_norm, DelW_poly_norm, DelW_norm, g, dt, k, c;
  double rest_density, rest_mass;
  double box_min[3], box_max[3];
  int nx, ny, nz;
} refh_params;

static void fill_params(struct Params *pp, const refh_params *rp) {
  memset(pp, 0, sizeof(*pp));
  pp->smoothing_radius = rp->h;
  pp->W_norm = rp->W_norm;
  pp->DelW_norm = rp->DelW_norm;
  pp->DelW_poly_norm = rp->DelW_poly_norm;
  pp->g = rp->g;
  pp->time_step = rp->dt;
  pp->k = rp->k;
  pp->c = rp->c;
  pp->dq = 0.1 * rp->h;
}

static void fill_aabb(struct AABB *b, const refh_params *rp) {
  b->min_x = rp->box_min[0]; b->max_x = rp->box_max[0];
  b->min_y = rp->box_min[1]; b->max_y = rp->box_max[1];
  b->min_z = rp->box_min[2]; b->max_z = rp->box_max[2];
}

typedef struct refh_arrays {
  double *xs, *ys, *zs, *x, *y, *z, *vx, *vy, *vz;
  double *dpx, *dpy, *dpz, *wx, *wy, *wz, *cx, *cy, *cz, *density, *lambda;
} refh_arrays;

static void fill_particles(struct Particles *P, const refh_arrays *a, int n,
                           const refh_params *rp) {
  memset(P, 0, sizeof(*P));
  P->x_star = a->xs; P->y_star = a->ys; P->z_star = a->zs;
  P->x = a->x; P->y = a->y; P->z = a->z;
  P->v_x = a->vx; P->v_y = a->vy; P->v_z = a->vz;
  P->dp_x = a->dpx; P->dp_y = a->dpy; P->dp_z = a->dpz;
  P->w_x = a->wx; P->w_y = a->wy; P->w_z = a->wz;
  P->color_x = a->cx; P->color_y = a->cy; P->color_z = a->cz;
  P->density = a->density; P->lambda = a->lambda;
  P->rest_density = rp->rest_density;
  P->rest_mass = rp->rest_mass;
  P->local_count = n;
  P->global_count = n;
  P->max_local = n;
  P->halo_count_left = 0;
  P->halo_count_right = 0;
}

static struct NeighborBucket *make_buckets(int n, const uint32_t *nbr, const int *cnt,
                                           int maxn) {
  struct NeighborBucket *b =
      (struct NeighborBucket *)calloc((size_t)(n > 0 ? n : 1), sizeof(struct NeighborBucket));
  for (int i = 0; i < n; ++i) {
    b[i].count = cnt[i];
    for (int j = 0; j < cnt[i] && j < MAX_NEIGHBORS; ++j)
      b[i].neighbor_indices[j] = nbr[(size_t)i * (size_t)maxn + (size_t)j];
  }
  return b;
}

static void fill_neighbors(struct Neighbors *N, const refh_params *rp) {
  memset(N, 0, sizeof(*N));
  N->hash_size_x = rp->nx;
  N->hash_size_y = rp->ny;
  N->hash_size_z = rp->nz;
  N->hash_spacing = rp->h;
}

int refh_max_neighbors(void) { return MAX_NEIGHBORS; }

void refh_hash_particles(int n, double *xs, double *ys, double *zs, const refh_params *rp,
                         uint32_t *keys, uint32_t *ids) {
  refh_arrays a; memset(&a, 0, sizeof(a));
  a.xs = xs; a.ys = ys; a.zs = zs;
  struct Particles P; fill_particles(&P, &a, n, rp);
  struct Neighbors N; fill_neighbors(&N, rp);
  N.hash_values = keys;
  N.particle_ids = ids;
  HashParticles(&P, &N);
}

/* start/end: cell tables; ids: sorted particle ids.  nbr is n*maxn. */
void refh_fill_neighbors(int n, double *xs, double *ys, double *zs, const refh_params *rp,
                         uint32_t *start, uint32_t *end, uint32_t *ids, uint32_t *nbr,
                         int *cnt, int maxn) {
  refh_arrays a; memset(&a, 0, sizeof(a));
  a.xs = xs; a.ys = ys; a.zs = zs;
  struct Particles P; fill_particles(&P, &a, n, rp);
  struct Params pp; fill_params(&pp, rp);
  struct Neighbors N; fill_neighbors(&N, rp);
  N.start_indices = start;
  N.end_indices = end;
  N.particle_ids = ids;
  N.neighbor_buckets =
      (struct NeighborBucket *)calloc((size_t)(n > 0 ? n : 1), sizeof(struct NeighborBucket));
  FillNeighbors(&P, &pp, &N);
  for (int i = 0; i < n; ++i) {
    cnt[i] = N.neighbor_buckets[i].count;
    for (int j = 0; j < cnt[i]; ++j)
      nbr[(size_t)i * (size_t)maxn + (size_t)j] = N.neighbor_buckets[i].neighbor_indices[j];
  }
  free(N.neighbor_buckets);
}

/* One reference stage on flat arrays.  stage selects the function. */
enum {
  REFH_GRAVITY = 0, REFH_PREDICT, REFH_DENSITY, REFH_LAMBDA, REFH_DPS,
  REFH_VELOCITIES, REFH_SURFACE, REFH_VORTICITY, REFH_CONFINEMENT, REFH_POSITIONS,
  REFH_BOUNDARY
};

void refh_stage(int stage, int n, refh_arrays *a, const refh_params *rp, const uint32_t *nbr,
                const int *cnt, int maxn, int substep) {
  struct Particles P; fill_particles(&P, a, n, rp);
  struct Params pp; fill_params(&pp, rp);
  struct AABB box; fill_aabb(&box, rp);
  struct Neighbors N; fill_neighbors(&N, rp);
  struct NeighborBucket *b = NULL;
  if (nbr) {
    b = make_buckets(n, nbr, cnt, maxn);
    N.neighbor_buckets = b;
  }
  switch (stage) {
    case REFH_GRAVITY: ApplyGravity(&P, &pp); break;
    case REFH_PREDICT: PredictPositions(&P, &pp, &box); break;
    case REFH_DENSITY: ComputeDensities(&P, &pp, &N); break;
    case REFH_LAMBDA: ComputeLambda(&P, &pp, &N); break;
    case REFH_DPS: UpdateDPs(&P, &pp, &N, substep); break;
    case REFH_VELOCITIES: UpdateVelocities(&P, &pp); break;
    case REFH_SURFACE: ComputeSurfaceTension(&P, &pp, &N); break;
    case REFH_VORTICITY: ComputeVorticity(&P, &pp, &N); break;
    case REFH_CONFINEMENT: ApplyVorticityConfinement(&P, &pp, &N); break;
    case REFH_POSITIONS: UpdatePositions(&P); break;
    case REFH_BOUNDARY:
      for (int i = 0; i < n; ++i) ApplyBoundaryConditions(&a->xs[i], &a->ys[i], &a->zs[i], &box, &pp);
      break;
    default: break;
  }
  free(b);
}

/* ApplyViscosity (particles.c:383-459) evaluated with Jacobi semantics: each
 * particle is run as a one-particle call on arrays offset to it, so every read
 * of a neighbour's velocity sees the value from before the pass.  Neighbour
 * indices are passed relative to the particle (two's-complement offsets),
 * which the reference converts back through `const int q_index`. */
void refh_viscosity_jacobi(int n, const double *xs, const double *ys, const double *zs,
                           const double *density, const double *vx, const double *vy,
                           const double *vz, const uint32_t *nbr, const int *cnt, int maxn,
                           const refh_params *rp, double *ovx, double *ovy, double *ovz) {
  double *wx = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double *wy = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double *wz = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  memcpy(wx, vx, sizeof(double) * (size_t)n);
  memcpy(wy, vy, sizeof(double) * (size_t)n);
  memcpy(wz, vz, sizeof(double) * (size_t)n);
  struct Params pp; fill_params(&pp, rp);
  struct Neighbors N; fill_neighbors(&N, rp);
  for (int i = 0; i < n; ++i) {
    struct NeighborBucket bucket;
    memset(&bucket, 0, sizeof(bucket));
    bucket.count = cnt[i];
    for (int j = 0; j < cnt[i]; ++j) {
      const int q = (int)nbr[(size_t)i * (size_t)maxn + (size_t)j];
      bucket.neighbor_indices[j] = (unsigned int)(q - i);
    }
    N.neighbor_buckets = &bucket;
    refh_arrays a; memset(&a, 0, sizeof(a));
    a.xs = (double *)xs + i; a.ys = (double *)ys + i; a.zs = (double *)zs + i;
    a.vx = wx + i; a.vy = wy + i; a.vz = wz + i;
    a.density = (double *)density + i;
    struct Particles P; fill_particles(&P, &a, 1, rp);
    ApplyViscosity(&P, &pp, &N);
    ovx[i] = wx[i]; ovy[i] = wy[i]; ovz[i] = wz[i];
    wx[i] = vx[i]; wy[i] = vy[i]; wz[i] = vz[i];
  }
  free(wx); free(wy); free(wz);
}
