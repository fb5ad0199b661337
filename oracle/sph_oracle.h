/*
 * FP64 CPU oracle for the position-based-fluids substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * entry point and bench.py's cpu_baseline leg may call it; the product library
 * (sphpc_amd/) never links, loads or calls it.
 *
 * It restates the reference's CPU path (AdamSimpson/SPHPC, Simulation/src/...)
 * in plain C with the reference's operation order, FP64 arithmetic and
 * -ffp-contract=off, so that it agrees bit-for-bit with the reference's own
 * particles.c/neighbors.c compiled on the host (oracle/_ref, built by the `ref` target of oracle/Makefile).
 * Each function cites the reference lines it follows.  The parts the reference
 * does not implement in C (Thrust sort/bounds, the CUDA texture collision) are
 * restated from their documented semantics:
 *   - SortByKey (thrust_c.cu:86-91): stable sort of (key, value) by key.
 *   - FindLowerBounds/FindUpperBounds (thrust_c.cu:93-120): lower/upper bound
 *     of every cell index 0..cells-1 in the sorted keys.
 *   - CalculateParticleCollisions (obstacle_cuda.cu:56-116): tex3D linear
 *     filtering is restated as exact trilinear interpolation with texel centres
 *     at (i+0.5)/dim and clamp addressing.
 *
 * Deviations from the reference, each switchable or documented:
 *   - ApplyViscosity (particles.c:383-459) updates v in place while reading
 *     neighbours' v (a data race on the GPU, Gauss-Seidel on one CPU thread).
 *     The oracle reads v_old and writes v_new (Jacobi).
 *   - obstacle.c:74-75 sets world max_y/max_z without adding min_y/min_z; the
 *     oracle takes world bounds as given (callers pass min + extent).
 *   - A zero SDF gradient (obstacle_cuda.cu:15, rsqrtf(0)) leaves the particle
 *     where it is instead of writing NaN.
 *   - vort_model 1 is an optional dimensionally consistent vorticity
 *     (omega_i = sum m/rho_j (v_j-v_i) x gradW, dv = dt*eps*(N x omega)),
 *     not in the reference; vort_model 0 is the reference formula.
 */
#ifndef SPH_ORACLE_H_
#define SPH_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  double h;               /* smoothing radius, simulation.c:152 */
  double W_norm;          /* simulation.c:175 */
  double DelW_poly_norm;  /* simulation.c:176 */
  double DelW_norm;       /* simulation.c:177 */
  double g;               /* ini PhysicalParameters.g */
  double dt;              /* ini SimParameters.time_step */
  double k;               /* surface tension gamma, ini PhysicalParameters.k */
  double c;               /* XSPH coefficient, ini PhysicalParameters.c */
  double rest_density;    /* particles.c:1034 */
  double rest_mass;       /* particles.c:1035-1037 */
  double lambda_eps;      /* particles.c:663 (0.01) */
  double vort_eps;        /* particles.c:289 (0.1*h) */
  double v_max;           /* particles.c:791 (100) */
  double k_stiff;         /* particles.c:693 (0.95) */
  double box_min[3];      /* BoundaryVolume min (x,y,z) */
  double box_max[3];      /* BoundaryVolume max */
  int nx, ny, nz;         /* hash grid dims, neighbors.c:23-28 */
  int max_neighbors;      /* neighbors.h:6 (60) */
  int vort_model;         /* 0 = reference formula */
  int surface_tension_on; /* run ComputeSurfaceTension */
  int viscosity_on;       /* run ApplyViscosity */
  int vorticity_on;       /* run ComputeVorticity + confinement */
} orc_params;

typedef struct orc_obstacle {
  int dim[3];            /* SDF_dim_x/y/z, obstacle.c:25-27 */
  double spacing;        /* SDF_spacing (model units per texel), obstacle.c:46 */
  double wmin[3];        /* world AABB of the SDF grid */
  double wmax[3];
  const float *sdf;      /* dim[0]*dim[1]*dim[2] values, x fastest */
} orc_obstacle;

/* Fill derived kernel norms and grid dims the way SetParameters and
 * AllocInitNeighbors do (simulation.c:152-177, neighbors.c:23-28). */
void orc_params_derive(orc_params *p);

/* Kernels, particles.c:18-65. */
double orc_W(double r, double h, double norm);
double orc_DelWPoly(double r, double h, double norm);
double orc_DelW(double r, double h, double norm);
double orc_C(double r, double h, double norm);

/* Neighbour search, neighbors.c:79-298, thrust_c.cu:86-120. */
uint32_t orc_hash_val(const orc_params *p, double x, double y, double z);
void orc_hash_particles(int n, const double *xs, const double *ys, const double *zs,
                        const orc_params *p, uint32_t *keys, uint32_t *ids);
void orc_sort_by_key(int n, uint32_t *keys, uint32_t *vals);
void orc_find_lower_bounds(const uint32_t *keys, int n, int cells, uint32_t *out);
void orc_find_upper_bounds(const uint32_t *keys, int n, int cells, uint32_t *out);
/* Returns the number of (particle, neighbour) hits dropped by the cap. */
long orc_fill_neighbors(int n, const double *xs, const double *ys, const double *zs,
                        const orc_params *p, const uint32_t *start, const uint32_t *end,
                        const uint32_t *ids, uint32_t *nbr, int *cnt);
long orc_find_all_neighbors(int n, const double *xs, const double *ys, const double *zs,
                            const orc_params *p, uint32_t *nbr, int *cnt);

/* Physics stages, particles.c. nbr is n*max_neighbors, cnt is n. */
void orc_apply_gravity(int n, double *vz, const orc_params *p);
void orc_apply_boundary(double *x, double *y, double *z, const orc_params *p);
void orc_predict_positions(int n, const double *x, const double *y, const double *z,
                           const double *vx, const double *vy, const double *vz,
                           double *xs, double *ys, double *zs, const orc_params *p);
void orc_compute_densities(int n, const double *xs, const double *ys, const double *zs,
                           const uint32_t *nbr, const int *cnt, const orc_params *p,
                           double *density);
void orc_compute_lambda(int n, const double *xs, const double *ys, const double *zs,
                        const double *density, const uint32_t *nbr, const int *cnt,
                        const orc_params *p, double *lambda);
double orc_stiffness(const orc_params *p, int substep);
void orc_update_dps(int n, const double *xs, const double *ys, const double *zs,
                    const double *lambda, const uint32_t *nbr, const int *cnt,
                    const orc_params *p, int substep, double *dpx, double *dpy, double *dpz);
void orc_collide(int n, double *xs, double *ys, double *zs, const orc_obstacle *o);
void orc_update_position_stars(int n, double *xs, double *ys, double *zs,
                               const double *dpx, const double *dpy, const double *dpz,
                               const orc_params *p, const orc_obstacle *o);
void orc_update_velocities(int n, const double *xs, const double *ys, const double *zs,
                           const double *x, const double *y, const double *z,
                           double *vx, double *vy, double *vz, const orc_params *p);
void orc_compute_color(int n, const double *xs, const double *ys, const double *zs,
                       const double *density, const uint32_t *nbr, const int *cnt,
                       const orc_params *p, double *cx, double *cy, double *cz);
void orc_apply_surface_force(int n, const double *xs, const double *ys, const double *zs,
                             const double *density, const double *cx, const double *cy,
                             const double *cz, const uint32_t *nbr, const int *cnt,
                             const orc_params *p, double *vx, double *vy, double *vz);
void orc_apply_viscosity(int n, const double *xs, const double *ys, const double *zs,
                         const double *density, const double *vx, const double *vy,
                         const double *vz, const uint32_t *nbr, const int *cnt,
                         const orc_params *p, double *ovx, double *ovy, double *ovz);
void orc_compute_vorticity(int n, const double *xs, const double *ys, const double *zs,
                           const double *density, const double *vx, const double *vy,
                           const double *vz, const uint32_t *nbr, const int *cnt,
                           const orc_params *p, double *wx, double *wy, double *wz);
void orc_apply_vorticity_confinement(int n, const double *xs, const double *ys,
                                     const double *zs, const double *wx, const double *wy,
                                     const double *wz, const uint32_t *nbr, const int *cnt,
                                     const orc_params *p, double *vx, double *vy, double *vz);
void orc_update_positions(int n, const double *xs, const double *ys, const double *zs,
                          double *x, double *y, double *z);

/* Max density error max_i max(rho_i - rho0, 0)/rho0 (PrintMaxDensity,
 * particles.c:884-896, expressed as a compression error). */
double orc_max_density_error(int n, const double *density, double rest_density);

/* Full substep in the driver's order (simulation.c:59-108).
 * iters: fixed Jacobi iteration count (simulation.c:72).  If tol > 0 the loop
 * stops before the constraint update once the max density error is <= tol
 * (at most iters iterations).  Returns the number of iterations run.
 * Outputs of the last stage values are written into the optional arrays. */
typedef struct orc_state {
  double *x, *y, *z, *vx, *vy, *vz;   /* canonical state, in/out */
  double *xs, *ys, *zs;               /* x* (out) */
  double *density, *lambda;           /* last Jacobi values (out) */
  uint32_t *nbr; int *cnt;            /* neighbour lists (out) */
} orc_state;
/* OpenMP threads of the following calls (timing only; results do not depend on it). */
void orc_set_threads(int threads);
int orc_substep(int n, orc_state *s, const orc_params *p, const orc_obstacle *o,
                int iters, double tol, double *max_err_out);
/* The same substep with every array a stage writes rounded to float32 after
 * the stage (x*, velocities, density, lambda, delta p, colour and vorticity
 * fields): the storage of an FP32 solver, for the long-run envelope. */
int orc_substep_f32_storage(int n, orc_state *s, const orc_params *p, const orc_obstacle *o,
                            int iters, double tol, double *max_err_out);

#ifdef __cplusplus
}
#endif
#endif
