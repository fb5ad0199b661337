/*
 * sphpc_amd -- MI355X (gfx950) position-based-fluids substep behind a C ABI.
 *
 * Drop-in GPU path for the single-rank solver of AdamSimpson/SPHPC.  The
 * reference's driver (Simulation/src/simulation.c:54-116) calls, per substep,
 * the stage functions of particles.h:44-92 and neighbors.h:35-47 on struct
 * pointers, and those drop to native code through thrust_c.h:47-69 and
 * obstacle_cuda.h:21-29.  This header exports the same stages, in the same
 * order and with the same meaning, on an opaque context that owns FP32 device
 * buffers, plus the native sort/bounds/collision entry points.
 *
 * Conventions
 *  - Every function returns 0 (SPH_OK) or a negative status; sph_last_error()
 *    gives the message.  The reference aborts on fatal errors (EXIT_PRINT,
 *    Common/print_macros.h:14-16); the bundled driver keeps that behaviour by
 *    exiting on a non-zero status.
 *  - Positions, velocities and all per-particle fields are FP32 on the device.
 *    Host<->device copies happen only in upload/download calls.
 *  - All work of one context runs in order on the context's own HIP stream;
 *    calls on one context must not overlap from several host threads.
 *  - The boundary box must start at the origin (the reference hashes absolute
 *    coordinates, neighbors.c:86-95, and tests x < 0, neighbors.c:257-260).
 */
#ifndef SPHPC_AMD_H_
#define SPHPC_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_MAX_NEIGHBORS 60 /* neighbors.h:6 */

enum sph_status {
  SPH_OK = 0,
  SPH_ERR_ARG = -1,      /* invalid argument */
  SPH_ERR_HIP = -2,      /* HIP runtime error (message names the call) */
  SPH_ERR_STATE = -3,    /* stage called out of order (e.g. no neighbour lists) */
  SPH_ERR_CAPACITY = -4, /* more particles than the context was created for */
  SPH_ERR_NOMEM = -5     /* device allocation failed */
};

/* Run parameters.  Fields of struct Params (simulation.h:6-21) plus the
 * constants the reference hard-codes; sph_default_params() returns the
 * reference's values. */
typedef struct sph_params {
  double smoothing_radius;  /* h; simulation.c:152 (2*float(spacing)) */
  double time_step;         /* dt; ini SimParameters.time_step */
  double g;                 /* gravity magnitude; ini PhysicalParameters.g */
  double k;                 /* surface tension gamma; ini PhysicalParameters.k */
  double c;                 /* XSPH viscosity; ini PhysicalParameters.c */
  double rest_density;      /* particles.c:1034 (1000) */
  double rest_mass;         /* particles.c:1035-1037 */
  double lambda_epsilon;    /* relaxation, particles.c:663 (0.01) */
  double vorticity_epsilon; /* confinement, particles.c:289 (0.1*h) */
  double v_max;             /* velocity clamp, particles.c:791 (100) */
  double stiffness_k;       /* stiffness schedule, particles.c:693 (0.95) */
  double density_tolerance; /* 0: fixed iterations (reference); >0: stop when
                               max_i max(rho_i-rho0,0)/rho0 <= tolerance */
  int solve_iterations;     /* Jacobi iterations (cap), simulation.c:72 (10) */
  int surface_tension;      /* run ComputeSurfaceTension (reference: 1) */
  int viscosity;            /* run ApplyViscosity (reference: 1) */
  int vorticity;            /* run ComputeVorticity+Confinement (reference: 1) */
  int vorticity_model;      /* 0: reference formula (particles.c:207-381);
                               1: omega weighted by m/rho_j, dv = dt*eps*(N x omega) */
} sph_params;

/* struct AABB, geometry.h:9-16 */
typedef struct sph_aabb {
  double min_x, max_x, min_y, max_y, min_z, max_z;
} sph_aabb;

/* Signed-distance obstacle (struct Obstacle, obstacle.h:8-29): an SDFGen grid
 * (x fastest) of dim_x*dim_y*dim_z FP32 values in model units with texel size
 * `spacing`, mapped onto the world box `world_bounds`. */
typedef struct sph_obstacle_desc {
  int dim_x, dim_y, dim_z;
  double spacing;
  sph_aabb world_bounds;
  const float *sdf; /* host pointer */
} sph_obstacle_desc;

/* Per-substep counters read back by sph_get_stats. */
typedef struct sph_stats {
  int particles;
  int iterations;            /* Jacobi updates applied in the last substep */
  int converged;             /* tolerance mode: reached the tolerance */
  float max_density_error;   /* last computed max_i max(rho_i-rho0,0)/rho0 */
  unsigned long long neighbor_overflow; /* hits dropped by the 60 cap (silent in
                                           the reference, neighbors.c:285-286),
                                           summed over every search the context ran */
  unsigned long long cell_clamps;       /* cell coordinates clamped into the grid,
                                           summed likewise */
  long long substeps;        /* substeps run on this context */
  unsigned long long neighbor_pairs;    /* sum of list lengths of the last search */
  int warm_fallback;         /* warm start reverted to the plain solve in the last substep:
                                1, or in a batch the number of scenes reverted (each
                                scene decides on its own maximum error) */
} sph_stats;

typedef struct sph_ctx sph_ctx;

const char *sph_last_error(void);
int sph_get_device_count(int *count);

/* Reference defaults for a given smoothing radius (simulation.c:130-178,
 * particles.c:289,663,691-694,791). */
void sph_default_params(sph_params *p, double smoothing_radius);

/* --- lifecycle: AllocInitParticles/AllocInitNeighbors/AllocInitObstacle
 *     (particles.h:104-108, neighbors.h:35-40, obstacle.h:32-33) ---------
 * A context holds at most SPH_MAX_PARTICLES particles (a batch: all scenes
 * together); its particle arrays are addressed with 32-bit byte offsets. */
#define SPH_MAX_PARTICLES ((1 << 28) - 64)
int sph_create(const sph_params *params, const sph_aabb *boundary, int max_particles,
               sph_ctx **out);
int sph_destroy(sph_ctx *ctx);  /* FinalizeParticles/FinalizeNeighbors */

/* Batched independent scenes (no counterpart in the reference, which runs one
 * scene per job): every stage runs once over all scenes' particles, each scene
 * with its own box [0, size) in its own coordinates, its own hash grid and
 * obstacle, and the shared physical parameters.  sph_upload_state then takes
 * all scenes' particles concatenated in scene order, and sph_download_state
 * returns them the same way. */
typedef struct sph_scene {
  int particles;
  double size_x, size_y, size_z;  /* box [0, size) */
  sph_obstacle_desc obstacle;     /* .sdf NULL: no obstacle */
} sph_scene;
int sph_create_batch(const sph_params *params, int nscenes, const sph_scene *scenes,
                     sph_ctx **out);
int sph_batch_scene_count(sph_ctx *ctx, int *nscenes);
/* Max density error max_i max(rho_i - rho0, 0) / rho0 of each scene at the
 * last density pass. */
int sph_batch_scene_errors(sph_ctx *ctx, float *max_err);
int sph_set_params(sph_ctx *ctx, const sph_params *params);
int sph_get_params(sph_ctx *ctx, sph_params *params);
int sph_grid_dims(sph_ctx *ctx, int *nx, int *ny, int *nz);
/* AllocInitObstacle_CUDA (obstacle_cuda.cu:133-183); sdf == NULL removes it */
int sph_set_obstacle(sph_ctx *ctx, const sph_obstacle_desc *obstacle);
/* Host -> device state (the acc enter data of particles.c:1056-1078). */
int sph_upload_state(sph_ctx *ctx, int n, const float *x, const float *y, const float *z,
                     const float *vx, const float *vy, const float *vz);
/* Device -> host state in particle-id order (acc update host, fileio_adios.c:73-76).
 * NULL arrays are skipped (x, y, z only for a frame); the copy goes through
 * pinned staging that the context allocates on its first transfer. */
int sph_download_state(sph_ctx *ctx, float *x, float *y, float *z, float *vx, float *vy,
                       float *vz);
int sph_particle_count(sph_ctx *ctx, int *n);

/* --- per-substep stages, in the driver's order (simulation.c:59-108) ---- */
int sph_apply_gravity(sph_ctx *ctx);               /* particles.c:516 */
int sph_predict_positions(sph_ctx *ctx);           /* particles.c:748 */
int sph_find_all_neighbors(sph_ctx *ctx);          /* neighbors.c:68 */
int sph_compute_densities(sph_ctx *ctx);           /* particles.c:461 */
int sph_compute_lambda(sph_ctx *ctx);              /* particles.c:588 */
int sph_update_dps(sph_ctx *ctx, int substep);     /* particles.c:668 */
int sph_update_position_stars(sph_ctx *ctx);       /* particles.c:529 */
int sph_update_velocities(sph_ctx *ctx);           /* particles.c:787 */
int sph_compute_surface_tension(sph_ctx *ctx);     /* particles.c:67 */
int sph_apply_viscosity(sph_ctx *ctx);             /* particles.c:383 */
int sph_compute_vorticity(sph_ctx *ctx);           /* particles.c:207 */
int sph_apply_vorticity_confinement(sph_ctx *ctx); /* particles.c:285 */
int sph_update_positions(sph_ctx *ctx);            /* particles.c:566 */

/* Whole substeps with fused kernels: the loop body of simulation.c:54-116
 * (output excluded), `substeps` times. */
int sph_step(sph_ctx *ctx, int substeps);
/* Replay whole substeps of sph_step from a HIP graph (captured after one eager
 * substep; rebuilt after any change of parameters, obstacle, particle count,
 * warm start, logging or profiling).  Results are bit-identical to the eager
 * launches.  Off by default; timing with sph_profile_enable uses eager runs. */
int sph_set_graph(sph_ctx *ctx, int enable);
int sph_synchronize(sph_ctx *ctx);
int sph_get_stats(sph_ctx *ctx, sph_stats *stats);
/* Max density error of every density pass of the last fused substep:
 * err[k] = max_i max(rho_i - rho0, 0) / rho0 after k position updates
 * (*count passes, at most cap). */
int sph_get_residuals(sph_ctx *ctx, float *err, int cap, int *count);
/* FP64 reductions: kinetic energy sum(m/2 |v|^2) and max density error. */
int sph_compute_energy(sph_ctx *ctx, double *kinetic_energy);

/* --- inspection (particle-id order; valid after sph_find_all_neighbors) -- */
enum sph_field {
  SPH_FIELD_POSITION_STAR = 0, /* x* (3 floats per particle) */
  SPH_FIELD_DENSITY = 1,       /* rho (1) */
  SPH_FIELD_LAMBDA = 2,        /* lambda (1) */
  SPH_FIELD_DP = 3,            /* delta p of sph_update_dps (3) */
  SPH_FIELD_VELOCITY = 4,      /* v in the velocity passes (3) */
  SPH_FIELD_COLOR = 5,         /* colour-field normal (3) */
  SPH_FIELD_VORTICITY = 6,     /* omega (3) */
  SPH_FIELD_POSITION_OLD = 7,  /* x at substep start (3) */
  SPH_FIELD_LAMBDA_SUM = 8     /* accumulated multiplier of the last warm-started or
                                  logged solve (1; see sph_set_warm_start) */
};
int sph_download_field(sph_ctx *ctx, int field, float *out);
/* Neighbour lists as particle ids in the reference order (NeighborBucket,
 * neighbors.h:25-28): ids[i*60 + j], j < counts[i].  ids NULL: counts only. */
int sph_download_neighbors(sph_ctx *ctx, uint32_t *ids, int *counts);
/* Sorted (cell key, particle id) pairs of the last neighbour search. */
int sph_download_sorted(sph_ctx *ctx, uint32_t *keys, uint32_t *ids);
/* Cell tables: start[c] = lower bound, end[c] = upper bound (cells entries). */
int sph_download_cell_bounds(sph_ctx *ctx, uint32_t *start, uint32_t *end);

/* --- native sub-boundary (device pointers; stream NULL = default) -------
 * These keep their scratch per device between calls (allocated on the first
 * call of a size, held until the process ends); calls on one device take
 * turns, and each returns after its work on `stream` is done. */
/* SortByKey, thrust_c.h:47-50: stable sort of n (key, value) pairs by key.
 * The radix passes follow the keys' highest set bit (one OR-reduction first). */
int sph_sort_by_key(uint32_t *keys, uint32_t *values, int n, void *stream);
/* Same, sorting only the low `key_bits` bits (keys must be < 2^key_bits). */
int sph_sort_by_key_bits(uint32_t *keys, uint32_t *values, int n, int key_bits,
                         void *stream);
/* No counterpart in the reference: sph_sort_by_key_bits for callers that
 * cannot allocate device memory (code under stream capture, serving loops,
 * harnesses that forbid hipMalloc).  Its scratch is a file-scope device array
 * sized for SPH_SORT_NOALLOC_MAX_PAIRS pairs, which every GPU the library runs
 * on holds until the process ends; larger n is rejected (SPH_ERR_ARG).  The
 * work is only enqueued on `stream` (no synchronize).  All calls on one GPU
 * share that array, so calls on one GPU that overlap (different streams
 * without ordering, or different threads) race. */
#define SPH_SORT_NOALLOC_MAX_PAIRS (1 << 21)
int sph_sort_by_key_bits_noalloc(uint32_t *keys, uint32_t *values, int n, int key_bits,
                                 void *stream);
/* FindLowerBounds/FindUpperBounds, thrust_c.h:52-56,65-69: for every cell
 * c in [0, cells) the lower/upper bound of c in n sorted keys. */
int sph_find_lower_bounds(const uint32_t *keys, int n, int cells, uint32_t *out,
                          void *stream);
int sph_find_upper_bounds(const uint32_t *keys, int n, int cells, uint32_t *out,
                          void *stream);
/* CalculateParticleCollisions, obstacle_cuda.h:21-29: push particles inside
 * the obstacle out along the SDF gradient.  The obstacle is an uploaded grid. */
typedef struct sph_sdf sph_sdf;
int sph_sdf_create(const sph_obstacle_desc *desc, sph_sdf **out);
int sph_sdf_destroy(sph_sdf *sdf);
int sph_calculate_particle_collisions(float *x, float *y, float *z, int n, const sph_sdf *sdf,
                                      void *stream);
/* The same on the reference's own FP64 device arrays, as CalculateParticle-
 * Collisions (obstacle_cuda.h:25-29, obstacle_cuda.cu:56-116) takes them: the
 * bounds test and normalised coordinates in double, the SDF arithmetic in
 * FP32, the push added to the double coordinates.  The reference's by-value
 * struct Obstacle (world box, SDF dims and spacing, a CUDA texture) is the
 * sph_sdf made from the same fields by sph_sdf_create. */
int sph_calculate_particle_collisions_f64(double *x, double *y, double *z, int n,
                                          const sph_sdf *sdf, void *stream);

/* Device memory plumbing for callers that hold raw device pointers. */
int sph_device_malloc(void **ptr, size_t bytes);
int sph_device_free(void *ptr);
int sph_memcpy_htod(void *dst, const void *src, size_t bytes);
int sph_memcpy_dtoh(void *dst, const void *src, size_t bytes);

/* Plain device-to-device copy kernel of `bytes` (a multiple of 16); used to
 * calibrate hardware byte counters against a known traffic volume. */
int sph_device_copy(void *dst, const void *src, size_t bytes, void *stream);

/* --- scene and obstacle inputs (host side) ------------------------------
 * Fluid blocks: the lattice of ConstructFluidVolume (geometry.c:42-55),
 * x fastest, centres at origin + (i + 1/2) * spacing; optional seeded
 * jitter U(-jitter, jitter) * spacing per axis. */
typedef struct sph_block {
  int nx, ny, nz;
  double origin_x, origin_y, origin_z;
} sph_block;
/* With x/y/z NULL only *count is set. */
int sph_scene_lattice(const sph_block *blocks, int nblocks, double spacing, double jitter,
                      uint64_t seed, float *x, float *y, float *z, int cap, int *count);
/* Uniform [0,1) draw of the jitter generator (seed, particle, axis). */
double sph_scene_uniform(uint64_t seed, uint64_t index, int axis);

enum sph_shape_kind { SPH_SHAPE_SPHERE = 0, SPH_SHAPE_TORUS = 1, SPH_SHAPE_BOX = 2 };
typedef struct sph_shape {
  int kind;
  int axis;          /* torus axis (0 x, 1 y, 2 z) */
  double center[3];
  double radius;     /* sphere radius; torus ring radius */
  double radius2;    /* torus tube radius */
  double half[3];    /* box half extents */
} sph_shape;
/* Union of signed distances sampled at the texel centres of a dim grid over
 * `bounds` (x fastest), the points tex3D filtering reads. */
int sph_sdf_sample_shapes(const sph_shape *shapes, int nshapes, int dim_x, int dim_y,
                          int dim_z, const sph_aabb *bounds, float *out);
/* SDFGen ASCII format (obstacle.c:13-66). read allocates *values (sph_host_free). */
int sph_sdf_write_sdfgen(const char *path, int dim_x, int dim_y, int dim_z,
                         const double origin[3], double dx, const float *values);
int sph_sdf_read_sdfgen(const char *path, int dims[3], double origin[3], double *dx,
                        float **values);
void sph_host_free(void *p);

/* sim-config.ini (ReadParameters, input_parser.cc:27-62) plus an optional
 * [Solver] section for the constants the reference hard-codes. */
typedef struct sph_run_config {
  int number_steps;
  double time_step;
  int number_particles;
  double g, c, k;
  sph_aabb boundary, initial_fluid;
  int has_obstacle;
  char sdf_file[256];
  double obstacle_min_x, obstacle_max_x, obstacle_min_y, obstacle_min_z;
  int solve_iterations;     /* simulation.c:72 (10) */
  double lambda_epsilon;    /* particles.c:663 (0.01) */
  double vorticity_epsilon; /* < 0: 0.1 h (particles.c:289) */
  double mass_factor;       /* particles.c:1037 (0.991) */
  double density_tolerance; /* 0: fixed iterations */
  int vorticity_model;      /* 0: reference */
  int surface_tension, viscosity, vorticity;  /* 1: on (reference) */
  double jitter;            /* lattice jitter in units of spacing (0) */
  unsigned long long seed;
  int output_every;         /* 0: int(1/(30 dt)) (simulation.c:113); < 0: never */
  char warm_start_net[256]; /* trained warm-start network file ("" = none) */
} sph_run_config;
int sph_config_load(const char *path, sph_run_config *cfg);
const char *sph_config_error(void);

/* --- per-kernel timing (HIP events on the context stream) -------------- */
int sph_profile_enable(sph_ctx *ctx, int enable);
/* Sum of event-timed durations (ms) and launch count of one kernel since the
 * last reset; kernel names as in sph_profile_kernel_name. */
int sph_profile_read(sph_ctx *ctx, int kernel, double *total_ms, long long *launches);
int sph_profile_reset(sph_ctx *ctx);
const char *sph_profile_kernel_name(int kernel);
int sph_profile_kernel_count(void);

/* --- launch autotuner (no counterpart in the reference) ------------------
 * Times `substeps` substeps for each candidate launch configuration from the
 * context's current state (restored afterwards): 64/128/256 threads per block
 * and 8/16/32 neighbour gathers in flight per wavefront for every
 * neighbour-gather kernel, with an obstacle the push fused into the
 * position update or as a separate pass, and compact (16-bit offsets) versus
 * wide (32-bit indices) neighbour lists.  Keeps the fastest choice per
 * kernel.  Results do not depend on the choice. */
#define SPH_TUNE_KERNELS 32
typedef struct sph_tuning {
  int split_collide;              /* obstacle push as a separate pass */
  int wide_lists;                 /* 32-bit list entries for every wavefront */
  int block[SPH_TUNE_KERNELS];    /* threads per block, by sph_profile_kernel_name index */
  int depth[SPH_TUNE_KERNELS];    /* gathers in flight per wavefront, same index */
  double ms_per_substep_default;  /* event-timed kernel sum before tuning */
  double ms_per_substep_tuned;    /* and with the chosen configuration */
} sph_tuning;
int sph_autotune(sph_ctx *ctx, int substeps, sph_tuning *result);
/* The context's current launch choices (timings zero), and applying a saved
 * set: block 0 / depth 0 keep the default (256 threads, 16 gathers); other
 * values must be 64, 128 or 256 and 8, 16 or 32.  Entries of kernels that
 * are not tuned are ignored. */
int sph_get_tuning(sph_ctx *ctx, sph_tuning *out);
int sph_set_tuning(sph_ctx *ctx, const sph_tuning *tuning);

/* --- learned lambda warm start ------------------------------------------
 * No counterpart in the reference: a 12 -> 64 -> 64 -> 1 SiLU network that
 * predicts, per particle, how much the iterations after the first add to the
 * accumulated multiplier of the constraint solve, Lambda = sum over
 * iterations of stiffness(it) * lambda_it (particles.c:664,693-694), from
 * quantities of the first density pass; the first step then applies that
 * remainder at once.  The plain Jacobi loop stays the acceptance test: with a
 * tolerance, iterations continue until the density error is met.
 *
 * Dataset rows are SPH_MLP_STRIDE floats, device memory, 16-byte aligned:
 *   0 neighbour count / 32           (NeighborBucket.count, neighbors.h:27)
 *   1 rho / rho0 - 1                 (ComputeDensities, particles.c:512)
 *   2 C_prev: -Lambda_prev * (|grad C|^2 + eps), previous substep
 *   3 |grad C|^2 / (|grad C|^2 + eps) (sum_C, particles.c:654-661)
 *   4 dt * div v (m/rho0-weighted spiky gradient of the predicted displacement)
 *   5 SDF distance / h at x*, clamped to [-4, 4] (4 outside the grid)
 *   6-8 normalised SDF gradient (0 outside the grid)
 *   9-11 first density moment sum_j m W_ij (x_j - x_i) / (h rho_i)
 *   12 |grad C|^2 + eps (not a network input; converts targets)
 *   13 lambda_0, the plain first-iteration lambda (not a network input)
 * The target is the remainder R = -(Lambda - lambda_0) (|grad C|^2 + eps)
 * (dimensionless, >= 0): what iterations after the first add to the
 * multiplier.  The warm-started first step uses lambda_0 - s max(R_pred, 0) /
 * (|grad C|^2 + eps), s the network's step scale; with R_pred = 0 it is the
 * plain first iteration. */
#define SPH_MLP_FEATURES 12
#define SPH_MLP_STRIDE 16
#define SPH_MLP_PARAMS 5313
typedef struct sph_mlp sph_mlp;
int sph_mlp_create(sph_mlp **out);  /* on the current device, own stream */
int sph_mlp_destroy(sph_mlp *net);
int sph_mlp_init(sph_mlp *net, unsigned long long seed);
/* FP32 master parameters: W1[16][64], b1[64], W2[64][64], b2[64], w3[64], b3
 * (row-major, row = input; W1 rows 12..15 are padding and stay zero). */
int sph_mlp_set_weights(sph_mlp *net, const float *theta);
int sph_mlp_get_weights(sph_mlp *net, float *theta);
/* Inputs are normalised as (x - mean) / stddev (stddev 0: input ignored),
 * the output is y * target_std + target_mean. */
int sph_mlp_set_normalization(sph_mlp *net, const float *mean, const float *stddev,
                              float target_mean, float target_std);
int sph_mlp_get_normalization(sph_mlp *net, float *mean, float *stddev, float *target_mean,
                              float *target_std);
int sph_mlp_fit_normalization(sph_mlp *net, const float *x_dev, const float *t_dev,
                              size_t count);
/* Caps the output used by the solver at max_target (0: no cap); saved with
 * the network.  Training clips its targets at the same bound. */
int sph_mlp_set_output_clip(sph_mlp *net, float max_target);
float sph_mlp_get_output_clip(sph_mlp *net);
/* The solver applies scale * Lambda_pred (default 1; a damping fitted on held-out
 * scenes by the trainer); saved with the network. */
int sph_mlp_set_step_scale(sph_mlp *net, float scale);
float sph_mlp_get_step_scale(sph_mlp *net);
/* Full-batch gradient of 0.5/B sum (y_n - t_n)^2 in normalised target units
 * (host outputs; either may be NULL). */
int sph_mlp_gradient(sph_mlp *net, const float *x_dev, const float *t_dev, size_t count,
                     float *grad, float *loss);
/* One Adam step (lr; beta 0.9/0.999, eps 1e-8) on `batch` rows drawn with
 * replacement from `count` by a counter-based generator keyed on `seed`
 * (batch 0: all rows in order).  loss != NULL waits for and returns the loss
 * before the update. */
int sph_mlp_train_step(sph_mlp *net, const float *x_dev, const float *t_dev, size_t count,
                       size_t batch, float lr, unsigned long long seed, float *loss);
int sph_mlp_reset_optimizer(sph_mlp *net);
int sph_mlp_forward(sph_mlp *net, const float *x_dev, size_t count, float *y_dev);   /* FP32 */
int sph_mlp_infer_bf16(sph_mlp *net, const float *x_dev, size_t count, float *y_dev); /* MFMA */
int sph_mlp_synchronize(sph_mlp *net);
int sph_mlp_save(sph_mlp *net, const char *path);
int sph_mlp_load(sph_mlp *net, const char *path);

/* Warm-start the solve of every following fused substep with `net` (its
 * current weights are copied; NULL turns the warm start off).  If the
 * predicted step raises the maximum density error, the substep falls back to
 * the plain first iteration (sph_stats.warm_fallback). */
int sph_set_warm_start(sph_ctx *ctx, sph_mlp *net);
/* The non-learned baseline of the same warm start: each following fused
 * substep starts its solve from `scale` times the particle's accumulated
 * multiplier of the previous substep, with the same plain iterations after it
 * and the same fallback (the reference restarts every solve from zero,
 * particles.c:588-666).  0 turns it off; a network warm start is turned off
 * by it and turns it off.  Scale in [0, 4]. */
int sph_set_warm_start_previous(sph_ctx *ctx, float scale);
/* Record features and accumulated multipliers of every following fused
 * substep (plain solve; takes precedence over the warm start). */
int sph_set_sample_logging(sph_ctx *ctx, int enable);
/* Copies every `every`-th particle's features and target of the last logged
 * substep to device rows x_dev[offset..], t_dev[offset..]; *written = rows. */
int sph_warm_samples(sph_ctx *ctx, float *x_dev, float *t_dev, size_t offset,
                     size_t capacity, int every, size_t *written);

#ifdef __cplusplus
}
#endif
#endif
