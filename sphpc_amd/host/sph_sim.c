/*
 * sph_sim: the reference's driver (Simulation/src/simulation.c:19-128) over
 * the C ABI of libsphpc_amd.  Same configuration file (./sim-config.ini or
 * argv[1]), same derived parameters (SetParameters, simulation.c:130-178),
 * same lattice (ConstructFluidVolume, geometry.c:19-74), same per-substep call
 * order (simulation.c:59-108) and the same 30 FPS output cadence
 * (simulation.c:113) with the MPI-IO writer's frame layout
 * (fileio_MPI_IO.c:63-118: interleaved FP64 x,y,z per particle, sim-<k>.bin).
 *
 * Options: --fused runs each substep through sph_step (fused kernels) instead
 * of the staged calls; --no-output disables frame writing; --steps N
 * overrides number_steps.  Any non-zero status prints the library's message
 * and exits, like EXIT_PRINT (Common/print_macros.h:14-16).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "sphpc_amd.h"

#define CHECK(call)                                                          \
  do {                                                                       \
    int rc_ = (call);                                                        \
    if (rc_ != SPH_OK) {                                                     \
      fprintf(stderr, "ERROR: %s:%d: %s failed (%d): %s\n", __FILE__, __LINE__, \
              #call, rc_, sph_last_error());                                 \
      exit(EXIT_FAILURE);                                                    \
    }                                                                        \
  } while (0)

static double now(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

static void write_frame(int k, int n, const float *x, const float *y, const float *z) {
  char path[256];
  mkdir("output", 0755);
  snprintf(path, sizeof(path), "output/sim-%d.bin", k);
  FILE *f = fopen(path, "wb");
  if (!f) {
    fprintf(stderr, "ERROR: cannot open %s\n", path);
    exit(EXIT_FAILURE);
  }
  for (int i = 0; i < n; ++i) {
    const double p[3] = {x[i], y[i], z[i]};
    fwrite(p, sizeof(double), 3, f);
  }
  fclose(f);
}

int main(int argc, char **argv) {
  const char *ini = "./sim-config.ini";
  int fused = 0, output = 1, steps_override = -1;
  for (int i = 1; i < argc; ++i) {
    if (!strcmp(argv[i], "--fused")) fused = 1;
    else if (!strcmp(argv[i], "--no-output")) output = 0;
    else if (!strcmp(argv[i], "--steps") && i + 1 < argc) steps_override = atoi(argv[++i]);
    else ini = argv[i];
  }
  sph_run_config cfg;
  if (sph_config_load(ini, &cfg) != SPH_OK) {
    printf("%s\n", sph_config_error());
    exit(-1);
  }
  if (steps_override >= 0) cfg.number_steps = steps_override;

  /* SetParameters, simulation.c:141-155 */
  const sph_aabb *fl = &cfg.initial_fluid;
  const double volume = (fl->max_x - fl->min_x) * (fl->max_y - fl->min_y) * (fl->max_z - fl->min_z);
  const float spacing = (float)pow(volume / cfg.number_particles, 1.0 / 3.0);
  const double rest_radius = spacing / 2.0;
  printf("particle radius: %f\n", rest_radius);
  const double h = 2.0 * spacing;
  printf("smoothing radius: %f\n", h);

  /* ConstructFluidVolume, geometry.c:23-74 (single rank) */
  const double s = 2.0 * rest_radius;
  sph_block block;
  block.nx = (int)floor((fl->max_x - fl->min_x) / s);
  block.ny = (int)floor((fl->max_y - fl->min_y) / s);
  block.nz = (int)floor((fl->max_z - fl->min_z) / s);
  block.origin_x = fl->min_x;
  block.origin_y = fl->min_y;
  block.origin_z = fl->min_z;
  int n = 0;
  CHECK(sph_scene_lattice(&block, 1, s, cfg.jitter, cfg.seed, NULL, NULL, NULL, 0, &n));
  float *buf = (float *)calloc((size_t)(n > 0 ? n : 1) * 6, sizeof(float));
  float *x = buf, *y = buf + n, *z = buf + 2 * (size_t)n;
  float *vx = buf + 3 * (size_t)n, *vy = buf + 4 * (size_t)n, *vz = buf + 5 * (size_t)n;
  CHECK(sph_scene_lattice(&block, 1, s, cfg.jitter, cfg.seed, x, y, z, n, &n));
  const double fluid_volume = (block.nx * s) * (block.ny * s) * (block.nz * s);

  /* AllocInitParticles, particles.c:1034-1040 */
  sph_params P;
  sph_default_params(&P, h);
  P.time_step = cfg.time_step;
  P.g = cfg.g;
  P.c = cfg.c;
  P.k = cfg.k;
  P.rest_density = 1000.0;
  P.rest_mass = fluid_volume * P.rest_density / (double)n * cfg.mass_factor;
  P.solve_iterations = cfg.solve_iterations;
  P.lambda_epsilon = cfg.lambda_epsilon;
  P.vorticity_epsilon = cfg.vorticity_epsilon < 0.0 ? 0.1 * h : cfg.vorticity_epsilon;
  P.density_tolerance = cfg.density_tolerance;
  P.vorticity_model = cfg.vorticity_model;
  P.surface_tension = cfg.surface_tension;
  P.viscosity = cfg.viscosity;
  P.vorticity = cfg.vorticity;
  printf("Rest mass: %f\n", P.rest_mass);
  printf("Rest Density: %f\n", P.rest_density);

  sph_ctx *ctx = NULL;
  CHECK(sph_create(&P, &cfg.boundary, n, &ctx));
  CHECK(sph_upload_state(ctx, n, x, y, z, vx, vy, vz));

  /* AllocInitObstacle (obstacle.c:13-79); world max_y/max_z follow the grid's
   * aspect ratio from min_y/min_z (the reference omits min_y/min_z there). */
  if (cfg.has_obstacle) {
    int dims[3];
    double origin[3], dx;
    float *values = NULL;
    if (sph_sdf_read_sdfgen(cfg.sdf_file, dims, origin, &dx, &values) != SPH_OK) {
      fprintf(stderr, "ERROR: Error parsing SDF file: %s\n", cfg.sdf_file);
      exit(EXIT_FAILURE);
    }
    sph_obstacle_desc od;
    od.dim_x = dims[0];
    od.dim_y = dims[1];
    od.dim_z = dims[2];
    od.spacing = dx;
    const double len = cfg.obstacle_max_x - cfg.obstacle_min_x;
    od.world_bounds.min_x = cfg.obstacle_min_x;
    od.world_bounds.max_x = cfg.obstacle_max_x;
    od.world_bounds.min_y = cfg.obstacle_min_y;
    od.world_bounds.max_y = cfg.obstacle_min_y + len * dims[1] / (float)dims[0];
    od.world_bounds.min_z = cfg.obstacle_min_z;
    od.world_bounds.max_z = cfg.obstacle_min_z + len * dims[2] / (float)dims[0];
    od.sdf = values;
    CHECK(sph_set_obstacle(ctx, &od));
    sph_host_free(values);
  }

  const int every = cfg.output_every > 0 ? cfg.output_every
                                         : (int)(1.0 / (cfg.time_step * 30.0));
  int frame = 0;
  if (output && cfg.output_every >= 0) write_frame(frame++, n, x, y, z);

  /* Learned warm start of the fused solve (no counterpart in the reference). */
  sph_mlp *net = NULL;
  if (fused && cfg.warm_start_net[0]) {
    CHECK(sph_mlp_create(&net));
    CHECK(sph_mlp_load(net, cfg.warm_start_net));
    CHECK(sph_set_warm_start(ctx, net));
    printf("Warm start: %s\n", cfg.warm_start_net);
  }

  const double start = now();
  for (int step = 0; step < cfg.number_steps; ++step) {
    if (fused) {
      CHECK(sph_step(ctx, 1));
    } else {
      CHECK(sph_apply_gravity(ctx));
      CHECK(sph_predict_positions(ctx));
      CHECK(sph_find_all_neighbors(ctx));
      for (int it = 0; it < P.solve_iterations; ++it) {
        CHECK(sph_compute_densities(ctx));
        CHECK(sph_compute_lambda(ctx));
        CHECK(sph_update_dps(ctx, it));
        CHECK(sph_update_position_stars(ctx));
      }
      CHECK(sph_update_velocities(ctx));
      if (P.surface_tension) CHECK(sph_compute_surface_tension(ctx));
      if (P.viscosity) CHECK(sph_apply_viscosity(ctx));
      if (P.vorticity) {
        CHECK(sph_compute_vorticity(ctx));
        CHECK(sph_apply_vorticity_confinement(ctx));
      }
      CHECK(sph_update_positions(ctx));
    }
    if (output && cfg.output_every >= 0 && every > 0 && step % every == 0) {
      /* a frame holds positions only (fileio_MPI_IO.c:79-106) */
      CHECK(sph_download_state(ctx, x, y, z, NULL, NULL, NULL));
      write_frame(frame++, n, x, y, z);
    }
  }
  CHECK(sph_synchronize(ctx));
  printf("Total loop time: %f seconds\n", now() - start);
  sph_stats st;
  CHECK(sph_get_stats(ctx, &st));
  printf("particles %d  substeps %lld  last max density error %.6f  neighbour overflow %llu\n",
         st.particles, st.substeps, st.max_density_error, st.neighbor_overflow);
  CHECK(sph_destroy(ctx));
  if (net) CHECK(sph_mlp_destroy(net));
  free(buf);
  return 0;
}
