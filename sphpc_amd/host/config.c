/*
 * sim-config.ini reader without Boost.  Required keys are the ones
 * ReadParameters reads (input_parser.cc:27-62); a missing one fails with the
 * reference's "Aborting:" message (input_parser.cc:64-67) as an error status.
 * [Obstacle] is optional here (Simulation/sim-config.ini has none, which makes
 * the reference abort), and a [Solver] section carries the constants the
 * reference hard-codes, with the reference's values as defaults.
 */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sphpc_amd.h"

#define MAX_ENTRIES 256

typedef struct {
  char key[96];  /* "Section.key" */
  char value[256];
} entry;

static __thread char g_msg[512];  /* per thread, like sph_last_error */

const char *sph_config_error(void) { return g_msg; }

static void trim(char *s) {
  size_t n = strlen(s);
  while (n && isspace((unsigned char)s[n - 1])) s[--n] = 0;
  size_t i = 0;
  while (s[i] && isspace((unsigned char)s[i])) ++i;
  if (i) memmove(s, s + i, n - i + 1);
}

static const char *find(const entry *e, int n, const char *key) {
  for (int i = 0; i < n; ++i)
    if (strcmp(e[i].key, key) == 0) return e[i].value;
  return NULL;
}

static int get_double(const entry *e, int n, const char *key, double *out, int required) {
  const char *v = find(e, n, key);
  if (!v) {
    if (required) {
      snprintf(g_msg, sizeof(g_msg), "Aborting: No such node (%s)", key);
      return SPH_ERR_ARG;
    }
    return SPH_OK;
  }
  char *end = NULL;
  const double d = strtod(v, &end);
  if (end == v || *end) {
    snprintf(g_msg, sizeof(g_msg), "Aborting: conversion of data to type \"d\" failed (%s)",
             key);
    return SPH_ERR_ARG;
  }
  *out = d;
  return SPH_OK;
}

static int get_int(const entry *e, int n, const char *key, int *out, int required) {
  double d = *out;
  int rc = get_double(e, n, key, &d, required);
  if (rc) return rc;
  if (d != floor(d) || !(d >= -2147483648.0 && d <= 2147483647.0)) {
    snprintf(g_msg, sizeof(g_msg), "Aborting: conversion of data to type \"i\" failed (%s)",
             key);
    return SPH_ERR_ARG;
  }
  *out = (int)d;
  return SPH_OK;
}

static int parse_entries(const entry *E, int n, sph_run_config *cfg);

int sph_config_load(const char *path, sph_run_config *cfg) {
  g_msg[0] = 0;
  if (!path || !cfg) return SPH_ERR_ARG;
  FILE *f = fopen(path, "r");
  if (!f) {
    snprintf(g_msg, sizeof(g_msg), "Aborting: %s: cannot open file", path);
    return SPH_ERR_ARG;
  }
  entry *E = (entry *)calloc(MAX_ENTRIES, sizeof(entry));
  if (!E) {
    fclose(f);
    snprintf(g_msg, sizeof(g_msg), "Aborting: out of memory reading %s", path);
    return SPH_ERR_NOMEM;
  }
  int n = 0;
  char section[64] = "";
  char line[512];
  while (fgets(line, sizeof(line), f)) {
    char *c = strpbrk(line, ";#");
    if (c) *c = 0;
    trim(line);
    if (!line[0]) continue;
    if (line[0] == '[') {
      char *r = strchr(line, ']');
      if (!r) continue;
      *r = 0;
      snprintf(section, sizeof(section), "%s", line + 1);
      trim(section);
      continue;
    }
    char *eq = strchr(line, '=');
    if (!eq || n >= MAX_ENTRIES) continue;
    *eq = 0;
    char k[64], v[256];
    snprintf(k, sizeof(k), "%s", line);
    snprintf(v, sizeof(v), "%s", eq + 1);
    trim(k);
    trim(v);
    snprintf(E[n].key, sizeof(E[n].key), "%s.%s", section, k);
    snprintf(E[n].value, sizeof(E[n].value), "%s", v);
    ++n;
  }
  fclose(f);
  const int rc_all = parse_entries(E, n, cfg);
  free(E);
  return rc_all;
}

static int parse_entries(const entry *E, int n, sph_run_config *cfg) {
  memset(cfg, 0, sizeof(*cfg));
  int rc = 0;
  rc |= get_int(E, n, "SimParameters.number_steps", &cfg->number_steps, 1);
  if (!rc) rc |= get_double(E, n, "SimParameters.time_step", &cfg->time_step, 1);
  if (!rc) rc |= get_int(E, n, "SimParameters.number_particles", &cfg->number_particles, 1);
  if (!rc) rc |= get_double(E, n, "PhysicalParameters.g", &cfg->g, 1);
  if (!rc) rc |= get_double(E, n, "PhysicalParameters.c", &cfg->c, 1);
  if (!rc) rc |= get_double(E, n, "PhysicalParameters.k", &cfg->k, 1);
  const char *bnames[6] = {"min_x", "min_y", "min_z", "max_x", "max_y", "max_z"};
  double *bv[6] = {&cfg->boundary.min_x, &cfg->boundary.min_y, &cfg->boundary.min_z,
                   &cfg->boundary.max_x, &cfg->boundary.max_y, &cfg->boundary.max_z};
  double *fv[6] = {&cfg->initial_fluid.min_x, &cfg->initial_fluid.min_y,
                   &cfg->initial_fluid.min_z, &cfg->initial_fluid.max_x,
                   &cfg->initial_fluid.max_y, &cfg->initial_fluid.max_z};
  char key[96];
  for (int i = 0; i < 6 && !rc; ++i) {
    snprintf(key, sizeof(key), "BoundaryVolume.%s", bnames[i]);
    rc |= get_double(E, n, key, bv[i], 1);
  }
  for (int i = 0; i < 6 && !rc; ++i) {
    snprintf(key, sizeof(key), "InitialFluidVolume.%s", bnames[i]);
    rc |= get_double(E, n, key, fv[i], 1);
  }
  if (rc) return SPH_ERR_ARG;

  const char *sdf = find(E, n, "Obstacle.SDF_file_name");
  if (sdf) {
    cfg->has_obstacle = 1;
    snprintf(cfg->sdf_file, sizeof(cfg->sdf_file), "%s", sdf);
    rc |= get_double(E, n, "Obstacle.min_x", &cfg->obstacle_min_x, 1);
    if (!rc) rc |= get_double(E, n, "Obstacle.max_x", &cfg->obstacle_max_x, 1);
    if (!rc) rc |= get_double(E, n, "Obstacle.min_y", &cfg->obstacle_min_y, 1);
    if (!rc) rc |= get_double(E, n, "Obstacle.min_z", &cfg->obstacle_min_z, 1);
    if (rc) return SPH_ERR_ARG;
  }

  /* [Solver]: constants the reference hard-codes (defaults = reference) */
  cfg->solve_iterations = 10;       /* simulation.c:72 */
  cfg->lambda_epsilon = 0.01;       /* particles.c:663 */
  cfg->vorticity_epsilon = -1.0;    /* < 0: 0.1 h, particles.c:289 */
  cfg->mass_factor = 0.991;         /* particles.c:1037 */
  cfg->density_tolerance = 0.0;
  cfg->vorticity_model = 0;
  cfg->surface_tension = 1;
  cfg->viscosity = 1;
  cfg->vorticity = 1;
  cfg->jitter = 0.0;
  cfg->seed = 0;
  cfg->output_every = 0;            /* 0: every int(1/(30 dt)) steps, simulation.c:113 */
  double seed = 0.0;
  rc |= get_int(E, n, "Solver.solve_iterations", &cfg->solve_iterations, 0);
  rc |= get_double(E, n, "Solver.lambda_epsilon", &cfg->lambda_epsilon, 0);
  rc |= get_double(E, n, "Solver.vorticity_epsilon", &cfg->vorticity_epsilon, 0);
  rc |= get_double(E, n, "Solver.mass_factor", &cfg->mass_factor, 0);
  rc |= get_double(E, n, "Solver.density_tolerance", &cfg->density_tolerance, 0);
  rc |= get_int(E, n, "Solver.vorticity_model", &cfg->vorticity_model, 0);
  rc |= get_int(E, n, "Solver.surface_tension", &cfg->surface_tension, 0);
  rc |= get_int(E, n, "Solver.viscosity", &cfg->viscosity, 0);
  rc |= get_int(E, n, "Solver.vorticity", &cfg->vorticity, 0);
  rc |= get_double(E, n, "Solver.jitter", &cfg->jitter, 0);
  rc |= get_double(E, n, "Solver.seed", &seed, 0);
  rc |= get_int(E, n, "Solver.output_every", &cfg->output_every, 0);
  cfg->warm_start_net[0] = '\0';
  const char *net = find(E, n, "Solver.warm_start_net");
  if (net) snprintf(cfg->warm_start_net, sizeof(cfg->warm_start_net), "%s", net);
  cfg->seed = (unsigned long long)seed;
  return rc ? SPH_ERR_ARG : SPH_OK;
}
