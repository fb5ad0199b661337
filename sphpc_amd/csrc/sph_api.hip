// C ABI of libsphpc_amd: context lifecycle, parameters, the per-substep stages
// in the reference's call order (simulation.c:59-108), the fused substep,
// inspection and event timing.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "sph_internal.h"

namespace sph {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }

int hip_fail(hipError_t e, const char *what) {
  g_error = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? SPH_ERR_NOMEM : SPH_ERR_HIP;
}

static int arg_error(const char *msg) {
  g_error = msg;
  return SPH_ERR_ARG;
}

static int state_error(const char *msg) {
  g_error = msg;
  return SPH_ERR_STATE;
}

void prof_begin(sph_ctx *c, int kernel, hipEvent_t *a) {
  (void)kernel;
  *a = nullptr;
  if (!c || !c->prof.on) return;
  hipEvent_t e;
  if (!c->prof.pool.empty()) {
    e = c->prof.pool.back();
    c->prof.pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    return;
  }
  if (hipEventRecord(e, c->stream) != hipSuccess) {
    c->prof.pool.push_back(e);
    return;
  }
  *a = e;
}

void prof_end(sph_ctx *c, int kernel, hipEvent_t a) {
  if (!c || !c->prof.on || !a) return;
  hipEvent_t e;
  if (!c->prof.pool.empty()) {
    e = c->prof.pool.back();
    c->prof.pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    c->prof.pool.push_back(a);
    return;
  }
  (void)hipEventRecord(e, c->stream);
  c->prof.pending.push_back({a, e, kernel});
}

static void prof_collect(sph_ctx *c) {
  if (c->prof.pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto &p : c->prof.pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      c->prof.total_ms[p.kernel] += ms;
      c->prof.launches[p.kernel] += 1;
    }
    c->prof.pool.push_back(p.a);
    c->prof.pool.push_back(p.b);
  }
  c->prof.pending.clear();
}

static void prof_zero(sph_ctx *c) {
  prof_collect(c);
  for (int k = 0; k < K_COUNT; ++k) {
    c->prof.total_ms[k] = 0.0;
    c->prof.launches[k] = 0;
  }
}

static const char *kKernelNames[K_COUNT] = {
    "predict_hash", "hash", "gravity", "predict", "radix_upsweep", "radix_downsweep",
    "scan", "reorder", "build_lists", "density_lambda", "density", "lambda", "dp",
    "dp_update", "position_stars", "velocity", "color", "surface", "xsph", "vorticity",
    "confine", "commit", "aux", "warm_lambda", "collide", "lambda_copy"};

// Derived FP32 constants (SetParameters, simulation.c:147-177; the constants
// of particles.c named in DevParams).  Computed in double, rounded once.
static void derive(sph_ctx *c) {
  const sph_params &P = c->P;
  DevParams &d = c->dev;
  const double h = (double)(float)P.smoothing_radius;
  const double W_norm = 315.0 / (64.0 * M_PI * pow(h, 9.0));
  const double DelW_poly_norm = -3.0 * 2.0 * W_norm;
  const double DelW_norm = -45.0 / (M_PI * pow(h, 6.0));
  const double C_norm = 32.0 / (M_PI * pow(h, 9.0));
  const double m = P.rest_mass, rho0 = P.rest_density;
  d.h = (float)h;
  d.h2 = (float)(h * h);
  d.inv_h = (float)(1.0 / h);
  d.h2_lo = (float)(h * h * (1.0 - 1e-5));
  d.h2_hi = (float)(h * h * (1.0 + 1e-5));
  d.h2d = h * h;
  d.rmin = (float)(h * 0.0001);
  d.rmin2 = (float)(h * 0.0001 * h * 0.0001);
  d.W_norm = (float)W_norm;
  d.DelW_norm = (float)DelW_norm;
  d.DelW_poly_norm = (float)DelW_poly_norm;
  d.C_norm = (float)C_norm;
  d.mass = (float)m;
  d.rest_density = (float)rho0;
  d.inv_rest_density = (float)(1.0 / rho0);
  d.mW0 = (float)(m * W_norm * pow(h, 6.0));
  d.mWn = (float)(m * W_norm);
  {
    // h is a float, so h * h is exact in double
    d.dens_k3 = pow(2.0, 26.0 / 3.0) / (h * h);
    const double k3 = d.dens_k3;
    d.dens_c1 = k3 * (h * h);
    const double a = m * W_norm / (rho0 * k3 * k3 * k3);
    const double c0 = m * W_norm * pow(h, 6.0) / rho0 - 1.0;
    d.dens_a_hi = (float)a;
    d.dens_a_lo = (float)(a - (double)d.dens_a_hi);
    d.dens_c0_hi = (float)c0;
    d.dens_c0_lo = (float)(c0 - (double)d.dens_c0_hi);
  }
  d.lam_k2 = (float)((m * DelW_norm / rho0) * (m * DelW_norm / rho0));
  d.lambda_eps = (float)P.lambda_epsilon;
  d.dt = (float)P.time_step;
  d.g_dt = (float)(-P.g * P.time_step);
  d.v_max = (float)P.v_max;
  d.xsph_c = (float)P.c;
  d.gamma = (float)P.k;
  d.coh_k = (float)(-20.0 * P.k * m * m * C_norm);
  d.curv_k = (float)(-P.k * m);
  d.h6_64 = (float)(pow(h, 6.0) / 64.0);
  d.vort_eps = (float)P.vorticity_epsilon;
  d.vort_model = P.vorticity_model;
  d.conf_k = (float)(P.vorticity_model == 1 ? P.time_step * P.vorticity_epsilon
                                            : P.time_step * P.vorticity_epsilon / m);
  d.color_k = (float)(h * m * DelW_poly_norm);
  d.div_k = (float)(m / rho0 * DelW_norm);
  const double lo[3] = {c->box.min_x, c->box.min_y, c->box.min_z};
  const double hi[3] = {c->box.max_x, c->box.max_y, c->box.max_z};
  for (int a = 0; a < 3; ++a) {
    d.box_lo[a] = (float)lo[a];
    d.box_hi[a] = (float)hi[a];
    d.box_hi_clamp[a] = (float)(hi[a] - h * 0.0001);
  }
}

static int check_params(const sph_params *p) {
  if (!p) return arg_error("params is NULL");
  const float hf = (float)p->smoothing_radius;
  if (!(p->smoothing_radius > 0.0) || (double)hf != p->smoothing_radius)
    return arg_error("smoothing_radius must be positive and exactly representable in FP32 "
                     "(the reference uses h = 2*float(spacing))");
  if (!(p->time_step > 0.0)) return arg_error("time_step must be positive");
  if (!(p->rest_density > 0.0) || !(p->rest_mass > 0.0))
    return arg_error("rest_density and rest_mass must be positive");
  if (p->solve_iterations < 0 || p->solve_iterations > kMaxIters)
    return arg_error("solve_iterations out of range [0, 128]");
  if (p->vorticity_model != 0 && p->vorticity_model != 1)
    return arg_error("vorticity_model must be 0 or 1");
  return SPH_OK;
}

static void graph_reset(sph_ctx *c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_valid = false;
}

static void free_all(sph_ctx *c) {
  graph_reset(c);
  void *bufs[] = {c->x_can, c->v_can, c->xs_can, c->keys, c->keys_alt, c->ids, c->ids_alt,
                  c->hist, c->cell_end, c->cell_start, c->pos_a, c->pos_b, c->xold,
                  c->vel_a, c->vel_b, c->aux, c->dpbuf, c->rho, c->cstage, c->lam_stage,
                  c->list, c->list_head,
                  c->stats, c->pair_partial, c->red, c->sdf_buf, c->scan.tile_sums,
                  c->pw, c->sx, c->lam_acc, c->lam_prev, c->feat, c->scenes_dev, c->scene_of,
                  c->batch_sdf, c->scene_err, c->scene_err0, c->scene_chunks,
                  c->warm_net};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (c->io_dev) (void)hipFree(c->io_dev);
  if (c->io_host) (void)hipHostFree(c->io_host);
  for (auto &p : c->prof.pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->prof.pool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
}

static int require_state(sph_ctx *c) {
  if (!c) return arg_error("context is NULL");
  c->gen++;
  if (!c->state_valid) return state_error("no particle state uploaded");
  return SPH_OK;
}

template <class T>
static int fetch(sph_ctx *c, const T *dev, size_t count, std::vector<T> *out) {
  out->resize(count);
  if (count)
    SPH_HIP(hipMemcpyAsync(out->data(), dev, sizeof(T) * count, hipMemcpyDeviceToHost,
                           c->stream));
  SPH_HIP(hipStreamSynchronize(c->stream));
  return SPH_OK;
}

static int require_lists(sph_ctx *c) {
  int rc = require_state(c);
  if (rc) return rc;
  if (!c->lists_valid) return state_error("neighbour lists missing: call sph_find_all_neighbors");
  return SPH_OK;
}

}  // namespace sph

using namespace sph;

extern "C" const char *sph_last_error(void) { return g_error.c_str(); }

extern "C" int sph_get_device_count(int *count) {
  if (!count) return arg_error("count is NULL");
  SPH_HIP(hipGetDeviceCount(count));
  return SPH_OK;
}

extern "C" void sph_default_params(sph_params *p, double h) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->smoothing_radius = h;
  p->time_step = 1.0 / 300.0;
  p->g = 9.8;
  p->k = 0.0;
  p->c = 0.0;
  p->rest_density = 1000.0;
  p->rest_mass = 0.0;
  p->lambda_epsilon = 0.01;         // particles.c:663
  p->vorticity_epsilon = 0.1 * h;   // particles.c:289
  p->v_max = 100.0;                 // particles.c:791
  p->stiffness_k = 0.95;            // particles.c:693
  p->density_tolerance = 0.0;
  p->solve_iterations = 10;         // simulation.c:72
  p->surface_tension = 1;
  p->viscosity = 1;
  p->vorticity = 1;
  p->vorticity_model = 0;
}

// Device buffers, stream and kernel set-up of a context whose grid (cells,
// key_bits) and capacity are known.  Frees the context on failure.
static int ctx_allocate(sph_ctx *c, int max_particles, sph_ctx **out) {
  auto fail = [&](int code) {
    free_all(c);
    delete c;
    return code;
  };
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) return fail(hip_fail(e, "hipGetDevice"));
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return fail(hip_fail(e, "hipStreamCreate"));

  const size_t cap = (size_t)std::max(max_particles, 1);
  c->cap = max_particles;
  const size_t f4 = sizeof(float4) * cap, u4 = sizeof(uint32_t) * cap;
  const size_t groups = (cap + kWave - 1) / kWave;
  struct {
    void **p;
    size_t bytes;
  } allocs[] = {
      {(void **)&c->x_can, f4},       {(void **)&c->v_can, f4},
      {(void **)&c->xs_can, f4},      {(void **)&c->keys, u4},
      {(void **)&c->keys_alt, u4},    {(void **)&c->ids, u4},
      {(void **)&c->ids_alt, u4},     {(void **)&c->cell_end, sizeof(uint32_t) * (c->cells + 1ull)},
      {(void **)&c->cell_start, sizeof(uint32_t) * (c->cells + 1ull)},
      {(void **)&c->pos_a, f4 + sizeof(float4) * kPosPad},
      {(void **)&c->pos_b, f4 + sizeof(float4) * kPosPad},
      {(void **)&c->xold, f4},        {(void **)&c->vel_a, f4},
      {(void **)&c->vel_b, f4},       {(void **)&c->aux, f4},
      {(void **)&c->dpbuf, f4},       {(void **)&c->rho, sizeof(float) * cap},
      {(void **)&c->cstage, sizeof(float) * cap},
      {(void **)&c->lam_stage, sizeof(float) * cap},
      {(void **)&c->list, sizeof(uint32_t) * groups * kListStride},
      {(void **)&c->list_head, sizeof(uint4) * cap},
      {(void **)&c->pw, f4},
      {(void **)&c->sx, sizeof(float) * 3 * (cap + kPosPad)},
      {(void **)&c->stats, sizeof(uint32_t) * STAT_WORDS},
      {(void **)&c->pair_partial, sizeof(uint32_t) * blocks_for((uint32_t)cap, kWave)},
      {(void **)&c->red, sizeof(double) * std::max<size_t>(4, blocks_for((uint32_t)cap))},
  };
  for (auto &a : allocs) {
    e = hipMalloc(a.p, a.bytes);
    if (e != hipSuccess) return fail(hip_fail(e, "hipMalloc (context buffers)"));
  }
  e = hipMemsetAsync(c->cell_end, 0, sizeof(uint32_t) * (c->cells + 1ull), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->stats, 0, sizeof(uint32_t) * STAT_WORDS, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->list_head, 0, sizeof(uint4) * cap, c->stream);
  // C = 0 until a density stage runs (lambda from it is zero, as the
  // reference's from its zero-initialised density array)
  if (e == hipSuccess) e = hipMemsetAsync(c->cstage, 0, sizeof(float) * cap, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->pos_a, 0, f4 + sizeof(float4) * kPosPad, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->pos_b, 0, f4 + sizeof(float4) * kPosPad, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->sx, 0, sizeof(float) * 3 * (cap + kPosPad), c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(hip_fail(e, "context initialisation"));
  c->pos = c->pos_a;
  c->soa_stride = (uint32_t)(cap + kPosPad);
  c->vel = c->vel_a;
  *out = c;
  return SPH_OK;
}

extern "C" int sph_create(const sph_params *params, const sph_aabb *boundary, int max_particles,
                          sph_ctx **out) {
  if (!out) return arg_error("out is NULL");
  *out = nullptr;
  int rc = check_params(params);
  if (rc) return rc;
  if (!boundary) return arg_error("boundary is NULL");
  if (boundary->min_x != 0.0 || boundary->min_y != 0.0 || boundary->min_z != 0.0)
    return arg_error("boundary must start at the origin (the reference hashes absolute "
                     "coordinates, neighbors.c:86-95)");
  if (!(boundary->max_x > 0.0 && boundary->max_y > 0.0 && boundary->max_z > 0.0))
    return arg_error("boundary extent must be positive");
  if (max_particles < 0) return arg_error("max_particles must be >= 0");
  if ((uint32_t)max_particles > kMaxParticles)
    return arg_error("max_particles exceeds 2^28 - 64 (particle arrays are addressed with "
                     "32-bit byte offsets)");

  sph_ctx *c = new sph_ctx();
  c->P = *params;
  c->box = *boundary;
  const double h = params->smoothing_radius;
  // AllocInitNeighbors, neighbors.c:23-28: ceil(L/h) + 1 cells per axis
  const double nxd = ceil((boundary->max_x - boundary->min_x) / h) + 1;
  const double nyd = ceil((boundary->max_y - boundary->min_y) / h) + 1;
  const double nzd = ceil((boundary->max_z - boundary->min_z) / h) + 1;
  if (nxd * nyd * nzd >= 4294967295.0) {
    delete c;
    return arg_error("hash grid exceeds 2^32 cells");
  }
  c->dev.nx = (uint32_t)nxd;
  c->dev.ny = (uint32_t)nyd;
  c->dev.nz = (uint32_t)nzd;
  c->dev.nxy = c->dev.nx * c->dev.ny;
  c->cells = c->dev.nxy * c->dev.nz;
  c->key_bits = 1;
  while (c->key_bits < 32 && ((uint64_t)1 << c->key_bits) < (uint64_t)c->cells) c->key_bits++;
  derive(c);
  return ctx_allocate(c, max_particles, out);
}

extern "C" int sph_destroy(sph_ctx *c) {
  if (!c) return SPH_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_all(c);
  delete c;
  return SPH_OK;
}

static int sdf_fields(const sph_obstacle_desc *o, SdfDev *d);

// Batched independent scenes in one context: one particle array, one sort and
// one launch per stage for all scenes, each with its own box, grid (a
// contiguous range of cell keys) and obstacle; scene k's particles are
// indices [first_k, first_k + particles_k) in upload order and in the sorted
// working order alike.
extern "C" int sph_create_batch(const sph_params *params, int nscenes, const sph_scene *scenes,
                                sph_ctx **out) {
  if (!out) return arg_error("out is NULL");
  *out = nullptr;
  int rc = check_params(params);
  if (rc) return rc;
  if (nscenes < 1 || nscenes > 65535 || !scenes)
    return arg_error("need 1..65535 scenes");
  const double h = params->smoothing_radius;
  std::vector<SceneDev> hs((size_t)nscenes);
  std::vector<uint32_t> first((size_t)nscenes + 1, 0u);
  uint64_t cells = 0, total = 0;
  size_t sdf_words = 0;
  double ext[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < nscenes; ++k) {
    const sph_scene &sc = scenes[k];
    const double L[3] = {sc.size_x, sc.size_y, sc.size_z};
    if (sc.particles < 0) return arg_error("scene particle count must be >= 0");
    if (!(L[0] > 0.0 && L[1] > 0.0 && L[2] > 0.0)) return arg_error("scene size must be positive");
    SceneDev &d = hs[k];
    memset(&d, 0, sizeof(d));
    // AllocInitNeighbors, neighbors.c:23-28: ceil(L/h) + 1 cells per axis
    d.nx = (uint32_t)(ceil(L[0] / h) + 1);
    d.ny = (uint32_t)(ceil(L[1] / h) + 1);
    d.nz = (uint32_t)(ceil(L[2] / h) + 1);
    d.nxy = d.nx * d.ny;
    d.cell_base = (uint32_t)cells;
    cells += (uint64_t)d.nxy * d.nz;
    for (int a = 0; a < 3; ++a) {
      d.box_hi[a] = (float)L[a];
      d.box_hi_clamp[a] = (float)(L[a] - h * 0.0001);
      ext[a] = std::max(ext[a], L[a]);
    }
    if (sc.obstacle.sdf) {
      if ((rc = sdf_fields(&sc.obstacle, &d.sdf))) return rc;
      sdf_words += (size_t)sc.obstacle.dim_x * sc.obstacle.dim_y * sc.obstacle.dim_z;
    }
    total += (uint64_t)sc.particles;
    first[k + 1] = (uint32_t)std::min<uint64_t>(total, 0xffffffffu);
    if (cells >= 0xffffffffull) return arg_error("hash grids exceed 2^32 cells");
  }
  if (total > kMaxParticles)
    return arg_error("too many particles: the scenes' total exceeds 2^28 - 64 (particle "
                     "arrays are addressed with 32-bit byte offsets)");
  sph_ctx *c = new sph_ctx();
  c->P = *params;
  c->box = sph_aabb{0.0, ext[0], 0.0, ext[1], 0.0, ext[2]};
  c->cells = (uint32_t)cells;
  c->key_bits = 1;
  while (c->key_bits < 32 && ((uint64_t)1 << c->key_bits) < cells) c->key_bits++;
  c->dev.nx = c->dev.ny = c->dev.nz = c->dev.nxy = 1;  // per scene instead
  derive(c);
  rc = ctx_allocate(c, (int)total, out);
  if (rc) return rc;
  // scene table, per-particle scene index, obstacle grids
  const size_t np = std::max<size_t>((size_t)total, 1);
  std::vector<uint16_t> owner(np, 0);
  std::vector<uint4> chunks;
  for (int k = 0; k < nscenes; ++k) {
    for (uint32_t i = first[k]; i < first[k + 1]; ++i) owner[i] = (uint16_t)k;
    for (uint32_t i = first[k]; i < first[k + 1]; i += kSceneChunk)
      chunks.push_back(make_uint4((uint32_t)k, i, std::min(i + kSceneChunk, first[k + 1]), 0u));
  }
  c->nchunks = (uint32_t)chunks.size();
  hipError_t e = hipMalloc((void **)&c->scenes_dev, sizeof(SceneDev) * nscenes);
  if (e == hipSuccess) e = hipMalloc((void **)&c->scene_of, sizeof(uint16_t) * np);
  if (e == hipSuccess) e = hipMalloc((void **)&c->scene_err, sizeof(uint32_t) * nscenes);
  if (e == hipSuccess) e = hipMalloc((void **)&c->scene_err0, sizeof(uint32_t) * nscenes);
  if (e == hipSuccess && c->nchunks)
    e = hipMalloc((void **)&c->scene_chunks, sizeof(uint4) * c->nchunks);
  if (e == hipSuccess && c->nchunks)
    e = hipMemcpyAsync(c->scene_chunks, chunks.data(), sizeof(uint4) * c->nchunks,
                       hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && sdf_words) e = hipMalloc((void **)&c->batch_sdf, sizeof(float) * sdf_words);
  size_t off = 0;
  for (int k = 0; k < nscenes && e == hipSuccess; ++k) {
    const sph_obstacle_desc &o = scenes[k].obstacle;
    if (!o.sdf) continue;
    const size_t cnt = (size_t)o.dim_x * o.dim_y * o.dim_z;
    hs[k].sdf.data = c->batch_sdf + off;
    e = hipMemcpyAsync(c->batch_sdf + off, o.sdf, sizeof(float) * cnt, hipMemcpyHostToDevice,
                       c->stream);
    off += cnt;
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->scenes_dev, hs.data(), sizeof(SceneDev) * nscenes,
                       hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->scene_of, owner.data(), sizeof(uint16_t) * np, hipMemcpyHostToDevice,
                       c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    const int code = hip_fail(e, "batch set-up");
    sph_ctx *done = *out;
    *out = nullptr;
    free_all(done);
    delete done;
    return code;
  }
  c->nscenes = nscenes;
  c->scene_first = first;
  return SPH_OK;
}

extern "C" int sph_batch_scene_count(sph_ctx *c, int *nscenes) {
  if (!c || !nscenes) return arg_error("NULL argument");
  *nscenes = c->nscenes;
  return SPH_OK;
}

extern "C" int sph_batch_scene_errors(sph_ctx *c, float *max_err) {
  if (!c || !max_err) return arg_error("NULL argument");
  if (!c->nscenes) return state_error("not a batched context");
  int rc = require_lists(c);
  if (rc) return rc;
  rc = launch_scene_errors(c);
  if (rc) return rc;
  std::vector<uint32_t> w((size_t)c->nscenes);
  SPH_HIP(hipMemcpyAsync(w.data(), c->scene_err, sizeof(uint32_t) * w.size(),
                         hipMemcpyDeviceToHost, c->stream));
  SPH_HIP(hipStreamSynchronize(c->stream));
  for (int k = 0; k < c->nscenes; ++k) memcpy(&max_err[k], &w[k], sizeof(float));
  return SPH_OK;
}

extern "C" int sph_set_params(sph_ctx *c, const sph_params *p) {
  if (c) c->gen++;
  if (!c) return arg_error("context is NULL");
  int rc = check_params(p);
  if (rc) return rc;
  if (p->smoothing_radius != c->P.smoothing_radius)
    return arg_error("smoothing_radius fixes the hash grid; create a new context to change it");
  graph_reset(c);
  c->P = *p;
  derive(c);
  return SPH_OK;
}

extern "C" int sph_get_params(sph_ctx *c, sph_params *p) {
  if (!c || !p) return arg_error("NULL argument");
  *p = c->P;
  return SPH_OK;
}

extern "C" int sph_grid_dims(sph_ctx *c, int *nx, int *ny, int *nz) {
  if (!c) return arg_error("context is NULL");
  if (nx) *nx = (int)c->dev.nx;
  if (ny) *ny = (int)c->dev.ny;
  if (nz) *nz = (int)c->dev.nz;
  return SPH_OK;
}

// Fields of an obstacle grid (AllocInitObstacle_CUDA, obstacle_cuda.cu:
// 133-183) except the device pointer.
static int sdf_fields(const sph_obstacle_desc *o, SdfDev *d) {
  if (o->dim_x < 1 || o->dim_y < 1 || o->dim_z < 1 || !(o->spacing > 0.0))
    return arg_error("obstacle dims and spacing must be positive");
  const sph_aabb &b = o->world_bounds;
  if (!(b.max_x > b.min_x && b.max_y > b.min_y && b.max_z > b.min_z))
    return arg_error("obstacle world bounds must be non-empty");
  d->dim[0] = o->dim_x;
  d->dim[1] = o->dim_y;
  d->dim[2] = o->dim_z;
  const double lo[3] = {b.min_x, b.min_y, b.min_z}, hi[3] = {b.max_x, b.max_y, b.max_z};
  for (int a = 0; a < 3; ++a) {
    d->wmin[a] = (float)lo[a];
    d->wmax[a] = (float)hi[a];
    d->len[a] = (float)(hi[a] - lo[a]);
    d->inv_dim[a] = (float)(1.0 / (double)d->dim[a]);
  }
  const float len_model = (float)(o->dim_x * o->spacing);  // obstacle_cuda.cu:101
  d->m2w = d->len[0] / len_model;                          // obstacle_cuda.cu:105
  return SPH_OK;
}

static int make_sdf(const sph_obstacle_desc *o, SdfDev *d, float **buf, hipStream_t s) {
  int rc = sdf_fields(o, d);
  if (rc) return rc;
  const size_t count = (size_t)o->dim_x * (size_t)o->dim_y * (size_t)o->dim_z;
  SPH_HIP(hipMalloc(buf, sizeof(float) * count));
  SPH_HIP(hipMemcpyAsync(*buf, o->sdf, sizeof(float) * count, hipMemcpyHostToDevice, s));
  SPH_HIP(hipStreamSynchronize(s));
  d->data = *buf;
  return SPH_OK;
}

extern "C" int sph_set_obstacle(sph_ctx *c, const sph_obstacle_desc *o) {
  if (c) c->gen++;
  if (!c) return arg_error("context is NULL");
  if (c->nscenes) return state_error("a batched context takes its obstacles from sph_create_batch");
  graph_reset(c);
  if (c->sdf_buf) {
    SPH_HIP(hipStreamSynchronize(c->stream));
    SPH_HIP(hipFree(c->sdf_buf));
    c->sdf_buf = nullptr;
  }
  memset(&c->sdf, 0, sizeof(c->sdf));
  if (!o || !o->sdf) return SPH_OK;
  int rc = make_sdf(o, &c->sdf, &c->sdf_buf, c->stream);
  if (rc) {
    memset(&c->sdf, 0, sizeof(c->sdf));
    if (c->sdf_buf) (void)hipFree(c->sdf_buf);
    c->sdf_buf = nullptr;
  }
  return rc;
}

// SoA <-> the float4 canonical state, on the device (coalesced on both sides).
__global__ void k_pack_state(uint32_t n, const float *__restrict__ soa,
                             float4 *__restrict__ x, float4 *__restrict__ v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = make_float4(soa[i], soa[n + i], soa[2 * (size_t)n + i], 0.0f);
  v[i] = make_float4(soa[3 * (size_t)n + i], soa[4 * (size_t)n + i], soa[5 * (size_t)n + i],
                     0.0f);
}

__global__ void k_unpack_state(uint32_t n, const float4 *__restrict__ x,
                               const float4 *__restrict__ v, float *__restrict__ soa,
                               int with_velocity) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 a = x[i];
  soa[i] = a.x;
  soa[n + i] = a.y;
  soa[2 * (size_t)n + i] = a.z;
  if (with_velocity) {
    const float4 b = v[i];
    soa[3 * (size_t)n + i] = b.x;
    soa[4 * (size_t)n + i] = b.y;
    soa[5 * (size_t)n + i] = b.z;
  }
}

// Device and pinned host staging for state transfers (6 floats per particle of
// the capacity), allocated on first use.
static int io_buffers(sph_ctx *c) {
  if (c->io_dev && c->io_host) return SPH_OK;
  const size_t bytes = sizeof(float) * 6 * (size_t)std::max(c->cap, 1);
  if (!c->io_dev) {
    hipError_t e = hipMalloc(&c->io_dev, bytes);
    if (e != hipSuccess) {
      c->io_dev = nullptr;
      return hip_fail(e, "hipMalloc (state transfer staging)");
    }
  }
  hipError_t e = hipHostMalloc(&c->io_host, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    c->io_host = nullptr;
    return hip_fail(e, "hipHostMalloc (state transfer staging)");
  }
  return SPH_OK;
}

extern "C" int sph_upload_state(sph_ctx *c, int n, const float *x, const float *y,
                                const float *z, const float *vx, const float *vy,
                                const float *vz) {
  if (c) c->gen++;
  if (!c) return arg_error("context is NULL");
  if (n < 0) return arg_error("n must be >= 0");
  if (n > c->cap) {
    set_error("particle count exceeds the context capacity");
    return SPH_ERR_CAPACITY;
  }
  if (n > 0 && (!x || !y || !z || !vx || !vy || !vz)) return arg_error("NULL state array");
  if (c->nscenes && n != c->cap)
    return arg_error("a batched context takes all scenes' particles at once, in scene order");
  if (n != c->n) graph_reset(c);
  SPH_HIP(hipStreamSynchronize(c->stream));
  if (n) {
    int rc = io_buffers(c);
    if (rc) return rc;
    const size_t un = (size_t)n;
    const float *src[6] = {x, y, z, vx, vy, vz};
    for (int k = 0; k < 6; ++k) memcpy(c->io_host + k * un, src[k], sizeof(float) * un);
    SPH_HIP(hipMemcpyAsync(c->io_dev, c->io_host, sizeof(float) * 6 * un,
                           hipMemcpyHostToDevice, c->stream));
    k_pack_state<<<blocks_for((uint32_t)n), kBlock, 0, c->stream>>>((uint32_t)n, c->io_dev,
                                                                    c->x_can, c->v_can);
    SPH_HIP(hipGetLastError());
  }
  if (c->lam_prev) SPH_HIP(hipMemsetAsync(c->lam_prev, 0, sizeof(float) * c->cap, c->stream));
  SPH_HIP(hipStreamSynchronize(c->stream));
  c->n = n;
  c->state_valid = true;
  c->sorted_valid = false;
  c->lists_valid = false;
  c->feat_substep = -1;
  return SPH_OK;
}

extern "C" int sph_download_state(sph_ctx *c, float *x, float *y, float *z, float *vx,
                                  float *vy, float *vz) {
  int rc = require_state(c);
  if (rc) return rc;
  const int n = c->n;
  float *dst[6] = {x, y, z, vx, vy, vz};
  const bool with_v = vx || vy || vz;
  if (n && (x || y || z || with_v)) {
    if ((rc = io_buffers(c))) return rc;
    const size_t un = (size_t)n;
    k_unpack_state<<<blocks_for((uint32_t)n), kBlock, 0, c->stream>>>(
        (uint32_t)n, c->x_can, c->v_can, c->io_dev, with_v ? 1 : 0);
    SPH_HIP(hipGetLastError());
    SPH_HIP(hipMemcpyAsync(c->io_host, c->io_dev, sizeof(float) * (with_v ? 6 : 3) * un,
                           hipMemcpyDeviceToHost, c->stream));
    SPH_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 6; ++k)
      if (dst[k]) memcpy(dst[k], c->io_host + k * un, sizeof(float) * un);
    return SPH_OK;
  }
  SPH_HIP(hipStreamSynchronize(c->stream));
  return SPH_OK;
}

extern "C" int sph_particle_count(sph_ctx *c, int *n) {
  if (!c || !n) return arg_error("NULL argument");
  *n = c->n;
  return SPH_OK;
}

// --------------------------------------------------------------------------
// Stages in the driver's order
// --------------------------------------------------------------------------
extern "C" int sph_apply_gravity(sph_ctx *c) {
  int rc = require_state(c);
  return rc ? rc : launch_gravity(c);
}

extern "C" int sph_predict_positions(sph_ctx *c) {
  int rc = require_state(c);
  return rc ? rc : launch_predict(c);
}

extern "C" int sph_find_all_neighbors(sph_ctx *c) {
  int rc = require_state(c);
  if (rc) return rc;
  rc = launch_substep_init(c);
  if (rc) return rc;
  c->staged_slot = 0;
  return launch_neighbor_search(c, false);
}

extern "C" int sph_compute_densities(sph_ctx *c) {
  int rc = require_lists(c);
  if (rc) return rc;
  const int slot = std::min(c->staged_slot, kMaxIters);
  c->staged_slot++;
  c->last_err_slot = slot;
  // density_to_w keeps rho in pos.w for the velocity passes.  With a fixed
  // iteration count the same gather also computes lambda and keeps it aside
  // for an immediately following sph_compute_lambda.
  if (c->P.density_tolerance > 0.0) return launch_density(c, 3, slot);
  if ((rc = launch_density(c, 4, slot))) return rc;
  c->lam_gen = c->gen;
  return SPH_OK;
}

extern "C" int sph_compute_lambda(sph_ctx *c) {
  // lambda from the density pass just before, when no call came in between
  // (the same values: one gather computes both from the same positions)
  const bool fresh = c && c->lam_gen == c->gen;
  int rc = require_lists(c);
  if (rc) return rc;
  return fresh ? launch_lambda_from_density(c) : launch_density(c, 2, -1);
}

extern "C" int sph_update_dps(sph_ctx *c, int substep) {
  int rc = require_lists(c);
  if (rc) return rc;
  if (substep < 0) return arg_error("substep must be >= 0");
  // the same pass also writes x* + delta p (clamped, pushed out of the
  // obstacle) into the other position buffer, for an immediately following
  // sph_update_position_stars; x* itself is unchanged until then
  float4 *next = (c->pos == c->pos_a) ? c->pos_b : c->pos_a;
  if ((rc = launch_dp(c, substep, next))) return rc;
  c->dp_next = next;
  c->dp_gen = c->gen;
  return SPH_OK;
}

extern "C" int sph_update_position_stars(sph_ctx *c) {
  const bool fresh = c && c->dp_gen == c->gen;
  int rc = require_lists(c);
  if (rc) return rc;
  if (fresh) {
    c->pos = c->dp_next;
    return SPH_OK;
  }
  return launch_position_stars(c);
}

extern "C" int sph_update_velocities(sph_ctx *c) {
  int rc = require_lists(c);
  return rc ? rc : launch_velocity(c);
}

extern "C" int sph_compute_surface_tension(sph_ctx *c) {
  int rc = require_lists(c);
  if (rc) return rc;
  rc = launch_color(c, false);
  return rc ? rc : launch_surface(c);
}

extern "C" int sph_apply_viscosity(sph_ctx *c) {
  int rc = require_lists(c);
  return rc ? rc : launch_xsph(c);
}

extern "C" int sph_compute_vorticity(sph_ctx *c) {
  int rc = require_lists(c);
  return rc ? rc : launch_vorticity(c);
}

extern "C" int sph_apply_vorticity_confinement(sph_ctx *c) {
  int rc = require_lists(c);
  return rc ? rc : launch_confine(c, false);
}

extern "C" int sph_update_positions(sph_ctx *c) {
  int rc = require_lists(c);
  if (rc) return rc;
  rc = launch_commit(c);
  if (rc) return rc;
  c->substeps++;
  return SPH_OK;
}

// --------------------------------------------------------------------------
// Fused substep
// --------------------------------------------------------------------------
// Device work of one fused substep, all on the context stream (capturable).
static int substep_launch(sph_ctx *c, bool *warm_out) {
  int rc = launch_substep_init(c);
  if (rc) return rc;
  rc = launch_neighbor_search(c, true);
  if (rc) return rc;
  const int iters = c->P.solve_iterations;
  // Logged or warm-started solve: the first iteration's lambda pass also
  // computes the network features (and, warm, replaces lambda by the
  // network's accumulated multiplier); later iterations are the plain ones.
  const bool prev = c->warm_prev_scale > 0.0f && !c->log_on;
  const bool warm = iters > 0 && (c->log_on || c->warm_on || prev);
  *warm_out = warm;
  int it0 = 0;
  const float4 *pre_warm = nullptr;
  const bool fallback = warm && !c->log_on && iters > 1;
  if (warm) {
    if ((rc = prev ? launch_warm_prev(c) : launch_warm_lambda(c, c->log_on))) return rc;
    // a batch decides the fallback per scene: errors before the step
    if (fallback && c->nscenes && (rc = launch_scene_errors(c, c->scene_err0))) return rc;
    pre_warm = c->pos;
    if ((rc = launch_dp_update(c, 0, 0, true))) return rc;
    it0 = 1;
  }
  for (int it = it0; it < iters; ++it) {
    if ((rc = launch_density(c, 0, it))) return rc;
    if (it == 1 && fallback) {
      if (c->nscenes && (rc = launch_scene_errors(c))) return rc;
      if ((rc = launch_warm_fallback(c, pre_warm))) return rc;
    }
    if ((rc = launch_dp_update(c, it, it, warm))) return rc;
  }
  c->last_err_slot = iters > 0 ? iters - 1 : -1;
  if (c->P.density_tolerance > 0.0) {
    // density after the last allowed update; skipped when already converged
    if ((rc = launch_density(c, 3, iters))) return rc;
    c->last_err_slot = iters;
  }
  if (c->P.surface_tension) {
    if ((rc = launch_color(c, true))) return rc;
    if ((rc = launch_surface(c))) return rc;
  } else {
    if ((rc = launch_velocity(c))) return rc;
  }
  if (c->P.viscosity && (rc = launch_xsph(c))) return rc;
  if (c->P.vorticity) {
    if ((rc = launch_vorticity(c))) return rc;
    if ((rc = launch_confine(c, true))) return rc;
  } else {
    if ((rc = launch_commit(c))) return rc;
  }
  if (warm && (rc = launch_scatter_lambda(c))) return rc;
  return SPH_OK;
}

static void substep_finish(sph_ctx *c, bool warm) {
  c->substeps++;
  if (warm && c->log_on) c->feat_substep = c->substeps;
}

static int substep_fused(sph_ctx *c) {
  bool warm = false;
  int rc = substep_launch(c, &warm);
  if (rc) return rc;
  substep_finish(c, warm);
  return SPH_OK;
}

// Records one substep into a graph without running it.  Every lazily sized
// buffer is allocated by then (an eager substep of the same size ran first),
// so the capture contains kernel launches and no allocation or memset.
static int capture_substep(sph_ctx *c) {
  graph_reset(c);
  SPH_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  bool warm = false;
  const int rc = substep_launch(c, &warm);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return rc ? rc : hip_fail(e, "hipStreamEndCapture");
  }
  const hipError_t ei = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    c->graph_exec = nullptr;
    return hip_fail(ei, "hipGraphInstantiate");
  }
  c->graph = g;
  c->graph_after = {c->pos, c->vel, c->key_s, c->id_s, c->pair_blocks,
                    c->last_err_slot, warm};
  c->graph_valid = true;
  return SPH_OK;
}

static int substep_graph(sph_ctx *c) {
  SPH_HIP(hipGraphLaunch(c->graph_exec, c->stream));
  const auto &a = c->graph_after;
  c->pos = a.pos;
  c->vel = a.vel;
  c->key_s = a.key_s;
  c->id_s = a.id_s;
  c->pair_blocks = a.pair_blocks;
  c->last_err_slot = a.last_err_slot;
  c->sorted_valid = c->lists_valid = true;
  substep_finish(c, a.warm);
  return SPH_OK;
}

// Buffers of the warm start / logging path, allocated on first use.
static int ensure_warm_buffers(sph_ctx *c, bool features) {
  const size_t cap = (size_t)std::max(c->cap, 1);
  if (!c->lam_acc) SPH_HIP(hipMalloc((void **)&c->lam_acc, sizeof(float) * cap));
  if (!c->lam_prev) {
    SPH_HIP(hipMalloc((void **)&c->lam_prev, sizeof(float) * cap));
    SPH_HIP(hipMemsetAsync(c->lam_prev, 0, sizeof(float) * cap, c->stream));
  }
  if (features && !c->feat)
    SPH_HIP(hipMalloc((void **)&c->feat, sizeof(float) * kMlpPad * cap));
  SPH_HIP(hipStreamSynchronize(c->stream));
  return SPH_OK;
}

extern "C" int sph_set_warm_start(sph_ctx *c, sph_mlp *net) {
  if (!c) return arg_error("context is NULL");
  graph_reset(c);
  if (!net) {
    c->warm_on = false;
    return SPH_OK;
  }
  c->warm_prev_scale = 0.0f;
  int rc = ensure_warm_buffers(c, false);
  if (rc) return rc;
  if (!c->warm_net) SPH_HIP(hipMalloc((void **)&c->warm_net, sizeof(MlpPacked)));
  rc = mlp_pack(net, c->warm_net, c->stream);
  if (rc) return rc;
  SPH_HIP(hipStreamSynchronize(c->stream));
  c->warm_on = true;
  return SPH_OK;
}

extern "C" int sph_set_warm_start_previous(sph_ctx *c, float scale) {
  if (!c) return arg_error("context is NULL");
  if (!(scale >= 0.0f) || scale > 4.0f) return arg_error("scale must be in [0, 4]");
  graph_reset(c);
  if (scale > 0.0f) {
    int rc = ensure_warm_buffers(c, false);
    if (rc) return rc;
    c->warm_on = false;
  }
  c->warm_prev_scale = scale;
  return SPH_OK;
}

extern "C" int sph_set_sample_logging(sph_ctx *c, int enable) {
  if (!c) return arg_error("context is NULL");
  graph_reset(c);
  if (enable) {
    int rc = ensure_warm_buffers(c, true);
    if (rc) return rc;
  }
  c->log_on = enable != 0;
  return SPH_OK;
}

extern "C" int sph_warm_samples(sph_ctx *c, float *x, float *t, size_t offset, size_t capacity,
                                int every, size_t *written) {
  if (!c || !x || !t) return arg_error("NULL argument");
  if (every < 1) return arg_error("every must be >= 1");
  if (written) *written = 0;
  if (c->feat_substep < 0 || c->feat_substep != c->substeps)
    return state_error("the last substep was not logged (sph_set_sample_logging)");
  if (offset > capacity) return arg_error("offset beyond capacity");
  if ((((uintptr_t)x) & 15u) != 0) return arg_error("feature rows must be 16-byte aligned");
  size_t count = ((size_t)c->n + (size_t)every - 1) / (size_t)every;
  count = std::min(count, capacity - offset);
  int rc = launch_warm_samples(c, x + offset * kMlpPad, t + offset, (uint32_t)count, every);
  if (rc) return rc;
  SPH_HIP(hipStreamSynchronize(c->stream));
  if (written) *written = count;
  return SPH_OK;
}

extern "C" int sph_step(sph_ctx *c, int substeps) {
  int rc = require_state(c);
  if (rc) return rc;
  if (substeps < 0) return arg_error("substeps must be >= 0");
  for (int s = 0; s < substeps; ++s) {
    if (c->graph_on && !c->prof.on && c->graph_valid) {
      rc = substep_graph(c);
    } else {
      rc = substep_fused(c);
      if (!rc && c->graph_on && !c->prof.on && c->n > 0) rc = capture_substep(c);
    }
    if (rc) return rc;
  }
  return SPH_OK;
}

extern "C" int sph_set_graph(sph_ctx *c, int enable) {
  if (!c) return arg_error("context is NULL");
  graph_reset(c);
  c->graph_on = enable != 0;
  return SPH_OK;
}

extern "C" int sph_synchronize(sph_ctx *c) {
  if (!c) return arg_error("context is NULL");
  SPH_HIP(hipStreamSynchronize(c->stream));
  return SPH_OK;
}

extern "C" int sph_get_stats(sph_ctx *c, sph_stats *st) {
  if (!c || !st) return arg_error("NULL argument");
  uint32_t w[STAT_WORDS];
  SPH_HIP(hipMemcpyAsync(w, c->stats, sizeof(w), hipMemcpyDeviceToHost, c->stream));
  SPH_HIP(hipStreamSynchronize(c->stream));
  memset(st, 0, sizeof(*st));
  st->particles = c->n;
  st->iterations = (int)w[STAT_ITERS];
  st->converged = (int)w[STAT_CONVERGED];
  st->warm_fallback = (int)w[STAT_FALLBACK];
  int slot = c->last_err_slot;
  if (c->P.density_tolerance > 0.0 && st->converged) slot = st->iterations;
  if (slot >= 0 && slot <= kMaxIters) {
    float f;
    memcpy(&f, &w[STAT_ERR0 + slot], sizeof(f));
    st->max_density_error = f;
  }
  st->neighbor_overflow = (unsigned long long)w[STAT_DROPPED] |
                          ((unsigned long long)w[STAT_DROPPED + 1] << 32);
  st->cell_clamps =
      (unsigned long long)w[STAT_CLAMPS] | ((unsigned long long)w[STAT_CLAMPS + 1] << 32);
  st->substeps = c->substeps;
  if (c->pair_blocks) {
    std::vector<uint32_t> part;
    int rc = fetch(c, (const uint32_t *)c->pair_partial, c->pair_blocks, &part);
    if (rc) return rc;
    unsigned long long sum = 0;
    for (uint32_t v : part) sum += v;
    st->neighbor_pairs = sum;
  }
  return SPH_OK;
}

extern "C" int sph_get_residuals(sph_ctx *c, float *err, int cap, int *count) {
  if (!c || !err || !count) return arg_error("NULL argument");
  uint32_t w[STAT_WORDS];
  SPH_HIP(hipMemcpyAsync(w, c->stats, sizeof(w), hipMemcpyDeviceToHost, c->stream));
  SPH_HIP(hipStreamSynchronize(c->stream));
  // one density pass before each applied update; with a tolerance, one more
  // (the pass that met it, or the final pass after the capped loop)
  const int applied = (int)w[STAT_ITERS];
  int passes = c->P.density_tolerance > 0.0 ? applied + 1 : applied;
  passes = std::max(0, std::min(passes, std::min(cap, kMaxIters + 1)));
  for (int k = 0; k < passes; ++k) memcpy(&err[k], &w[STAT_ERR0 + k], sizeof(float));
  *count = passes;
  return SPH_OK;
}

extern "C" int sph_compute_energy(sph_ctx *c, double *ke) {
  int rc = require_state(c);
  if (rc) return rc;
  if (!ke) return arg_error("NULL argument");
  rc = launch_kinetic_energy(c, c->red);
  if (rc) return rc;
  // per-block partial sums, added here in block order: the same bits on every run
  std::vector<double> part(c->n ? blocks_for((uint32_t)c->n) : 0);
  if (!part.empty())
    SPH_HIP(hipMemcpyAsync(part.data(), c->red, sizeof(double) * part.size(),
                           hipMemcpyDeviceToHost, c->stream));
  SPH_HIP(hipStreamSynchronize(c->stream));
  double sum = 0.0;
  for (double v : part) sum += v;
  *ke = sum;
  return SPH_OK;
}

// --------------------------------------------------------------------------
// Inspection
// --------------------------------------------------------------------------
extern "C" int sph_download_field(sph_ctx *c, int field, float *out) {
  int rc = require_lists(c);
  if (rc) return rc;
  if (!out) return arg_error("out is NULL");
  const int n = c->n;
  std::vector<uint32_t> ids;
  if ((rc = fetch(c, c->id_s, (size_t)n, &ids))) return rc;
  std::vector<float4> v4;
  std::vector<float> v1;
  int comps = 3;
  switch (field) {
    case SPH_FIELD_POSITION_STAR: rc = fetch(c, (const float4 *)c->pos, n, &v4); break;
    case SPH_FIELD_DP: rc = fetch(c, (const float4 *)c->dpbuf, n, &v4); break;
    case SPH_FIELD_VELOCITY: rc = fetch(c, (const float4 *)c->vel, n, &v4); break;
    case SPH_FIELD_COLOR:
    case SPH_FIELD_VORTICITY: rc = fetch(c, (const float4 *)c->aux, n, &v4); break;
    case SPH_FIELD_POSITION_OLD: rc = fetch(c, (const float4 *)c->xold, n, &v4); break;
    case SPH_FIELD_DENSITY: comps = 1; rc = fetch(c, (const float *)c->rho, n, &v1); break;
    case SPH_FIELD_LAMBDA:
      comps = 1;
      rc = fetch(c, (const float4 *)c->pos, n, &v4);
      break;
    case SPH_FIELD_LAMBDA_SUM:
      // already in particle-id order
      if (!c->lam_prev) return state_error("no warm start or logging enabled");
      if ((rc = fetch(c, (const float *)c->lam_prev, n, &v1))) return rc;
      memcpy(out, v1.data(), sizeof(float) * n);
      return SPH_OK;
    default: return arg_error("unknown field");
  }
  if (rc) return rc;
  for (int s = 0; s < n; ++s) {
    const uint32_t id = ids[s];
    if (field == SPH_FIELD_DENSITY) {
      out[id] = v1[s];
    } else if (field == SPH_FIELD_LAMBDA) {
      out[id] = v4[s].w;
    } else {
      out[3 * (size_t)id + 0] = v4[s].x;
      out[3 * (size_t)id + 1] = v4[s].y;
      out[3 * (size_t)id + 2] = v4[s].z;
    }
  }
  (void)comps;
  return SPH_OK;
}

extern "C" int sph_download_neighbors(sph_ctx *c, uint32_t *out_ids, int *counts) {
  int rc = require_lists(c);
  if (rc) return rc;
  if (!counts) return arg_error("NULL argument");
  const int n = c->n;
  const size_t groups = ((size_t)n + kWave - 1) / kWave;
  std::vector<uint32_t> ids, list, head;
  if ((rc = fetch(c, c->id_s, (size_t)n, &ids))) return rc;
  if ((rc = fetch(c, (const uint32_t *)c->list_head, (size_t)n * 4, &head))) return rc;
  if (!out_ids) {
    for (int s = 0; s < n; ++s) counts[ids[s]] = (int)(head[(size_t)s * 4 + 3] & kHeadCount);
    return SPH_OK;
  }
  if ((rc = fetch(c, (const uint32_t *)c->list, groups * kListStride, &list))) return rc;
  for (int s = 0; s < n; ++s) {
    const uint32_t id = ids[s];
    const uint32_t *h = &head[(size_t)s * 4];
    const int cnt = (int)(h[3] & kHeadCount);
    const int e0 = (int)((h[3] >> 8) & kHeadCount), e1 = (int)((h[3] >> 16) & kHeadCount);
    counts[id] = cnt;
    const size_t base = (size_t)(s >> 6) * kListStride + (size_t)(s & 63) * 4;
    for (int k = 0; k < cnt; ++k) {
      uint32_t j;
      if (h[3] & kHeadWide) {
        j = list[base + (size_t)(k >> 2) * kWave * 4 + (k & 3)];
      } else {
        const uint32_t w = list[base + (size_t)(k >> 3) * kWave * 4 + ((k & 7) >> 1)];
        j = (k < e0 ? h[0] : (k < e1 ? h[1] : h[2])) + ((k & 1) ? w >> 16 : w & 0xFFFFu);
      }
      if (j >= (uint32_t)n) return state_error("neighbour list entry out of range");
      out_ids[(size_t)id * kMaxNb + k] = ids[j];
    }
  }
  return SPH_OK;
}

extern "C" int sph_download_sorted(sph_ctx *c, uint32_t *keys, uint32_t *ids) {
  int rc = require_lists(c);
  if (rc) return rc;
  const int n = c->n;
  if (n && keys) SPH_HIP(hipMemcpyAsync(keys, c->key_s, 4ull * n, hipMemcpyDeviceToHost, c->stream));
  if (n && ids) SPH_HIP(hipMemcpyAsync(ids, c->id_s, 4ull * n, hipMemcpyDeviceToHost, c->stream));
  SPH_HIP(hipStreamSynchronize(c->stream));
  return SPH_OK;
}

extern "C" int sph_download_cell_bounds(sph_ctx *c, uint32_t *start, uint32_t *end) {
  int rc = require_lists(c);
  if (rc) return rc;
  std::vector<uint32_t> cs;
  if ((rc = fetch(c, (const uint32_t *)c->cell_start, (size_t)c->cells + 1, &cs))) return rc;
  for (uint32_t i = 0; i < c->cells; ++i) {
    if (start) start[i] = cs[i];
    if (end) end[i] = cs[i + 1];
  }
  return SPH_OK;
}

// --------------------------------------------------------------------------
// Standalone collision and device plumbing
// --------------------------------------------------------------------------
struct sph_sdf {
  SdfDev dev;
  SdfBox64 box;  // the world box as given (double)
  float *buf;
};

extern "C" int sph_sdf_create(const sph_obstacle_desc *desc, sph_sdf **out) {
  if (!desc || !desc->sdf || !out) return arg_error("NULL argument");
  *out = nullptr;
  sph_sdf *s = new sph_sdf();
  s->buf = nullptr;
  const sph_aabb &wb = desc->world_bounds;
  s->box = SdfBox64{{wb.min_x, wb.min_y, wb.min_z}, {wb.max_x, wb.max_y, wb.max_z}};
  int rc = make_sdf(desc, &s->dev, &s->buf, nullptr);
  if (rc) {
    if (s->buf) (void)hipFree(s->buf);
    delete s;
    return rc;
  }
  *out = s;
  return SPH_OK;
}

extern "C" int sph_sdf_destroy(sph_sdf *s) {
  if (!s) return SPH_OK;
  if (s->buf) (void)hipFree(s->buf);
  delete s;
  return SPH_OK;
}

extern "C" int sph_calculate_particle_collisions(float *x, float *y, float *z, int n,
                                                 const sph_sdf *sdf, void *stream) {
  if (n < 0 || !sdf || (n > 0 && (!x || !y || !z))) return arg_error("invalid argument");
  int rc = launch_collide_soa(x, y, z, (uint32_t)n, sdf->dev, (hipStream_t)stream);
  if (rc) return rc;
  SPH_HIP(hipStreamSynchronize((hipStream_t)stream));
  return SPH_OK;
}

extern "C" int sph_calculate_particle_collisions_f64(double *x, double *y, double *z, int n,
                                                     const sph_sdf *sdf, void *stream) {
  if (n < 0 || !sdf || (n > 0 && (!x || !y || !z))) return arg_error("invalid argument");
  int rc = launch_collide_soa_f64(x, y, z, (uint32_t)n, sdf->dev, sdf->box, (hipStream_t)stream);
  if (rc) return rc;
  SPH_HIP(hipStreamSynchronize((hipStream_t)stream));
  return SPH_OK;
}

extern "C" int sph_device_malloc(void **ptr, size_t bytes) {
  if (!ptr) return arg_error("ptr is NULL");
  SPH_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return SPH_OK;
}

extern "C" int sph_device_free(void *ptr) {
  if (ptr) SPH_HIP(hipFree(ptr));
  return SPH_OK;
}

extern "C" int sph_memcpy_htod(void *dst, const void *src, size_t bytes) {
  if (bytes) SPH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return SPH_OK;
}

extern "C" int sph_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  if (bytes) SPH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return SPH_OK;
}

// --------------------------------------------------------------------------
// Timing
// --------------------------------------------------------------------------
// Launch autotuner.  Tries 64, 128 and 256 threads per block and then 8, 16
// and 32 gathers in flight per wave for every neighbour-gather kernel and,
// with an obstacle, the fused versus separate obstacle push, timing
// `substeps` substeps of each candidate with HIP events from the same saved
// state; keeps the fastest choice per kernel and restores the state.  None of
// the choices changes any result.
static const int kTunable[] = {K_BUILD_LISTS, K_DENSITY_LAMBDA, K_DENSITY, K_LAMBDA, K_DP,
                               K_DP_UPDATE, K_COLOR, K_SURFACE, K_XSPH, K_VORTICITY,
                               K_CONFINE, K_WARM_LAMBDA};

extern "C" int sph_autotune(sph_ctx *c, int substeps, sph_tuning *out) {
  if (c) c->gen++;
  int rc = require_state(c);
  if (rc) return rc;
  if (substeps < 1) return arg_error("substeps must be >= 1");
  const size_t n = (size_t)c->n;
  if (!n) return SPH_OK;
  float4 *sx = nullptr, *sv = nullptr;
  float *sl = nullptr;
  SPH_HIP(hipMalloc((void **)&sx, sizeof(float4) * n));
  hipError_t e = hipMalloc((void **)&sv, sizeof(float4) * n);
  if (e == hipSuccess && c->lam_prev) e = hipMalloc((void **)&sl, sizeof(float) * n);
  auto release = [&]() {
    if (sx) (void)hipFree(sx);
    if (sv) (void)hipFree(sv);
    if (sl) (void)hipFree(sl);
  };
  if (e != hipSuccess) {
    release();
    return hip_fail(e, "hipMalloc (autotune snapshot)");
  }
  auto copy_state = [&](bool save) -> hipError_t {
    hipError_t r = hipMemcpyAsync(save ? sx : c->x_can, save ? c->x_can : sx, sizeof(float4) * n,
                                  hipMemcpyDeviceToDevice, c->stream);
    if (r == hipSuccess)
      r = hipMemcpyAsync(save ? sv : c->v_can, save ? c->v_can : sv, sizeof(float4) * n,
                         hipMemcpyDeviceToDevice, c->stream);
    if (r == hipSuccess && sl)
      r = hipMemcpyAsync(save ? sl : c->lam_prev, save ? c->lam_prev : sl, sizeof(float) * n,
                         hipMemcpyDeviceToDevice, c->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(c->stream);
    return r;
  };
  const long long substeps0 = c->substeps;
  const bool prof0 = c->prof.on;
  const auto tune0 = c->tune;
  e = copy_state(true);
  // one trial: restore, one untimed substep, `substeps` timed ones
  auto trial = [&](double *per_kernel, double *total) -> int {
    hipError_t r = copy_state(false);
    if (r != hipSuccess) return hip_fail(r, "autotune restore");
    c->prof.on = false;
    int q = sph_step(c, 1);
    if (q) return q;
    prof_zero(c);
    c->prof.on = true;
    q = sph_step(c, substeps);
    if (q) return q;
    prof_collect(c);
    *total = 0.0;
    for (int k = 0; k < K_COUNT; ++k) {
      per_kernel[k] = c->prof.launches[k] ? c->prof.total_ms[k] / substeps : 0.0;
      *total += per_kernel[k];
    }
    c->prof.on = false;
    return SPH_OK;
  };
  double best[K_COUNT], cur[K_COUNT], total = 0.0, before = 0.0;
  uint16_t best_bs[K_COUNT] = {};
  for (int k = 0; k < K_COUNT; ++k) best[k] = 1e300;
  const uint16_t sizes[3] = {64, 128, 256};
  for (int b = 0; b < 3 && e == hipSuccess && !rc; ++b) {
    for (int k : kTunable) c->tune.block[k] = sizes[b];
    graph_reset(c);
    rc = trial(cur, &total);
    if (rc) break;
    if (sizes[b] == 256) before = total;
    for (int k : kTunable)
      if (cur[k] > 0.0 && cur[k] < best[k]) {
        best[k] = cur[k];
        best_bs[k] = sizes[b];
      }
  }
  // gathers in flight per wave, with the chosen block sizes
  for (int k : kTunable) c->tune.block[k] = best_bs[k] ? best_bs[k] : (uint16_t)kBlock;
  uint8_t best_depth[K_COUNT] = {};
  for (int k = 0; k < K_COUNT; ++k) best[k] = 1e300;
  const uint8_t depths[3] = {8, 16, 32};
  for (int d = 0; d < 3 && e == hipSuccess && !rc; ++d) {
    for (int k : kTunable) c->tune.depth[k] = depths[d];
    graph_reset(c);
    rc = trial(cur, &total);
    if (rc) break;
    for (int k : kTunable)
      if (cur[k] > 0.0 && cur[k] < best[k]) {
        best[k] = cur[k];
        best_depth[k] = depths[d];
      }
  }
  for (int k : kTunable) c->tune.depth[k] = best_depth[k] ? best_depth[k] : (uint8_t)16;
  bool split_best = c->tune.split_collide;
  double after = 0.0;
  if (!rc && e == hipSuccess) {
    const bool obstacle = c->nscenes ? c->batch_sdf != nullptr : c->sdf.data != nullptr;
    double t_split[2] = {1e300, 1e300};
    for (int sp = 0; sp < (obstacle ? 2 : 1) && !rc; ++sp) {
      c->tune.split_collide = obstacle ? (sp == 0) : tune0.split_collide;
      graph_reset(c);
      rc = trial(cur, &total);
      t_split[sp] = total;
    }
    if (obstacle && !rc) split_best = t_split[0] <= t_split[1];
    after = obstacle ? std::min(t_split[0], t_split[1]) : t_split[0];
    c->tune.split_collide = split_best;
  }
  if (!rc && e == hipSuccess) {
    // list format: the other one than the trials above used
    const bool wide0 = c->tune.wide_lists;
    c->tune.wide_lists = !wide0;
    graph_reset(c);
    rc = trial(cur, &total);
    if (!rc && total < after)
      after = total;
    else
      c->tune.wide_lists = wide0;
  }
  if (rc || e != hipSuccess) c->tune = tune0;
  const hipError_t er = copy_state(false);
  release();
  prof_zero(c);
  c->prof.on = prof0;
  c->substeps = substeps0;
  c->lists_valid = c->sorted_valid = false;
  graph_reset(c);
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "autotune snapshot");
  if (er != hipSuccess) return hip_fail(er, "autotune restore");
  if (out) {
    sph_get_tuning(c, out);
    out->ms_per_substep_default = before;
    out->ms_per_substep_tuned = after;
  }
  return SPH_OK;
}

extern "C" int sph_get_tuning(sph_ctx *c, sph_tuning *out) {
  if (!c || !out) return arg_error("NULL argument");
  memset(out, 0, sizeof(*out));
  out->split_collide = c->tune.split_collide;
  out->wide_lists = c->tune.wide_lists;
  for (int k = 0; k < K_COUNT && k < SPH_TUNE_KERNELS; ++k) {
    out->block[k] = c->tune.block[k] ? c->tune.block[k] : kBlock;
    out->depth[k] = c->tune.depth[k] ? c->tune.depth[k] : 16;
  }
  return SPH_OK;
}

extern "C" int sph_set_tuning(sph_ctx *c, const sph_tuning *t) {
  if (!c || !t) return arg_error("NULL argument");
  auto tune = c->tune;
  for (int k : kTunable) {
    const int b = t->block[k], d = t->depth[k];
    if (b != 0 && b != 64 && b != 128 && b != 256)
      return arg_error("threads per block must be 0, 64, 128 or 256");
    if (d != 0 && d != 8 && d != 16 && d != 32)
      return arg_error("gathers in flight must be 0, 8, 16 or 32");
    tune.block[k] = (uint16_t)b;
    tune.depth[k] = (uint8_t)d;
  }
  tune.split_collide = t->split_collide != 0;
  tune.wide_lists = t->wide_lists != 0;
  c->tune = tune;
  // lists already built keep their format; the next search uses the new one
  graph_reset(c);
  return SPH_OK;
}

extern "C" int sph_profile_enable(sph_ctx *c, int enable) {
  if (!c) return arg_error("context is NULL");
  if (!enable) prof_collect(c);
  c->prof.on = enable != 0;
  graph_reset(c);
  return SPH_OK;
}

extern "C" int sph_profile_read(sph_ctx *c, int kernel, double *total_ms, long long *launches) {
  if (!c) return arg_error("context is NULL");
  if (kernel < 0 || kernel >= K_COUNT) return arg_error("unknown kernel id");
  prof_collect(c);
  if (total_ms) *total_ms = c->prof.total_ms[kernel];
  if (launches) *launches = c->prof.launches[kernel];
  return SPH_OK;
}

extern "C" int sph_profile_reset(sph_ctx *c) {
  if (!c) return arg_error("context is NULL");
  prof_collect(c);
  for (int k = 0; k < K_COUNT; ++k) {
    c->prof.total_ms[k] = 0.0;
    c->prof.launches[k] = 0;
  }
  return SPH_OK;
}

extern "C" const char *sph_profile_kernel_name(int kernel) {
  return (kernel >= 0 && kernel < K_COUNT) ? kKernelNames[kernel] : "";
}

extern "C" int sph_profile_kernel_count(void) { return K_COUNT; }

// --------------------------------------------------------------------------
// Counter calibration: plain 16-byte-vector copy
// --------------------------------------------------------------------------
__global__ void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

extern "C" int sph_device_copy(void *dst, const void *src, size_t bytes, void *stream) {
  if (!dst || !src || bytes % 16) return arg_error("sph_device_copy: bytes must be a multiple of 16");
  const size_t n = bytes / 16;
  if (!n) return SPH_OK;
  const unsigned blocks = (unsigned)std::min<size_t>((n + kBlock - 1) / kBlock, 65536);
  k_copy16<<<blocks, kBlock, 0, (hipStream_t)stream>>>((const uint4 *)src, (uint4 *)dst, n);
  SPH_HIP(hipGetLastError());
  SPH_HIP(hipStreamSynchronize((hipStream_t)stream));
  return SPH_OK;
}
