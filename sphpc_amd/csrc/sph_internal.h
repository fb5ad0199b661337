// Internal declarations shared by the HIP translation units of libsphpc_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "sphpc_amd.h"

namespace sph {

constexpr int kMaxNb = SPH_MAX_NEIGHBORS;
constexpr int kBlock = 256;          // threads per block for elementwise kernels
constexpr int kWave = 64;            // CDNA wavefront
constexpr int kListStride = kWave * kMaxNb;  // list words per 64-particle group
// Largest context: float4 particle arrays (plus their padding) stay below
// 4 GiB, so that gathers address them with 32-bit byte offsets.
constexpr uint32_t kMaxParticles = SPH_MAX_PARTICLES;
static_assert((kMaxParticles + 64ull) * 16ull <= (1ull << 32), "32-bit byte offsets");
// Particles per block of the per-scene error reduction of a batch.
constexpr uint32_t kSceneChunk = 4096;
// Neighbour-list head per particle: the first candidate index of each of the
// three dz layers and kHeadCount | layer-0 count << 8 | layers-0..1 count << 16
// | kHeadWide.  A compact list holds 16-bit offsets from the layer's first
// index; a wide one (whole wavefronts, when a layer spans more than
// kCompactSpan indices or the tuning asks for it) holds 32-bit indices.
constexpr uint32_t kHeadCount = 0xFFu;
constexpr uint32_t kHeadWide = 1u << 31;
constexpr uint32_t kCompactSpan = 1u << 16;
// The sorted position arrays hold kPosPad readable entries past the capacity:
// the list build loads candidates 8 at a time and may read up to 7 entries
// past the end of a range (their tests are discarded).
constexpr size_t kPosPad = 8;


// Device-side constants of one run, all FP32 except the exact-predicate h^2.
struct DevParams {
  float h, h2, inv_h;
  float h2_lo, h2_hi;        // guard band around h^2 for the list predicate
  double h2d;                // (double)h * (double)h, exact
  float rmin, rmin2;         // r floor 1e-4 h (particles.c:125,265,344,636,725)
  float W_norm, DelW_norm, DelW_poly_norm, C_norm;
  float mass, rest_density, inv_rest_density;
  float mW0;                 // m * W(0)
  float mWn;                 // m * W_norm
  // Density sum in 32-bit fixed point: each neighbour adds the truncated
  // integer u^3 with u = k3 (h^2 - r^2) = fma(-r^2, k3, c1), computed in
  // double, where k3^3 h^6 = 2^26 bounds a term.  The constraint is
  // C = rho/rho0 - 1 = dens_a * S + dens_c0, both constants as float pairs
  // (hi + lo).
  double dens_k3, dens_c1;       // k3, k3 h^2
  float dens_a_hi, dens_a_lo;    // m W_norm / (rho0 k3^3)
  float dens_c0_hi, dens_c0_lo;  // m W(0) / rho0 - 1
  float lam_k2;              // (m * DelW_norm / rho0)^2
  float lambda_eps;
  float dt, g_dt;            // g_dt = -g * dt
  float v_max;
  float xsph_c, gamma;
  float coh_k;               // -20 * gamma * m^2 * C_norm
  float curv_k;              // -gamma * m
  float h6_64;               // h^6 / 64
  float vort_eps;
  float conf_k;              // dt * vort_eps / m  (model 0) or dt * vort_eps (model 1)
  float color_k;             // h * m * DelW_poly_norm
  float div_k;               // m / rho0 * DelW_norm (velocity divergence feature)
  float box_lo[3], box_hi[3], box_hi_clamp[3];  // clamp to [lo, hi - 1e-4 h)
  uint32_t nx, ny, nz, nxy;
  int vort_model;
};

// Obstacle grid in device memory, sampled with software trilinear filtering.
struct SdfDev {
  const float *data;  // nullptr: no obstacle
  int dim[3];
  float wmin[3], wmax[3];
  float len[3];       // float(wmax - wmin)
  float inv_dim[3];   // 1/dim (texel size in normalised coordinates)
  float m2w;          // world length / model length (obstacle_cuda.cu:101-105)
};

// The obstacle's world box in double, for callers holding FP64 positions.
struct SdfBox64 {
  double lo[3], hi[3];
};

// Batched scenes: each scene has its own box (local coordinates from the
// origin), hash grid (a contiguous range of global cell keys from cell_base)
// and obstacle.  A scene's particles are contiguous in the id order and, as
// its keys are contiguous, at the same indices in the sorted order.
struct SceneDev {
  uint32_t cell_base, nx, ny, nz, nxy;
  float box_hi[3], box_hi_clamp[3];
  SdfDev sdf;
};

struct BatchDev {
  const SceneDev *scenes;     // nullptr: one scene described by DevParams
  const uint16_t *scene_of;   // scene of particle index i (both orders)
};

enum StatSlot {
  STAT_DROPPED = 0,      // 64-bit counter (two words)
  STAT_CLAMPS = 2,       // 64-bit counter (two words)
  STAT_ITERS = 4,
  STAT_CONVERGED = 5,
  STAT_FALLBACK = 6,     // warm start reverted to the plain first iteration
  STAT_ERR0 = 8,         // per-iteration max error bits (kMaxIters + 1 slots)
  STAT_WORDS = 256
};
constexpr int kMaxIters = 128;

// Lambda warm-start network: 12 features (padded to 16) -> 64 SiLU -> 64 SiLU -> 1.
// FP32 master parameters in one vector: W1[16][64] (row = input), b1[64],
// W2[64][64] (row = input), b2[64], w3[64], b3.
constexpr int kMlpIn = 12, kMlpPad = 16, kMlpH = 64;
constexpr int kMlpW1 = 0, kMlpB1 = kMlpW1 + kMlpPad * kMlpH, kMlpW2 = kMlpB1 + kMlpH,
              kMlpB2 = kMlpW2 + kMlpH * kMlpH, kMlpW3 = kMlpB2 + kMlpH,
              kMlpB3 = kMlpW3 + kMlpH, kMlpParams = kMlpB3 + 1;  // 5313
// Normalisation words: feature means, reciprocal deviations, target shift/scale.
// kNormYMax > 0 caps the network output (the clipping bound of its training
// targets); kNormStep > 0 scales the predicted multiplier (damping, 0 means 1).
constexpr int kNormMu = 0, kNormIsd = 16, kNormYShift = 32, kNormYScale = 33, kNormYMax = 34,
              kNormStep = 35, kNormWords = 36;

// BF16 inference weights packed in MFMA operand order (layout in sph_mlp_dev.h).
struct MlpPacked {
  uint4 a1[2 * 64];        // layer 1 A operands [row block][lane], 8 bf16 each
  uint4 a2[2 * 4 * 64];    // layer 2 A operands [row block][k chunk][lane]
  float4 b1[2 * 2 * 4];    // layer 1 bias [row block][lane half][accumulator / 4]
  float4 b2[2 * 2 * 4];    // layer 2 bias, same order
  float4 w3[2 * 2 * 4];    // output weights (bf16 values), same order
  float norm[kNormWords];
  float b3;
};

enum KernelId {
  K_PREDICT_HASH = 0, K_HASH, K_GRAVITY, K_PREDICT, K_RADIX_UP, K_RADIX_DOWN, K_SCAN,
  K_REORDER, K_BUILD_LISTS, K_DENSITY_LAMBDA, K_DENSITY, K_LAMBDA, K_DP, K_DP_UPDATE,
  K_POS_STARS, K_VELOCITY, K_COLOR, K_SURFACE, K_XSPH, K_VORTICITY, K_CONFINE, K_COMMIT,
  K_AUX, K_WARM_LAMBDA, K_COLLIDE, K_LAMBDA_COPY, K_COUNT
};

struct ScanScratch {
  uint32_t *tile_sums = nullptr;
  size_t tile_cap = 0;
};

struct Profiler {
  bool on = false;
  struct Pair { hipEvent_t a, b; int kernel; };
  std::vector<Pair> pending;
  std::vector<hipEvent_t> pool;
  double total_ms[K_COUNT] = {};
  long long launches[K_COUNT] = {};
};

}  // namespace sph

struct sph_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  sph_params P{};
  sph_aabb box{};
  sph::DevParams dev{};
  sph::SdfDev sdf{};
  float *sdf_buf = nullptr;

  uint32_t cells = 0;
  int key_bits = 1;
  int cap = 0;
  int n = 0;

  // canonical (particle-id order) state
  float4 *x_can = nullptr, *v_can = nullptr, *xs_can = nullptr;
  // sort
  uint32_t *keys = nullptr, *keys_alt = nullptr, *ids = nullptr, *ids_alt = nullptr;
  uint32_t *hist = nullptr;
  size_t hist_cap = 0;
  const uint32_t *key_s = nullptr, *id_s = nullptr;  // final sorted buffers
  // cells + 1 entries; cell_end holds the run ends of the sorted keys, zero
  // between searches
  uint32_t *cell_end = nullptr, *cell_start = nullptr;
  sph::ScanScratch scan;
  // sorted working set
  float4 *pos_a = nullptr, *pos_b = nullptr, *pos = nullptr;  // pos: current x*
  float4 *xold = nullptr;
  float4 *vel_a = nullptr, *vel_b = nullptr, *vel = nullptr;  // vel: current v
  float4 *aux = nullptr;  // colour normal, then vorticity (w = |omega|)
  float4 *pw = nullptr;   // (x*, |omega|) for the confinement gather
  // sorted x* again as three float arrays (x, then y, then z, soa_stride
  // apart), written by the reorder pass for the list build's candidate loads
  float *sx = nullptr;
  uint32_t soa_stride = 0;
  float4 *dpbuf = nullptr;  // staged delta p
  float *rho = nullptr;
  // stage path: C = rho/rho0 - 1 of the last density stage (read by a lambda
  // stage that runs its own gather) and the lambda that density pass computed
  float *cstage = nullptr, *lam_stage = nullptr;
  uint32_t *list = nullptr;     // sorted neighbour indices, quads of slots per 64 particles
  uint4 *list_head = nullptr;   // per sorted particle: layer bases and counts
  // state transfers: SoA x, y, z, vx, vy, vz of up to cap particles, on the
  // device and in pinned host memory (allocated on first use)
  float *io_dev = nullptr, *io_host = nullptr;
  uint32_t *stats = nullptr;  // sph::STAT_WORDS words
  uint32_t *pair_partial = nullptr;  // per-block list-length sums of the last search
  uint32_t pair_blocks = 0;
  double *red = nullptr;      // FP64 reduction scratch
  // learned warm start
  float *lam_acc = nullptr;   // sorted: sum over iterations of stiffness * lambda
  float *lam_prev = nullptr;  // canonical: lam_acc of the previous substep
  float *feat = nullptr;      // sorted feature rows (kMlpPad floats) when logging
  sph::MlpPacked *warm_net = nullptr;
  bool warm_on = false;
  float warm_prev_scale = 0.0f;  // > 0: warm start from the previous multiplier
  bool log_on = false;
  long long feat_substep = -1;  // substep count after the last logged substep
  // batched scenes (sph_create_batch); nscenes == 0: single scene
  int nscenes = 0;
  sph::SceneDev *scenes_dev = nullptr;
  uint16_t *scene_of = nullptr;
  float *batch_sdf = nullptr;
  uint32_t *scene_err = nullptr;
  uint32_t *scene_err0 = nullptr;  // per scene: error before a warm-started step
  std::vector<uint32_t> scene_first;  // nscenes + 1 particle offsets
  // (scene, first, end) of particle chunks of at most kSceneChunk, each within
  // one scene (a scene's particles are one index range in both orders)
  uint4 *scene_chunks = nullptr;
  uint32_t nchunks = 0;
  sph::BatchDev batch() const { return sph::BatchDev{scenes_dev, scene_of}; }
  // launch choices (sph_autotune)
  struct {
    bool split_collide = true;  // obstacle push as its own pass after k_dp_update
    uint16_t block[sph::K_COUNT] = {};  // threads per block of gather kernels (0: 256)
    uint8_t depth[sph::K_COUNT] = {};   // gathers in flight per wave: 8, 16, 32 (0: 16)
    bool wide_lists = false;  // 32-bit list entries for every wave
  } tune;
  // whole-substep HIP graph (sph_set_graph)
  bool graph_on = false;
  bool graph_valid = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  struct {
    float4 *pos, *vel;
    const uint32_t *key_s, *id_s;
    uint32_t pair_blocks;
    int last_err_slot;
    bool warm;
  } graph_after{};
  int last_err_slot = -1;
  int staged_slot = 0;
  // state generation: bumped by every call that may change the state; the
  // stage path's density pass records it, and the lambda stage reuses that
  // pass's lambda when nothing ran in between
  unsigned long long gen = 0, lam_gen = ~0ull;
  unsigned long long dp_gen = ~0ull;  // likewise the delta-p stage's updated x*
  float4 *dp_next = nullptr;
  bool state_valid = false;
  bool sorted_valid = false;
  bool lists_valid = false;
  long long substeps = 0;
  sph::Profiler prof;
};

namespace sph {

void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define SPH_HIP(call)                                   \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return sph::hip_fail(e_, #call); \
  } while (0)

inline uint32_t blocks_for(uint32_t n, uint32_t per = kBlock) { return (n + per - 1) / per; }

// Launch helpers with optional event timing.
void prof_begin(sph_ctx *c, int kernel, hipEvent_t *a);
void prof_end(sph_ctx *c, int kernel, hipEvent_t a);

// Sort / scan (sph_sort.hip).  out[i] = in[0] + ... + in[i-1] (max_op: the
// maximum of those, 0 for none); write_total also sets out[len].
int exclusive_scan(uint32_t *in, uint32_t *out, uint32_t len, bool zero_in, bool write_total,
                   ScanScratch *scratch, hipStream_t s, sph_ctx *prof, bool max_op = false);
int radix_sort_pairs(const uint32_t *kin, const uint32_t *vin, uint32_t *k0, uint32_t *v0,
                     uint32_t *k1, uint32_t *v1, uint32_t n, int key_bits, uint32_t **hist,
                     size_t *hist_cap, ScanScratch *scratch, hipStream_t s, sph_ctx *prof,
                     const uint32_t **kout, const uint32_t **vout);
int count_keys(const uint32_t *keys, uint32_t n, uint32_t *count, uint32_t cells,
               hipStream_t s);

// Physics launchers (sph_physics.hip).
int launch_gravity(sph_ctx *c);
int launch_predict(sph_ctx *c);
int launch_neighbor_search(sph_ctx *c, bool fused);
// mode 0: density + lambda (fused substep), 1: density, 2: lambda, 3: density
// with rho into pos.w (tolerance loop), 4: density + lambda with rho into pos.w
// and lambda into dpbuf (the stage path; launch_lambda_from_density moves it)
int launch_density(sph_ctx *c, int mode, int err_slot);
int launch_lambda_from_density(sph_ctx *c);
// xnext: also the position update of launch_position_stars, into xnext
int launch_dp(sph_ctx *c, int substep, float4 *xnext = nullptr);
int launch_position_stars(sph_ctx *c);
int launch_dp_update(sph_ctx *c, int substep, int err_slot, bool accumulate = false);
int launch_velocity(sph_ctx *c);
int launch_color(sph_ctx *c, bool with_velocity);
int launch_surface(sph_ctx *c);
int launch_xsph(sph_ctx *c);
int launch_vorticity(sph_ctx *c);
int launch_confine(sph_ctx *c, bool commit);
int launch_commit(sph_ctx *c);
int launch_substep_init(sph_ctx *c);
int launch_collide_soa(float *x, float *y, float *z, uint32_t n, const SdfDev &o,
                       hipStream_t s);
int launch_collide_soa_f64(double *x, double *y, double *z, uint32_t n, const SdfDev &o,
                           const SdfBox64 &b, hipStream_t s);
int launch_kinetic_energy(sph_ctx *c, double *out_dev);
int launch_warm_lambda(sph_ctx *c, bool log);
int launch_warm_prev(sph_ctx *c);
int launch_scatter_lambda(sph_ctx *c);
int launch_warm_samples(sph_ctx *c, float *x, float *t, uint32_t count, int every);
int launch_warm_fallback(sph_ctx *c, const float4 *pre_warm);
int launch_scene_errors(sph_ctx *c, uint32_t *dst = nullptr);
int mlp_pack(sph_mlp *m, MlpPacked *dst, hipStream_t s);  // sph_mlp.hip

}  // namespace sph
