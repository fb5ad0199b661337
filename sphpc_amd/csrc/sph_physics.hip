// Position-based-fluids substep kernels for gfx950.
//
// Particles live in two orders.  The canonical state (x_can, v_can) is in
// particle-id order, like the reference's arrays.  After the neighbour search
// the substep works on copies reordered into (cell key, id) order, so that a
// particle's 27-cell neighbourhood is nine contiguous index ranges and the
// neighbour gathers of one wavefront hit the same cache lines.  The commit
// kernel scatters x* and v back to id order at the end of the substep.
//
// Vector fields are float4.  pos.w carries lambda between the density/lambda
// pass and the delta-p pass, and rho afterwards; aux.w carries |omega|.
//
// Neighbour lists are stored per 64-particle group in 16-byte quads: lane
// s%64's quad q is at word (s/64)*64*60 + q*256 + (s%64)*4, so a wave's load
// of one quad per lane covers 1 KB contiguously.  In (dz, dy, dx, id) order a
// particle's neighbours have strictly increasing sorted indices, and those of
// one dz layer lie within a few grid rows of the layer's first candidate, so
// a compact list stores 16-bit offsets from the three layer bases held in the
// particle's list head (8 neighbours per quad); a wide list stores the 32-bit
// indices (4 per quad).
#include <type_traits>

#include "sph_internal.h"
#include "sph_mlp_dev.h"

namespace sph {

// Logical block index for a launch of nb blocks so that each XCD (blocks are
// dispatched round-robin over 8 XCDs) walks one contiguous range of particles
// and keeps its neighbourhood in its own L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t q = nb >> 3, r = nb & 7u;
  const uint32_t x = b & 7u, idx = b >> 3;
  return x * q + min(x, r) + idx;
}

__device__ __forceinline__ uint32_t particle_index(uint32_t n_blocks) {
  return xcd_block(blockIdx.x, n_blocks) * blockDim.x + threadIdx.x;
}

__device__ __forceinline__ const uint32_t *list_ptr(const uint32_t *list, uint32_t s) {
  return list + (size_t)(s >> 6) * kListStride + (s & 63u) * 4u;
}

// One particle's list: its first quad, its head and the wave's format.
struct NbList {
  const uint32_t *L;
  uint4 h;      // layer bases (x, y, z) and counts (w)
  int c;        // neighbours
  int e0, e1;   // neighbours in layer 0, in layers 0 and 1
  bool wide;    // uniform over the wavefront
};

__device__ __forceinline__ NbList nb_list(const uint32_t *list, const uint4 *head, uint32_t s) {
  NbList r;
  r.L = list_ptr(list, s);
  r.h = head[s];
  r.c = (int)(r.h.w & kHeadCount);
  r.e0 = (int)((r.h.w >> 8) & kHeadCount);
  r.e1 = (int)((r.h.w >> 16) & kHeadCount);
  // every lane of a wavefront holds the same format bit
  r.wide = (__builtin_amdgcn_readfirstlane(r.h.w) & kHeadWide) != 0u;
  return r;
}

// Launch bounds of the delta-p, density, colour and confinement kernels: at
// up to 16 gathers in flight at most 128 VGPRs, i.e. 4 waves per SIMD.  With
// 64-bit gather addresses delta p took 139 VGPRs (3 waves) and the bound made
// it 7% faster at 8M particles despite a few spilled registers (colour and
// confinement, at 138-140, were 1-4% slower with it).  With the buffer gathers
// below delta p spills 20 bytes under the bound, and colour and confinement
// (130-133 VGPRs without it) 12-16 bytes, and run 2-3% faster than at 3 waves.
// Five waves (96 VGPRs) would spill 164-220 bytes per lane.
#define GATHER_BOUNDS(NB) __launch_bounds__(kBlock, (NB) <= 16 ? 4 : 1)

// Visits the c neighbours of a particle in list order, NB at a time: first the
// batch's list quads, then its NB dependent gathers, then the arithmetic.  The
// gathers are latency-bound, so the depth NB (gathers in flight per wave) sets
// the speed: 16 was 13-22% faster than 8 in every gather pass on MI355X; the
// autotuner chooses 8, 16 or 32 per kernel.  With PREFETCH the list quads of
// batch k+1 are issued right after the gathers of batch k, so a batch costs
// one memory round trip instead of two; that pays for 16-byte gathers, and of
// the 32-byte (Pair4) ones only in the XSPH pass.  Slots past c are never used; their gathers
// load a valid index (the particle itself in a wide list, the layer-2 base in
// a compact one, which the build sets to the particle itself when that layer
// has no candidates).
constexpr int kBatch = 8;  // candidate batch of the list build
constexpr uint32_t kQuadStride = kWave * 4;  // list words between a lane's quads
// Hits the list build buffers in LDS per thread (16 bits each) for the
// coalesced write-out; later hits of a compact list are stored directly.  At
// 40 a 256-thread block holds 20 KB, so LDS admits 7 waves per SIMD, which the
// kernel's launch bound (70 VGPRs, no spill) also fits: measured at 8M
// particles 0.855 -> 0.842 ms against 48 slots at 6 waves (60 slots, 30 KB,
// admitted 5).
constexpr int kLdsHits = 40;
static_assert(kLdsHits % 8 == 0 && kLdsHits <= kMaxNb, "whole quads in LDS");

// Entry e of a lane's compact list (16-bit offset v), stored on its own: quad
// e / 8 (kQuadStride words apart), 32-bit word (e % 8) / 2, low half first.
__device__ __forceinline__ void store_entry16(uint32_t *out, int e, uint32_t v) {
  uint16_t *w = reinterpret_cast<uint16_t *>(out + (size_t)(e >> 3) * kQuadStride + ((e & 7) >> 1));
  w[e & 1] = (uint16_t)v;
}

template <int NB>
__device__ __forceinline__ void load_wide(const NbList &nl, int k0, uint32_t self,
                                          uint32_t (&j)[NB]) {
  const uint4 *Q = reinterpret_cast<const uint4 *>(nl.L);
  const int q0 = k0 >> 2;
#pragma unroll
  for (int u = 0; u < NB / 4; ++u) {
    uint4 a = make_uint4(self, self, self, self);
    if (k0 + 4 * u < nl.c) a = Q[(size_t)(q0 + u) * kWave];
    j[4 * u] = a.x;
    j[4 * u + 1] = a.y;
    j[4 * u + 2] = a.z;
    j[4 * u + 3] = a.w;
  }
#pragma unroll
  for (int t = 0; t < NB; ++t)
    if (k0 + t >= nl.c) j[t] = self;
}

// Compact quads: entry k = base of its layer + 16-bit offset; padding entries
// are offset 0 from the layer-2 base.
template <int NB>
__device__ __forceinline__ void load_compact(const NbList &nl, int k0, uint32_t (&j)[NB]) {
  static_assert(NB % 8 == 0, "compact quads hold 8 entries");
  const uint4 *Q = reinterpret_cast<const uint4 *>(nl.L);
  const int q0 = k0 >> 3;
  // values, not member lvalues: a select between member addresses would
  // keep the struct in memory (and the compiler moves such memory to LDS)
  const uint32_t b0 = nl.h.x, b1 = nl.h.y, b2 = nl.h.z;
  const int e0 = nl.e0, e1 = nl.e1, c = nl.c;
#pragma unroll
  for (int u = 0; u < NB / 8; ++u) {
    uint4 a = make_uint4(0u, 0u, 0u, 0u);
    if (k0 + 8 * u < c) a = Q[(size_t)(q0 + u) * kWave];
    const uint32_t w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = k0 + 8 * u + t;
      const uint32_t off = (t & 1) ? (w[t >> 1] >> 16) : (w[t >> 1] & 0xFFFFu);
      const uint32_t lb = k < e1 ? b1 : b2;
      j[8 * u + t] = (k < e0 ? b0 : lb) + off;
    }
  }
}

// A wavefront runs each batch for its longest list.  Batches in which no lane's
// list ends skip the per-neighbour count test, and a last batch in which every
// lane ends within NB/2 neighbours runs only NB/2 of them (the wave's longest
// list, not the mean, sets the work: half batches cut the padding slots).
template <typename T, int NB, bool PREFETCH, bool WIDE, bool LAST_SKIP, typename Load,
          typename Body>
__device__ __forceinline__ void gather_loop(const NbList &nl, uint32_t self, Load &load,
                                            Body &body) {
  uint32_t jn[NB];
  auto fetch = [&](int k0) {
    if (WIDE)
      load_wide<NB>(nl, k0, self, jn);
    else
      load_compact<NB>(nl, k0, jn);
  };
  const int c = nl.c;
  fetch(0);
  for (int k0 = 0; k0 < c; k0 += NB) {
    // the lanes still in the loop have c > k0
    if (!PREFETCH && k0 > 0) fetch(k0);
    if (NB >= 16 && __ballot(k0 + NB / 2 < c) == 0ull) {
      T q[NB / 2];
#pragma unroll
      for (int b = 0; b < NB / 2; ++b) q[b] = load(jn[b]);
#pragma unroll
      for (int b = 0; b < NB / 2; ++b)
        if (k0 + b < c) body(q[b]);
    } else {
      T q[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) q[b] = load(jn[b]);
      // with LAST_SKIP not on the wave's last batch: the decode is VALU work,
      // but as a block of its own it no longer interleaves with the
      // arithmetic (faster for the density pass only, measured)
      if (PREFETCH && (!LAST_SKIP || __ballot(k0 + NB < c) != 0ull)) fetch(k0 + NB);
      if (__ballot(k0 + NB > c) == 0ull) {
#pragma unroll
        for (int b = 0; b < NB; ++b) body(q[b]);
      } else {
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (k0 + b < c) body(q[b]);
      }
    }
  }
}

template <typename T, int NB, bool PREFETCH = true, bool LAST_SKIP = false, typename Load,
          typename Body>
__device__ __forceinline__ void for_neighbors(const NbList &nl, uint32_t self, Load load,
                                              Body body) {
  if (nl.wide)
    gather_loop<T, NB, PREFETCH, true, LAST_SKIP>(nl, self, load, body);
  else
    gather_loop<T, NB, PREFETCH, false, LAST_SKIP>(nl, self, load, body);
}

struct Pair4 {
  float4 a, b;
};

// Gathers of 16-byte records through a buffer resource with a 16-byte stride
// (`buffer_load ... idxen`): the texture addresser forms base + 16 j from the
// index itself, so a gather costs no VALU (a 64-bit address costs two, the
// zero high word and the shift-add) and needs no address registers.  At 16
// gathers in flight: colour and confinement 138-140 -> 130-133 VGPRs, delta p
// spills 20 bytes instead of 52 and density+lambda none instead of 12, XSPH at
// 8 gathers 112 -> 90 VGPRs (5 waves per SIMD instead of 4), and the autotuner
// now picks 16 gathers in flight for XSPH and vorticity instead of 8 (measured
// at 8M particles: vorticity 0.70 -> 0.58 ms, XSPH 0.67 -> 0.60, delta p,
// colour and confinement 3% faster, density+lambda unchanged).  An index at or
// past `records` reads zeros instead of leaving the array.
typedef float sph_v4f __attribute__((ext_vector_type(4)));
__device__ sph_v4f sph_struct_load_v4f32(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset,
                                         int soffset, int aux)
    __asm("llvm.amdgcn.struct.ptr.buffer.load.v4f32");
// Word 3 of a gfx9-family buffer descriptor: 32-bit data format, no swizzle.
constexpr int kBufferWord3 = 0x00020000;

struct Records16 {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ Records16(const float4 *a, uint32_t records)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<float4 *>(a), (short)16, (int)records,
                                            kBufferWord3)) {}
  __device__ __forceinline__ float4 operator[](uint32_t j) const {
    const sph_v4f v = sph_struct_load_v4f32(r, (int)j, 0, 0, 0);
    return make_float4(v.x, v.y, v.z, v.w);
  }
};

__device__ __forceinline__ float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// x^2 + y^2 + z^2 with two fused multiply-adds (this file is compiled without
// implicit contraction, so every fma is written out).
__device__ __forceinline__ float sq3(float x, float y, float z) {
  return fmaf(z, z, fmaf(y, y, x * x));
}

// ApplyBoundaryConditions (particles.c:842-867).
__device__ __forceinline__ float clamp_axis(float x, float lo, float hi, float hi_clamp) {
  if (x < lo) return lo;
  if (x >= hi) return hi_clamp;
  return x;
}

// Box and hash grid of the scene a particle belongs to (the context's own for
// a single scene; the scene table's in a batch).  Boxes start at the origin.
struct Geo {
  uint32_t base, nx, ny, nz, nxy;
  float hi[3], hi_clamp[3];
};

template <bool BATCH>
__device__ __forceinline__ Geo geo_of(uint32_t i, const DevParams &P, const BatchDev &B) {
  Geo g;
  if (!BATCH) {
    g.base = 0u;
    g.nx = P.nx; g.ny = P.ny; g.nz = P.nz; g.nxy = P.nxy;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      g.hi[a] = P.box_hi[a];
      g.hi_clamp[a] = P.box_hi_clamp[a];
    }
  } else {
    const SceneDev &d = B.scenes[B.scene_of[i]];
    g.base = d.cell_base;
    g.nx = d.nx; g.ny = d.ny; g.nz = d.nz; g.nxy = d.nxy;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      g.hi[a] = d.box_hi[a];
      g.hi_clamp[a] = d.box_hi_clamp[a];
    }
  }
  return g;
}

// Obstacle of particle i's scene.
template <bool BATCH>
__device__ __forceinline__ const SdfDev &sdf_of(uint32_t i, const SdfDev &o, const BatchDev &B) {
  if (!BATCH) return o;
  return B.scenes[B.scene_of[i]].sdf;
}

__device__ __forceinline__ void apply_boundary(float &x, float &y, float &z, const Geo &g,
                                               const DevParams &P) {
  x = clamp_axis(x, P.box_lo[0], g.hi[0], g.hi_clamp[0]);
  y = clamp_axis(y, P.box_lo[1], g.hi[1], g.hi_clamp[1]);
  z = clamp_axis(z, P.box_lo[2], g.hi[2], g.hi_clamp[2]);
}

__device__ __forceinline__ void apply_boundary(float &x, float &y, float &z,
                                               const DevParams &P) {
  x = clamp_axis(x, P.box_lo[0], P.box_hi[0], P.box_hi_clamp[0]);
  y = clamp_axis(y, P.box_lo[1], P.box_hi[1], P.box_hi_clamp[1]);
  z = clamp_axis(z, P.box_lo[2], P.box_hi[2], P.box_hi_clamp[2]);
}

// Exact floor(x / h) for float x, h: a float estimate corrected with the sign
// of the exactly rounded residual fma(-k, h, x).  Equals the reference's
// floor of the double quotient (neighbors.c:86-88), which never rounds across
// an integer for float operands.
__device__ __forceinline__ uint32_t cell_coord(float x, float h, float inv_h, uint32_t dim,
                                               uint32_t *clamped) {
  float k = floorf(x * inv_h);
  if (fmaf(-k, h, x) < 0.0f)
    k -= 1.0f;
  else if (fmaf(-(k + 1.0f), h, x) >= 0.0f)
    k += 1.0f;
  if (!(k >= 0.0f)) {  // negative or NaN: keep the index inside the grid
    *clamped += 1u;
    return 0u;
  }
  if (k > (float)(dim - 1)) {
    *clamped += 1u;
    return dim - 1;
  }
  return (uint32_t)k;
}

// HashVal (neighbors.c:79-98), offset by the scene's first cell key.
__device__ __forceinline__ uint32_t hash_key(float x, float y, float z, const DevParams &P,
                                             const Geo &g, uint32_t *clamped) {
  const uint32_t gx = cell_coord(x, P.h, P.inv_h, g.nx, clamped);
  const uint32_t gy = cell_coord(y, P.h, P.inv_h, g.ny, clamped);
  const uint32_t gz = cell_coord(z, P.h, P.inv_h, g.nz, clamped);
  return g.base + (g.nxy * gz + (gy * g.nx + gx));
}

__device__ __forceinline__ void add_counter64(uint32_t *stats, int slot, uint32_t v) {
  if (v) atomicAdd(reinterpret_cast<unsigned long long *>(stats + slot), (unsigned long long)v);
}

// The same with one atomic per wavefront (its lanes' sum); every lane of the
// wave must call it.  The list build's drop count is nonzero for every
// particle at the 60-neighbour cap (tens of thousands in a late dam-break
// substep), and per-lane atomics on one word serialise.
__device__ __forceinline__ void add_counter64_wave(uint32_t *stats, int slot, uint32_t v) {
  if (__ballot(v != 0u) == 0ull) return;
  uint32_t t = v;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o, 64);
  if ((threadIdx.x & 63) == 0)
    atomicAdd(reinterpret_cast<unsigned long long *>(stats + slot), (unsigned long long)t);
}

// Block sum of v written to partial[blockIdx.x] (no atomics; the host adds the
// partials when it reads the statistics).  All threads of the block call it.
__device__ __forceinline__ void block_sum_to(uint32_t *partial, uint32_t v) {
  __shared__ uint32_t wsum[kBlock / kWave];
  uint32_t t = v;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
    for (uint32_t w = 0; w < blockDim.x / kWave; ++w) s += wsum[w];
    partial[blockIdx.x] = s;
  }
}

// ---------------------------------------------------------------------------
// Elementwise stages before the neighbour search
// ---------------------------------------------------------------------------
// ApplyGravity (particles.c:516-527).
__global__ __launch_bounds__(kBlock) void k_gravity(uint32_t n, float4 *__restrict__ v,
                                                    DevParams P) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  v[i].z = v[i].z + P.g_dt;
}

// PredictPositions (particles.c:748-784).
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_predict(uint32_t n, const float4 *__restrict__ x,
                                                    const float4 *__restrict__ v,
                                                    float4 *__restrict__ xs, DevParams P,
                                                    BatchDev B) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 a = x[i], u = v[i];
  float px = a.x + u.x * P.dt, py = a.y + u.y * P.dt, pz = a.z + u.z * P.dt;
  apply_boundary(px, py, pz, geo_of<BATCH>(i, P, B), P);
  xs[i] = make_float4(px, py, pz, 0.0f);
}

// HashParticles (neighbors.c:101-125).
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_hash(uint32_t n, const float4 *__restrict__ xs,
                                                 uint32_t *__restrict__ keys, DevParams P,
                                                 uint32_t *__restrict__ stats, BatchDev B) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 a = xs[i];
  uint32_t clamped = 0;
  const uint32_t key = hash_key(a.x, a.y, a.z, P, geo_of<BATCH>(i, P, B), &clamped);
  keys[i] = key;
  add_counter64(stats, STAT_CLAMPS, clamped);
}

// Fused ApplyGravity + PredictPositions + HashParticles for the fused
// substep: v after gravity is only read by the prediction (UpdateVelocities
// overwrites it), so it is not stored.
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_predict_hash(uint32_t n,
                                                         const float4 *__restrict__ x,
                                                         const float4 *__restrict__ v,
                                                         float4 *__restrict__ xs,
                                                         uint32_t *__restrict__ keys,
                                                         DevParams P,
                                                         uint32_t *__restrict__ stats,
                                                         BatchDev B) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 a = x[i], u = v[i];
  const float vz = u.z + P.g_dt;
  float px = a.x + u.x * P.dt, py = a.y + u.y * P.dt, pz = a.z + vz * P.dt;
  const Geo g = geo_of<BATCH>(i, P, B);
  apply_boundary(px, py, pz, g, P);
  xs[i] = make_float4(px, py, pz, 0.0f);
  uint32_t clamped = 0;
  const uint32_t key = hash_key(px, py, pz, P, g, &clamped);
  keys[i] = key;
  add_counter64(stats, STAT_CLAMPS, clamped);
}

// Gather x* and x into (cell key, id) order, and the cell ends for the cell
// tables from the sorted keys: the last particle of each run of equal keys
// writes end[key] = its index + 1 (no atomics; the entries of empty cells stay
// zero, as the scan leaves them).  The lower bound of cell c is then the
// largest end of the cells before it (an exclusive max-scan).
__global__ __launch_bounds__(kBlock) void k_reorder(uint32_t n, const uint32_t *__restrict__ id_s,
                                                    const uint32_t *__restrict__ key_s,
                                                    const float4 *__restrict__ xs_can,
                                                    const float4 *__restrict__ x_can,
                                                    float4 *__restrict__ pos,
                                                    float4 *__restrict__ xold,
                                                    uint32_t *__restrict__ cell_end,
                                                    float *__restrict__ sx, uint32_t stride) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const uint32_t id = id_s[s];
  const float4 q = xs_can[id];
  pos[s] = q;
  sx[s] = q.x;
  sx[stride + s] = q.y;
  sx[2u * stride + s] = q.z;
  xold[s] = x_can[id];
  const uint32_t k = key_s[s];
  if (s + 1 == n || key_s[s + 1] != k) cell_end[k] = s + 1;
}

// ---------------------------------------------------------------------------
// Neighbour lists: FillNeighbors (neighbors.c:217-298)
// ---------------------------------------------------------------------------
// r^2 < h^2 evaluated exactly as the reference does it, in double with the
// reference's association and no contraction (neighbors.c:278-285).
__device__ __noinline__ bool inside_exact(float px, float py, float pz, float qx, float qy,
                                          float qz, double h2d) {
#pragma clang fp contract(off)
  const double dx = (double)px - (double)qx;
  const double dy = (double)py - (double)qy;
  const double dz = (double)pz - (double)qz;
  const double r2 = dx * dx + dy * dy + dz * dz;
  return r2 < h2d;
}

// Four consecutive floats from a 4-byte-aligned address (one 16-byte load).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ f4u ld4u(const float *p) { return *reinterpret_cast<const f4u *>(p); }

// One candidate of the list build: each mask takes its compare as the next
// bit, m = 2 m + [r2 < bound], one add-with-carry with the compare in VCC as
// the carry (two VALU per mask, where the compiler's select, shift and or take
// two and a half).  Candidates are fed last to first, so candidate t ends at
// bit t.
__device__ __forceinline__ void candidate_bits(uint32_t &lo, uint32_t &hi, float r2,
                                               float h2_lo, float h2_hi) {
  asm("v_cmp_gt_f32_e32 vcc, %[hl], %[r]\n\t"
      "v_addc_co_u32_e32 %[lo], vcc, %[lo], %[lo], vcc\n\t"
      "v_cmp_gt_f32_e32 vcc, %[hh], %[r]\n\t"
      "v_addc_co_u32_e32 %[hi], vcc, %[hi], %[hi], vcc"
      : [lo] "+v"(lo), [hi] "+v"(hi)
      : [r] "v"(r2), [hl] "s"(h2_lo), [hh] "s"(h2_hi)
      : "vcc");
}

// One thread per sorted particle.  The 3 cells (dx = -1..1) of each (dz, dy)
// row are consecutive keys, so each row is one contiguous range of sorted
// indices; walking rows in (dz, dy) order and indices upward reproduces the
// reference's (dz, dy, dx, id) order and therefore its 60-cap truncation.
//
// Hits are buffered in LDS as 16-bit offsets from the first candidate of
// their dz layer and written out after the scan as compact quads (8 entries
// each): a wave then stores 1 KB of the list contiguously per instruction.
// Written as found, lanes store their k-th neighbour at different times and
// partially written lines are written back several times.  A wavefront in
// which some particle's layer spans more than kCompactSpan indices (never
// near rest density: a layer holds about two grid rows of particles), or
// every wavefront when `force_wide` is set, writes 32-bit indices directly.
// The head records the layer bases (the particle itself for a layer without
// candidates, so that padding entries decode to a valid index) and counts.
template <bool BATCH>
__global__ __launch_bounds__(kBlock, 7) void k_build_lists(
    uint32_t n, uint32_t nblocks, const float4 *__restrict__ pos,
    const float *__restrict__ sx, uint32_t stride,
    const uint32_t *__restrict__ key_s, const uint32_t *__restrict__ cstart, DevParams P,
    uint32_t *__restrict__ list, uint4 *__restrict__ head, uint32_t *__restrict__ stats,
    uint32_t *__restrict__ pair_partial, int force_wide, BatchDev B) {
  extern __shared__ uint16_t sbuf[];  // [kLdsHits][blockDim.x]
  const uint32_t s = particle_index(nblocks);
  int c = 0;
  uint32_t dropped = 0;
  bool wide = false;
  uint32_t *out = list + (size_t)(s >> 6) * kListStride + (s & 63u) * 4u;
  if (s < n) {
    const float4 p = pos[s];
    const Geo G = geo_of<BATCH>(s, P, B);
    const uint32_t key = key_s[s] - G.base;
    const uint32_t cz = key / G.nxy;
    const uint32_t rem = key - cz * G.nxy;
    const uint32_t cy = rem / G.nx;
    const uint32_t cx = rem - cy * G.nx;
    const uint32_t x0 = cx > 0 ? cx - 1 : 0;
    const uint32_t x1 = cx + 1 < G.nx ? cx + 1 : G.nx - 1;
    const uint32_t y0 = cy > 0 ? cy - 1 : 0;
    const uint32_t y1 = cy + 1 < G.ny ? cy + 1 : G.ny - 1;
    uint32_t base[3] = {s, s, s};
    bool direct = force_wide != 0;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const int z = (int)cz + l - 1;
      if (z < 0 || z >= (int)G.nz) continue;
      const uint32_t zrow = G.base + (uint32_t)z * G.nxy;
      const uint32_t first = cstart[zrow + y0 * G.nx + x0];
      const uint32_t span = cstart[zrow + y1 * G.nx + x1 + 1] - first;
      if (span > 0u) base[l] = first;
      if (span > kCompactSpan) direct = true;
    }
    // the format is per wavefront (the lanes' quads interleave)
    wide = __ballot(direct) != 0ull;
    int cl[3] = {0, 0, 0};
    uint32_t slot = threadIdx.x;  // c * blockDim.x + threadIdx.x
    for (int dz = -1; dz <= 1; ++dz) {
      const int z = (int)cz + dz;
      if (z >= 0 && z < (int)G.nz) {
        const uint32_t lb = dz < 0 ? base[0] : (dz == 0 ? base[1] : base[2]);
        // the layer's three row ranges in one round trip (a row outside the
        // grid is read at the particle's own row and left empty)
        uint32_t r0[3], r1[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int y = (int)cy + k - 1;
          const bool in = y >= 0 && y < (int)G.ny;
          const uint32_t row = G.base + (uint32_t)z * G.nxy + (in ? (uint32_t)y : cy) * G.nx;
          r0[k] = cstart[row + x0];
          r1[k] = cstart[row + x1 + 1];
          if (!in) r1[k] = r0[k];
        }
        for (int k = 0; k < 3; ++k) {
          const uint32_t j0 = k == 0 ? r0[0] : (k == 1 ? r0[1] : r0[2]);
          const uint32_t j1 = k == 0 ? r1[0] : (k == 1 ? r1[1] : r1[2]);
          // candidates 32 at a time: test all (8 loads in flight per group),
          // then append the hits of the chunk in j order
          for (uint32_t cb = j0; cb < j1; cb += 32) {
            const uint32_t len = j1 - cb;
            uint32_t lo = 0u, hi = 0u;  // bits: r2 < h2_lo, r2 < h2_hi
            // groups last to first: a lane skips a suffix of the groups, which
            // then come first, while its masks are still 0
#pragma unroll
            for (int g = 3; g >= 0; --g) {
              if (8u * g < len) {
                // 8 candidates as two 16-byte loads per coordinate array:
                // a wave's lanes (about 8 cells) then touch 2-3 cache lines per
                // load instead of about 8 for 16-byte records.  Reads up to 7
                // entries past j1: each array is padded by kPosPad readable
                // entries beyond the capacity.
                const float *X = sx + cb + 8 * g;
                const f4u x0 = ld4u(X), x1 = ld4u(X + 4);
                const f4u y0 = ld4u(X + stride), y1 = ld4u(X + stride + 4);
                const f4u z0 = ld4u(X + 2u * stride), z1 = ld4u(X + 2u * stride + 4);
                const float qx[kBatch] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
                const float qy[kBatch] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w};
                const float qz[kBatch] = {z0.x, z0.y, z0.z, z0.w, z1.x, z1.y, z1.z, z1.w};
#pragma unroll
                for (int t = kBatch - 1; t >= 0; --t) {
                  const float r2 = sq3(p.x - qx[t], p.y - qy[t], p.z - qz[t]);
                  candidate_bits(lo, hi, r2, P.h2_lo, P.h2_hi);
                }
              }
            }
            uint32_t keep = len >= 32u ? ~0u : (1u << len) - 1u;
            if (s - cb < 32u) keep &= ~(1u << (s - cb));  // not the particle itself
            lo &= keep;
            uint32_t band = hi & keep & ~lo;  // within the guard band of h^2
            while (band) {
              const int t = __builtin_ctz(band);
              band &= band - 1u;
              const float4 q = pos[cb + t];
              if (inside_exact(p.x, p.y, p.z, q.x, q.y, q.z, P.h2d)) lo |= 1u << t;
            }
            // append the chunk's hits in j order; the cap and the end of the
            // LDS buffer are tested per hit only when this chunk can reach
            // them (`slot` = the LDS index of entry c, kept as a running sum)
            const int nh = __builtin_popcount(lo);
            if (wide && c + nh <= kMaxNb) {
              while (lo) {
                const uint32_t j = cb + (uint32_t)__builtin_ctz(lo);
                lo &= lo - 1u;
                out[(size_t)(c >> 2) * kQuadStride + (c & 3)] = j;
                ++c;
              }
            } else if (!wide && c + nh <= kLdsHits) {
              while (lo) {
                const uint32_t j = cb + (uint32_t)__builtin_ctz(lo);
                lo &= lo - 1u;
                sbuf[slot] = (uint16_t)(j - lb);
                slot += blockDim.x;
              }
              c += nh;
            } else {
              while (lo) {
                const uint32_t j = cb + (uint32_t)__builtin_ctz(lo);
                lo &= lo - 1u;
                if (c < kMaxNb) {
                  if (wide)
                    out[(size_t)(c >> 2) * kQuadStride + (c & 3)] = j;
                  else if (c < kLdsHits)
                    sbuf[slot] = (uint16_t)(j - lb);
                  else
                    store_entry16(out, c, j - lb);
                  slot += blockDim.x;
                  ++c;
                } else {
                  ++dropped;
                }
              }
            }
          }
        }
      }
      cl[dz + 1] = c;
    }
    head[s] = make_uint4(base[0], base[1], base[2],
                         (uint32_t)c | ((uint32_t)cl[0] << 8) | ((uint32_t)cl[1] << 16) |
                             (wide ? kHeadWide : 0u));
    // the rest of a directly stored last quad is padding: offset 0 (readers
    // decode whole quads, and a stale offset could index past the particles)
    if (!wide && c > kLdsHits)
      for (int e = c; (e & 7) != 0; ++e) store_entry16(out, e, 0u);
  }
  // coalesced write-out of compact quads: every lane of the wave walks the
  // wave's longest list one quad (8 entries) at a time, up to the end of the
  // LDS buffer (later quads were stored directly)
  int cmax = wide ? 0 : c;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) cmax = max(cmax, __shfl_xor(cmax, o, 64));
  for (int k = 0; k < cmax && k < kLdsHits; k += 8) {
    if (k < c) {
      uint32_t w[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const uint32_t lo = (k + 2 * t < c) ? sbuf[(k + 2 * t) * blockDim.x + threadIdx.x] : 0u;
        const uint32_t hi =
            (k + 2 * t + 1 < c) ? sbuf[(k + 2 * t + 1) * blockDim.x + threadIdx.x] : 0u;
        w[t] = lo | (hi << 16);
      }
      *reinterpret_cast<uint4 *>(out + (size_t)(k >> 3) * kQuadStride) =
          make_uint4(w[0], w[1], w[2], w[3]);
    }
  }
  add_counter64_wave(stats, STAT_DROPPED, dropped);
  block_sum_to(pair_partial, (uint32_t)c);
}

// ---------------------------------------------------------------------------
// Jacobi constraint loop
// ---------------------------------------------------------------------------
// Max of a non-negative float over a block, then at most one atomic on the
// float's bit pattern (monotonic for non-negative floats).  The slot only
// grows during a launch, so a block whose maximum does not exceed the value it
// reads skips the atomic; without this every block would issue a device-scope
// atomic to one address.  Every thread of the block must call it.  With one
// read and atomic per wave instead, each wave waited for that round trip at
// its end: per block, the density pass measured 1.6% faster at 8M particles.
__device__ __forceinline__ void block_max_to(uint32_t *slot, float v, bool active) {
  __shared__ float bm[kBlock / kWave];
  float m = active ? v : 0.0f;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) bm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < blockDim.x / kWave; ++w) m = fmaxf(m, bm[w]);
    if (m > 0.0f) {
      const uint32_t bits = __float_as_uint(m);
      if (bits > *(volatile uint32_t *)slot) atomicMax(slot, bits);
    }
  }
}

__device__ __forceinline__ bool converged(const uint32_t *stats) {
  return stats[STAT_CONVERGED] != 0u;
}

// One neighbour's density term: the integer (k3 (h^2 - r^2))^3, truncated
// (W(r) = norm (h^2 - r^2)^3, 0 for r >= h).  u and u^3 are computed in
// double (half the FP32 rate on this target, a few VALU slots more per
// neighbour): in FP32 their roundings (about 4e-7 of a term, independent
// between terms) moved C = rho/rho0 - 1 by up to 2e-7, which near rest,
// where C itself is 1e-3..1e-2, is lambda's largest error.  Sums of at most
// 60 terms of at most 2^26 fit 32 bits; the sum is exact and independent of
// the order of the terms.
__device__ __forceinline__ uint32_t density_term(float r2, const DevParams &P) {
  const double u = fmax(fma(-(double)r2, P.dens_k3, P.dens_c1), 0.0);
  return (uint32_t)((u * u) * u);
}

// C = rho/rho0 - 1 from the fixed-point density sum S of cnt terms.  Each
// truncation drops half a unit on average, restored by cnt/2.  S is split
// into 20 high and 12 low bits, each exact in FP32, and C = a S + c0 is
// evaluated with a and c0 as float pairs: near rest C ~ 1e-3 is the
// difference of two terms ~0.8, and this keeps its error at the FP32
// rounding of the terms instead of that of rho ~ 1e3.
__device__ __forceinline__ float density_constraint(uint32_t S, int cnt, const DevParams &P) {
  S += (uint32_t)cnt >> 1;
  const float shi = (float)(S & 0xFFFFF000u), slo = (float)(S & 0xFFFu);
  return fmaf(P.dens_a_hi, shi, P.dens_c0_hi) +
         (P.dens_c0_lo + fmaf(P.dens_a_lo, shi, P.dens_a_hi * slo));
}

// ComputeDensities (particles.c:461-514) and ComputeLambda (particles.c:588-666)
// in one gather: rho_i needs only r^2, lambda_i needs C_i (own) and the
// spiky gradients.  Writes rho and pos.w = lambda.
// MODE 0: both, 1: density only, 2: lambda only (C from c_io).  With c_io,
// MODE 0/1 also store C there; with lam_out, lambda goes there and pos.w
// keeps rho (density_to_w).
template <int MODE, int NB>
__global__ GATHER_BOUNDS(NB) void k_density_lambda(
    uint32_t n, uint32_t nblocks, float4 *pos, const uint32_t *__restrict__ list,
    const uint4 *__restrict__ head, DevParams P, float *__restrict__ rho_out, float *c_io,
    uint32_t *stats, int err_slot, int check_converged, int density_to_w,
    float *__restrict__ lam_out) {
  if (check_converged && converged(stats)) return;
  const uint32_t s = particle_index(nblocks);
  const bool active = s < n;
  float err = 0.0f;
  if (active) {
    const float4 p = pos[s];
    const NbList nl = nb_list(list, head, s);
    uint32_t S = 0u;
    float sgx = 0.0f, sgy = 0.0f, sgz = 0.0f, sc = 0.0f;
    const Records16 G(pos, n);
    for_neighbors<float4, NB, true, true>(nl, s, [&](uint32_t j) { return G[j]; },
                                          [&](const float4 &q) {
      const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      const float r2 = sq3(dx, dy, dz);
      if (MODE != 2) S += density_term(r2, P);
      if (MODE != 1) {
        const float r2c = fmaxf(r2, P.rmin2);
        const float ir = fast_rsq(r2c);
        const float hr = fmaf(-r2c, ir, P.h);  // h - r, r floored at 1e-4 h
        const float hp = fmaxf(hr, 0.0f);  // a listed pair can be past h after an update
        const float g = hp * hp * ir;  // DelW(r)/r / norm
        sgx = fmaf(g, dx, sgx);
        sgy = fmaf(g, dy, sgy);
        sgz = fmaf(g, dz, sgz);
        sc = fmaf(g * g, r2, sc);
      }
    });
    float C;
    if (MODE != 2) {
      C = density_constraint(S, nl.c, P);
      const float rho = fmaf(C, P.rest_density, P.rest_density);
      rho_out[s] = rho;
      if (c_io) c_io[s] = C;
      // the reported error is that of the stored density, as every later
      // reader of rho (per-scene errors, the warm fallback) computes it
      err = fmaxf((rho - P.rest_density) * P.inv_rest_density, 0.0f);
      if (density_to_w) pos[s].w = rho;
    } else {
      C = c_io[s];
    }
    if (MODE != 1) {
      const float Ci = fmaxf(C, 0.0f);
      const float sum_c = P.lam_k2 * (sc + sq3(sgx, sgy, sgz));
      const float lam = -Ci / (sum_c + P.lambda_eps);
      if (lam_out)
        lam_out[s] = lam;  // pos.w keeps rho (density_to_w)
      else
        pos[s].w = lam;
    }
  }
  // every thread gets here: the only early return above is uniform
  if (MODE != 2 && err_slot >= 0) block_max_to(&stats[STAT_ERR0 + err_slot], err, active);
}

// Trilinear sample with normalised coordinates, texel centres at (i+0.5)/dim
// and clamp addressing, as tex3D with linear filtering (obstacle_cuda.cu:
// 174-180) defines it, with exact FP32 weights.
__device__ __forceinline__ float sdf_sample(const SdfDev &o, float u, float v, float w) {
  const float cu = u * (float)o.dim[0] - 0.5f;
  const float cv = v * (float)o.dim[1] - 0.5f;
  const float cw = w * (float)o.dim[2] - 0.5f;
  const float fu = floorf(cu), fv = floorf(cv), fw = floorf(cw);
  const float au = cu - fu, av = cv - fv, aw = cw - fw;
  const int iu = (int)fu, iv = (int)fv, iw = (int)fw;
  const int u0 = min(max(iu, 0), o.dim[0] - 1), u1 = min(max(iu + 1, 0), o.dim[0] - 1);
  const int v0 = min(max(iv, 0), o.dim[1] - 1), v1 = min(max(iv + 1, 0), o.dim[1] - 1);
  const int w0 = min(max(iw, 0), o.dim[2] - 1), w1 = min(max(iw + 1, 0), o.dim[2] - 1);
  const size_t sy = (size_t)o.dim[0], sz = (size_t)o.dim[0] * (size_t)o.dim[1];
  const float *d = o.data;
  const float c000 = d[u0 + v0 * sy + w0 * sz], c100 = d[u1 + v0 * sy + w0 * sz];
  const float c010 = d[u0 + v1 * sy + w0 * sz], c110 = d[u1 + v1 * sy + w0 * sz];
  const float c001 = d[u0 + v0 * sy + w1 * sz], c101 = d[u1 + v0 * sy + w1 * sz];
  const float c011 = d[u0 + v1 * sy + w1 * sz], c111 = d[u1 + v1 * sy + w1 * sz];
  const float c00 = c000 + au * (c100 - c000), c10 = c010 + au * (c110 - c010);
  const float c01 = c001 + au * (c101 - c001), c11 = c011 + au * (c111 - c011);
  const float c0 = c00 + av * (c10 - c00), c1 = c01 + av * (c11 - c01);
  return c0 + aw * (c1 - c0);
}

// The push of CalculateParticleCollisions_kernel (obstacle_cuda.cu:81-113)
// at normalised grid coordinates (u, v, w): false outside the obstacle,
// otherwise the world-space push along the SDF gradient.
__device__ __forceinline__ bool sdf_push(const SdfDev &o, float u, float v, float w,
                                         float &px, float &py, float &pz) {
  const float d = sdf_sample(o, u, v, w);
  if (!(d < 0.0f)) return false;
  const float gx = sdf_sample(o, u + o.inv_dim[0], v, w) - sdf_sample(o, u - o.inv_dim[0], v, w);
  const float gy = sdf_sample(o, u, v + o.inv_dim[1], w) - sdf_sample(o, u, v - o.inv_dim[1], w);
  const float gz = sdf_sample(o, u, v, w + o.inv_dim[2]) - sdf_sample(o, u, v, w - o.inv_dim[2]);
  const float dd = gx * gx + gy * gy + gz * gz;
  if (!(dd > 0.0f)) return false;  // the reference would normalise a zero vector to NaN
  const float inv = 1.0f / sqrtf(dd);
  const float dist = fabsf(d) * o.m2w;
  px = dist * (gx * inv);
  py = dist * (gy * inv);
  pz = dist * (gz * inv);
  return true;
}

// CalculateParticleCollisions_kernel (obstacle_cuda.cu:56-116).
__device__ __forceinline__ void collide(float &x, float &y, float &z, const SdfDev &o) {
  if (!o.data) return;
  if (!(x > o.wmin[0] && x < o.wmax[0] && y > o.wmin[1] && y < o.wmax[1] && z > o.wmin[2] &&
        z < o.wmax[2]))
    return;
  const float u = (x - o.wmin[0]) / o.len[0];
  const float v = (y - o.wmin[1]) / o.len[1];
  const float w = (z - o.wmin[2]) / o.len[2];
  float px, py, pz;
  if (!sdf_push(o, u, v, w, px, py, pz)) return;
  y += py;
  x += px;
  z += pz;
}

// UpdateDPs (particles.c:668-746): returns sum_j (lambda_i + lambda_j) DelW(r)/r d.
template <int NB>
__device__ __forceinline__ float3 dp_sum(const float4 &p, const float4 *pos, uint32_t n,
                                         const NbList &nl, uint32_t self, const DevParams &P) {
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  const Records16 G(pos, n);
  for_neighbors<float4, NB>(nl, self, [&](uint32_t j) { return G[j]; }, [&](const float4 &q) {
    const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    const float r2c = fmaxf(sq3(dx, dy, dz), P.rmin2);
    const float ir = fast_rsq(r2c);
    const float hr = fmaf(-r2c, ir, P.h);
    const float hp = fmaxf(hr, 0.0f);
    const float g = hp * hp * ir;
    const float f = (p.w + q.w) * g;
    ax = fmaf(f, dx, ax);
    ay = fmaf(f, dy, ay);
    az = fmaf(f, dz, az);
  });
  return make_float3(ax, ay, az);
}

// Staged UpdateDPs: dp = stiffness * m * DelW_norm / rho0 * sum.
// With xnext, also the position update of k_position_stars from the same dp
// (x* + dp, clamp, collision, rho into w), into another buffer.
template <bool BATCH, int NB>
__global__ GATHER_BOUNDS(NB) void k_dp(uint32_t n, uint32_t nblocks,
                                               const float4 *__restrict__ pos,
                                               const uint32_t *__restrict__ list,
                                               const uint4 *__restrict__ head, DevParams P,
                                               float scale, float4 *__restrict__ dp,
                                               float4 *__restrict__ xnext,
                                               const float *__restrict__ rho, SdfDev o,
                                               BatchDev B) {
  const uint32_t s = particle_index(nblocks);
  if (s >= n) return;
  const float4 p = pos[s];
  const float3 a = dp_sum<NB>(p, pos, n, nb_list(list, head, s), s, P);
  const float4 d = make_float4(scale * a.x, scale * a.y, scale * a.z, 0.0f);
  dp[s] = d;
  if (xnext) {
    float x = p.x + d.x, y = p.y + d.y, z = p.z + d.z;
    apply_boundary(x, y, z, geo_of<BATCH>(s, P, B), P);
    collide(x, y, z, sdf_of<BATCH>(s, o, B));
    xnext[s] = make_float4(x, y, z, rho[s]);
  }
}

// Staged UpdatePositionStars (particles.c:529-564): x* += dp, clamp, collide.
// Also stores rho in w for the velocity passes.
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_position_stars(uint32_t n, float4 *__restrict__ pos,
                                                           const float4 *__restrict__ dp,
                                                           const float *__restrict__ rho,
                                                           DevParams P, SdfDev o, BatchDev B) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const float4 p = pos[s], d = dp[s];
  float x = p.x + d.x, y = p.y + d.y, z = p.z + d.z;
  apply_boundary(x, y, z, geo_of<BATCH>(s, P, B), P);
  collide(x, y, z, sdf_of<BATCH>(s, o, B));
  pos[s] = make_float4(x, y, z, rho[s]);
}

// Fused UpdateDPs + UpdatePositionStars: reads pin (x*, lambda), writes pout
// (x*', rho).  COLLIDE false leaves the obstacle to k_collide_sorted: the
// inlined trilinear gradient doubles the kernel's registers (109 vs 58 VGPRs)
// and halves its occupancy, for a push that only particles inside the SDF box
// need.  In tolerance mode, once the density pass of this iteration is
// within tolerance, the particles are copied unchanged instead.  With lam_acc,
// adds stiff * lambda_i to the particle's accumulated multiplier.
template <bool COLLIDE, bool BATCH, int NB>
__global__ GATHER_BOUNDS(NB) void k_dp_update(
    uint32_t n, uint32_t nblocks, const float4 *__restrict__ pin, float4 *__restrict__ pout,
    const float *__restrict__ rho, const uint32_t *__restrict__ list,
    const uint4 *__restrict__ head, DevParams P, float scale, SdfDev o, uint32_t *stats,
    int err_slot, float tol, float *__restrict__ lam_acc, float stiff, BatchDev B) {
  bool done = false;
  if (tol > 0.0f) {
    done = converged(stats) ||
           __uint_as_float(stats[STAT_ERR0 + err_slot]) <= tol;
  }
  const uint32_t s = particle_index(nblocks);
  if (s >= n) return;
  const float4 p = pin[s];
  if (done) {
    if (s == 0) stats[STAT_CONVERGED] = 1u;
    pout[s] = make_float4(p.x, p.y, p.z, rho[s]);
    return;
  }
  if (s == 0) stats[STAT_ITERS] += 1u;
  const float3 a = dp_sum<NB>(p, pin, n, nb_list(list, head, s), s, P);
  float x = fmaf(scale, a.x, p.x), y = fmaf(scale, a.y, p.y), z = fmaf(scale, a.z, p.z);
  apply_boundary(x, y, z, geo_of<BATCH>(s, P, B), P);
  if (COLLIDE) collide(x, y, z, sdf_of<BATCH>(s, o, B));
  pout[s] = make_float4(x, y, z, rho[s]);
  if (lam_acc) lam_acc[s] = fmaf(stiff, p.w, lam_acc[s]);
}

// ---------------------------------------------------------------------------
// Learned warm start
// ---------------------------------------------------------------------------
// SDF distance (in h, clamped to [-4, 4]) and unit gradient at a point, with
// the sampling of the collision pass (obstacle_cuda.cu:33-53,88).
__device__ __forceinline__ void sdf_features(float x, float y, float z, const SdfDev &o,
                                             const DevParams &P, float *f) {
  f[0] = 4.0f;
  f[1] = f[2] = f[3] = 0.0f;
  if (!o.data) return;
  if (!(x > o.wmin[0] && x < o.wmax[0] && y > o.wmin[1] && y < o.wmax[1] && z > o.wmin[2] &&
        z < o.wmax[2]))
    return;
  const float u = (x - o.wmin[0]) / o.len[0];
  const float v = (y - o.wmin[1]) / o.len[1];
  const float w = (z - o.wmin[2]) / o.len[2];
  const float d = sdf_sample(o, u, v, w);
  f[0] = fminf(fmaxf(d * o.m2w * P.inv_h, -4.0f), 4.0f);
  const float gx = sdf_sample(o, u + o.inv_dim[0], v, w) - sdf_sample(o, u - o.inv_dim[0], v, w);
  const float gy = sdf_sample(o, u, v + o.inv_dim[1], w) - sdf_sample(o, u, v - o.inv_dim[1], w);
  const float gz = sdf_sample(o, u, v, w + o.inv_dim[2]) - sdf_sample(o, u, v, w - o.inv_dim[2]);
  const float dd = gx * gx + gy * gy + gz * gz;
  if (!(dd > 0.0f)) return;
  const float inv = 1.0f / sqrtf(dd);
  f[1] = gx * inv;
  f[2] = gy * inv;
  f[3] = gz * inv;
}

// First pass of a warm-started or logged solve.  ComputeDensities and
// ComputeLambda as in k_density_lambda, plus the network features, in one
// gather of (x*, x) per neighbour.  LOG: writes the feature rows and the
// plain lambda.  Otherwise runs the network on MFMA (BF16, FP32 accumulate),
// whose output R estimates what the later iterations of a converged solve add
// to the multiplier, and writes lambda_0 - scale * max(R, 0) / (|grad C|^2 + eps):
// the plain first step plus the predicted remainder.
// Either way the particle's accumulated multiplier restarts at zero.  Every
// lane of a wave reaches the network (no early return).
template <bool LOG, bool BATCH, int NB = 8>
__global__ __launch_bounds__(kBlock) void k_warm_lambda(
    uint32_t n, uint32_t nblocks, float4 *pos, const float4 *__restrict__ xold,
    const uint32_t *__restrict__ list, const uint4 *__restrict__ head, DevParams P,
    float *__restrict__ rho_io, uint32_t *stats, const uint32_t *__restrict__ id_s,
    const float *__restrict__ lam_prev, SdfDev o, const MlpPacked *__restrict__ net,
    float *__restrict__ lam_acc, float *__restrict__ feat, float2 *__restrict__ plain,
    BatchDev B) {
  const uint32_t s = particle_index(nblocks);
  const bool active = s < n;
  float f[kMlpPad];
#pragma unroll
  for (int i = 0; i < kMlpPad; ++i) f[i] = 0.0f;
  float err = 0.0f, lam0 = 0.0f, denom = 1.0f, rho = 0.0f;
  if (active) {
    const float4 p = pos[s];
    const float4 po = xold[s];
    const float ux = p.x - po.x, uy = p.y - po.y, uz = p.z - po.z;  // dt * v_i
    const NbList nl = nb_list(list, head, s);
    uint32_t S = 0u;
    float sgx = 0.0f, sgy = 0.0f, sgz = 0.0f, sc = 0.0f;
    float mx = 0.0f, my = 0.0f, mz = 0.0f, dv = 0.0f;
    const Records16 Gp(pos, n), Go(xold, n);
    auto load = [&](uint32_t j) { return Pair4{Gp[j], Go[j]}; };
    for_neighbors<Pair4, NB, false>(nl, s, load, [&](const Pair4 &qq) {
      const float4 &q = qq.a, &qo = qq.b;
      const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
      const float r2 = sq3(dx, dy, dz);
      S += density_term(r2, P);  // as k_density_lambda
      const float u = fmaxf(P.h2 - r2, 0.0f);
      const float w3 = u * u * u;
      mx = fmaf(-w3, dx, mx);
      my = fmaf(-w3, dy, my);
      mz = fmaf(-w3, dz, mz);
      const float r2c = fmaxf(r2, P.rmin2);
      const float ir = fast_rsq(r2c);
      const float hr = fmaf(-r2c, ir, P.h);
      const float hp = fmaxf(hr, 0.0f);
    const float g = hp * hp * ir;
      sgx = fmaf(g, dx, sgx);
      sgy = fmaf(g, dy, sgy);
      sgz = fmaf(g, dz, sgz);
      sc = fmaf(g * g, r2, sc);
      const float ex = (q.x - qo.x) - ux, ey = (q.y - qo.y) - uy, ez = (q.z - qo.z) - uz;
      dv = fmaf(g, fmaf(ez, dz, fmaf(ey, dy, ex * dx)), dv);
    });
    const float cr = density_constraint(S, nl.c, P);
    rho = fmaf(cr, P.rest_density, P.rest_density);
    rho_io[s] = rho;
    err = fmaxf((rho - P.rest_density) * P.inv_rest_density, 0.0f);  // as k_density_lambda
    const float gc2 = P.lam_k2 * (sc + sq3(sgx, sgy, sgz));
    denom = gc2 + P.lambda_eps;
    lam0 = -fmaxf(cr, 0.0f) / denom;  // as k_density_lambda
    f[0] = (float)nl.c * (1.0f / 32.0f);
    f[1] = cr;
    f[2] = -lam_prev[id_s[s]] * denom;
    f[3] = gc2 / denom;
    f[4] = P.div_k * dv;
    sdf_features(p.x, p.y, p.z, sdf_of<BATCH>(s, o, B), P, f + 5);
    const float mk = P.mWn / (rho * P.h);
    f[9] = mk * mx;
    f[10] = mk * my;
    f[11] = mk * mz;
  }
  block_max_to(&stats[STAT_ERR0], err, active);  // every thread gets here
  float lam = lam0;
  if (LOG) {
    if (active) {
      float4 *row = reinterpret_cast<float4 *>(feat + (size_t)s * kMlpPad);
      row[0] = make_float4(f[0], f[1], f[2], f[3]);
      row[1] = make_float4(f[4], f[5], f[6], f[7]);
      row[2] = make_float4(f[8], f[9], f[10], f[11]);
      row[3] = make_float4(denom, lam0, 0.0f, 0.0f);
    }
  } else {
    const float y = mlp_wave(f, net);
    const float ymax = net->norm[kNormYMax], step = net->norm[kNormStep];
    const float ce = ymax > 0.0f ? fminf(fmaxf(y, 0.0f), ymax) : fmaxf(y, 0.0f);
    lam = lam0 - (step > 0.0f ? step : 1.0f) * ce / denom;
    if (active) plain[s] = make_float2(lam0, rho);  // for the plain fallback
  }
  if (active) {
    pos[s].w = lam;
    lam_acc[s] = 0.0f;
  }
}

// Warm start from the previous substep's multiplier (no network), after a
// plain density + lambda pass (rho, pos.w = lambda_0): keeps (lambda_0, rho)
// for the plain fallback and starts the solve from `scale` times the
// particle's accumulated multiplier of the previous substep (id order).
__global__ __launch_bounds__(kBlock) void k_warm_prev(uint32_t n, float4 *__restrict__ pos,
                                                      const float *__restrict__ rho,
                                                      const uint32_t *__restrict__ id_s,
                                                      const float *__restrict__ lam_prev,
                                                      float scale, float2 *__restrict__ plain,
                                                      float *__restrict__ lam_acc) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  plain[s] = make_float2(pos[s].w, rho[s]);
  pos[s].w = scale * lam_prev[id_s[s]];
  lam_acc[s] = 0.0f;
}

// Plain fallback of a warm start, after the density pass of iteration 1: if
// the predicted step left the maximum density error above the error before it,
// the particles return to their positions before the step with the plain
// first-iteration lambda and density, and iteration 1 proceeds from there
// (its error slot then holds the pre-step error).
__global__ __launch_bounds__(kBlock) void k_warm_fallback(uint32_t n, uint32_t *stats,
                                                          const float4 *__restrict__ pre,
                                                          float4 *__restrict__ pos,
                                                          const float2 *__restrict__ plain,
                                                          float *__restrict__ rho,
                                                          float *__restrict__ lam_acc) {
  if (stats[STAT_CONVERGED]) return;
  const uint32_t e0 = stats[STAT_ERR0], e1 = stats[STAT_ERR0 + 1];
  if (e1 <= e0) return;  // non-negative floats compare as their bit patterns
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s < n) {
    const float4 p = pre[s];
    const float2 q = plain[s];
    pos[s] = make_float4(p.x, p.y, p.z, q.x);
    rho[s] = q.y;
    lam_acc[s] = 0.0f;
  }
  if (s == 0) {
    stats[STAT_FALLBACK] = 1u;
    stats[STAT_ERR0 + 1] = e0;
  }
}

// The same for a batch of scenes, scene by scene: a scene whose maximum error
// after the step (e1) exceeds the one before it (e0) is reverted.  A second
// kernel sets the iteration-1 error slot to the maximum over scenes of the
// error each scene continues from, and counts the reverted scenes.
__global__ __launch_bounds__(kBlock) void k_warm_fallback_scenes(
    uint32_t n, const uint32_t *stats, const uint16_t *__restrict__ scene_of,
    const uint32_t *__restrict__ e0, const uint32_t *__restrict__ e1,
    const float4 *__restrict__ pre, float4 *__restrict__ pos,
    const float2 *__restrict__ plain, float *__restrict__ rho, float *__restrict__ lam_acc) {
  if (stats[STAT_CONVERGED]) return;
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const uint32_t k = scene_of[s];
  if (e1[k] <= e0[k]) return;  // non-negative floats compare as their bit patterns
  const float4 p = pre[s];
  const float2 q = plain[s];
  pos[s] = make_float4(p.x, p.y, p.z, q.x);
  rho[s] = q.y;
  lam_acc[s] = 0.0f;
}

__global__ __launch_bounds__(kBlock) void k_warm_fallback_stats(uint32_t nscenes,
                                                                uint32_t *stats,
                                                                const uint32_t *__restrict__ e0,
                                                                const uint32_t *__restrict__ e1) {
  if (stats[STAT_CONVERGED]) return;
  uint32_t m = 0u, reverted = 0u;
  for (uint32_t k = threadIdx.x; k < nscenes; k += blockDim.x) {
    const bool back = e1[k] > e0[k];
    m = max(m, back ? e0[k] : e1[k]);
    reverted += back ? 1u : 0u;
  }
  __shared__ uint32_t sm[kBlock], sr[kBlock];
  sm[threadIdx.x] = m;
  sr[threadIdx.x] = reverted;
  __syncthreads();
  for (uint32_t o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      sm[threadIdx.x] = max(sm[threadIdx.x], sm[threadIdx.x + o]);
      sr[threadIdx.x] += sr[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    stats[STAT_ERR0 + 1] = sm[0];
    stats[STAT_FALLBACK] = sr[0];
  }
}

// Keeps each particle's accumulated multiplier in particle-id order.
__global__ __launch_bounds__(kBlock) void k_scatter_lambda(uint32_t n,
                                                           const uint32_t *__restrict__ id_s,
                                                           const float *__restrict__ lam_acc,
                                                           float *__restrict__ lam_prev) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s < n) lam_prev[id_s[s]] = lam_acc[s];
}

// Dataset rows of a logged substep: every `every`-th sorted particle's
// feature row and its target R = -(Lambda - lambda_0) * (|grad C|^2 + eps), the
// part of the accumulated multiplier after the first iteration.
__global__ __launch_bounds__(kBlock) void k_warm_samples(uint32_t count, uint32_t every,
                                                         const float *__restrict__ feat,
                                                         const float *__restrict__ lam_acc,
                                                         float *__restrict__ x,
                                                         float *__restrict__ t) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= count) return;
  const size_t s = (size_t)k * every;
  const float4 *src = reinterpret_cast<const float4 *>(feat + s * kMlpPad);
  float4 *dst = reinterpret_cast<float4 *>(x + (size_t)k * kMlpPad);
  const float4 r3 = src[3];
  dst[0] = src[0];
  dst[1] = src[1];
  dst[2] = src[2];
  dst[3] = r3;
  t[k] = -(lam_acc[s] - r3.y) * r3.x;
}

// ---------------------------------------------------------------------------
// Velocity stages
// ---------------------------------------------------------------------------
// UpdateVelocities (particles.c:787-838).
__device__ __forceinline__ float4 velocity_of(const float4 &p, const float4 &xo,
                                              const DevParams &P) {
  float vx = (p.x - xo.x) / P.dt, vy = (p.y - xo.y) / P.dt, vz = (p.z - xo.z) / P.dt;
  vx = fminf(fmaxf(vx, -P.v_max), P.v_max);
  vy = fminf(fmaxf(vy, -P.v_max), P.v_max);
  vz = fminf(fmaxf(vz, -P.v_max), P.v_max);
  return make_float4(vx, vy, vz, 0.0f);
}

__global__ __launch_bounds__(kBlock) void k_velocity(uint32_t n, const float4 *__restrict__ pos,
                                                     const float4 *__restrict__ xold,
                                                     float4 *__restrict__ vel, DevParams P) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  vel[s] = velocity_of(pos[s], xold[s], P);
}

// ComputeSurfaceTension colour pass (particles.c:94-140), optionally fused
// with UpdateVelocities (own particle only).  DelWPoly(r)/r = norm (h^2-r^2)^2
// with r floored at 1e-4 h.
template <bool WITH_VELOCITY, int NB>
__global__ GATHER_BOUNDS(NB) void k_color(uint32_t n, uint32_t nblocks,
                                                  const float4 *__restrict__ pos,
                                                  const float4 *__restrict__ xold,
                                                  const uint32_t *__restrict__ list,
                                                  const uint4 *__restrict__ head, DevParams P,
                                                  float4 *__restrict__ vel,
                                                  float4 *__restrict__ color) {
  const uint32_t s = particle_index(nblocks);
  if (s >= n) return;
  const float4 p = pos[s];
  if (WITH_VELOCITY) vel[s] = velocity_of(p, xold[s], P);
  const NbList nl = nb_list(list, head, s);
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  const Records16 G(pos, n);
  for_neighbors<float4, NB>(nl, s, [&](uint32_t j) { return G[j]; }, [&](const float4 &q) {
    const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    const float r2c = fmaxf(sq3(dx, dy, dz), P.rmin2);
    const float u = P.h2 - r2c;
    const float up = fmaxf(u, 0.0f);
    const float g = up * up * fast_rcp(q.w);
    ax = fmaf(g, dx, ax);
    ay = fmaf(g, dy, ay);
    az = fmaf(g, dz, az);
  });
  color[s] = make_float4(P.color_k * ax, P.color_k * ay, P.color_k * az, 0.0f);
}

// ComputeSurfaceTension force pass (particles.c:142-204): cohesion (Akinci
// spline C, particles.c:54-65) + curvature, v += dt F / rho_i in place.  The
// spline is computed without a branch per neighbour (the branch split the
// loop body and held 104 VGPRs at 8 gathers in flight, 89 without it); at 16
// gathers, bounded to 168 VGPRs (3 waves per SIMD, 20 bytes spilled), the
// pass measured 0.628 -> 0.561 ms at 8M particles.
template <int NB>
__global__ __launch_bounds__(kBlock, (NB) <= 16 ? 3 : 1) void k_surface(uint32_t n, uint32_t nblocks,
                                                    const float4 *__restrict__ pos,
                                                    const float4 *__restrict__ color,
                                                    const uint32_t *__restrict__ list,
                                                    const uint4 *__restrict__ head, DevParams P,
                                                    float4 *__restrict__ vel) {
  const uint32_t s = particle_index(nblocks);
  if (s >= n) return;
  const float4 p = pos[s];
  const float4 ci = color[s];
  const NbList nl = nb_list(list, head, s);
  const float hh = 0.5f * P.h;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
  const Records16 Gp(pos, n), Gc(color, n);
  auto load = [&](uint32_t j) { return Pair4{Gp[j], Gc[j]}; };
  for_neighbors<Pair4, NB, false>(nl, s, load, [&](const Pair4 &qc) {
    const float4 &q = qc.a, &cq = qc.b;
    const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    const float r2c = fmaxf(sq3(dx, dy, dz), P.rmin2);
    const float ir = fast_rsq(r2c);
    const float r = r2c * ir;
    // (h - r) r, and 0 for r > h without a branch: max(h - r, 0) is h - r
    // exactly for r <= h, and for r > h the spline is 0 either way
    const float t = fmaxf(P.h - r, 0.0f) * r;
    const float t3 = t * t * t;
    const float spline = r <= hh ? fmaf(2.0f, t3, -P.h6_64) : t3;
    const float coh = P.coh_k * spline * ir;
    const float K = 2.0f * P.rest_density * fast_rcp(p.w + q.w);
    fx = fmaf(K, fmaf(coh, dx, P.curv_k * (ci.x - cq.x)), fx);
    fy = fmaf(K, fmaf(coh, dy, P.curv_k * (ci.y - cq.y)), fy);
    fz = fmaf(K, fmaf(coh, dz, P.curv_k * (ci.z - cq.z)), fz);
  });
  const float k = P.dt / p.w;
  float4 v = vel[s];
  v.x = fmaf(k, fx, v.x);
  v.y = fmaf(k, fy, v.y);
  v.z = fmaf(k, fz, v.z);
  vel[s] = v;
}

// ApplyViscosity (particles.c:383-459), reading v_old and writing v_new.  The
// distance keeps the reference's typo r^2 = dx^2 + dy^2 dz^2 (particles.c:439).
// The only two-array pass with the list prefetch: at 16 gathers in flight it
// takes 175 VGPRs (2 waves per SIMD), bounded to 168 (3 waves, 32 bytes
// spilled) it ran 8% faster than without the prefetch at 8M particles; the
// surface and vorticity passes were 0-15% slower with it.
template <int NB>
__global__ __launch_bounds__(kBlock, (NB) <= 16 ? 3 : 1) void k_xsph(uint32_t n, uint32_t nblocks,
                                                 const float4 *__restrict__ pos,
                                                 const float4 *__restrict__ vin,
                                                 const uint32_t *__restrict__ list,
                                                 const uint4 *__restrict__ head, DevParams P,
                                                 float4 *__restrict__ vout) {
  const uint32_t s = particle_index(nblocks);
  if (s >= n) return;
  const float4 p = pos[s];
  const float4 vi = vin[s];
  const NbList nl = nb_list(list, head, s);
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  const Records16 Gp(pos, n), Gv(vin, n);
  auto load = [&](uint32_t j) { return Pair4{Gp[j], Gv[j]}; };
  for_neighbors<Pair4, NB, true>(nl, s, load, [&](const Pair4 &qv) {
    const float4 &q = qv.a, &vq = qv.b;
    const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    const float u = P.h2 - fmaf(dy * dy * dz, dz, dx * dx);
    const float up = fmaxf(u, 0.0f);
    const float w = up * up * up * fast_rcp(q.w);
    ax = fmaf(vq.x - vi.x, w, ax);
    ay = fmaf(vq.y - vi.y, w, ay);
    az = fmaf(vq.z - vi.z, w, az);
  });
  const float k = P.xsph_c * P.mWn;
  vout[s] = make_float4(fmaf(k, ax, vi.x), fmaf(k, ay, vi.y), fmaf(k, az, vi.z), 0.0f);
}

// ComputeVorticity (particles.c:207-283); aux = (omega, |omega|), and
// pw = (x*, |omega|) so that the confinement gather is one float4 per neighbour.
template <int NB>
__global__ __launch_bounds__(kBlock) void k_vorticity(uint32_t n, uint32_t nblocks,
                                                      const float4 *__restrict__ pos,
                                                      const float4 *__restrict__ vel,
                                                      const uint32_t *__restrict__ list,
                                                      const uint4 *__restrict__ head, DevParams P,
                                                      float4 *__restrict__ omega,
                                                      float4 *__restrict__ pw) {
  const uint32_t s = particle_index(nblocks);
  if (s >= n) return;
  const float4 p = pos[s];
  const float4 vi = vel[s];
  const NbList nl = nb_list(list, head, s);
  float ox = 0.0f, oy = 0.0f, oz = 0.0f;
  const Records16 Gp(pos, n), Gv(vel, n);
  auto load = [&](uint32_t j) { return Pair4{Gp[j], Gv[j]}; };
  for_neighbors<Pair4, NB, false>(nl, s, load, [&](const Pair4 &qv) {
    const float4 &q = qv.a, &vq = qv.b;
    const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    const float r2c = fmaxf(sq3(dx, dy, dz), P.rmin2);
    const float ir = fast_rsq(r2c);
    const float hr = fmaf(-r2c, ir, P.h);
    const float hp = fmaxf(hr, 0.0f);
    float g = hp * hp * ir;
    if (P.vort_model == 1) g *= fast_rcp(q.w);
    const float wx = g * dx, wy = g * dy, wz = g * dz;
    const float vx = vq.x - vi.x, vy = vq.y - vi.y, vz = vq.z - vi.z;
    ox = fmaf(vy, wz, fmaf(-vz, wy, ox));
    oy = fmaf(vz, wx, fmaf(-vx, wz, oy));
    oz = fmaf(vx, wy, fmaf(-vy, wx, oz));
  });
  const float k = P.vort_model == 1 ? P.DelW_norm * P.mass : P.DelW_norm;
  ox *= k;
  oy *= k;
  oz *= k;
  const float wn = sqrtf(sq3(ox, oy, oz));
  omega[s] = make_float4(ox, oy, oz, wn);
  pw[s] = make_float4(p.x, p.y, p.z, wn);
}

// ApplyVorticityConfinement (particles.c:285-381): eta = sum |omega_j| gradW,
// N = eta / max(|eta|, 1e-4 h), v += dt eps (N x omega_i) / m.  When COMMIT
// is set the result is also UpdatePositions (particles.c:566-586): x* and v
// are scattered back to particle-id order.
template <bool COMMIT, int NB>
__global__ GATHER_BOUNDS(NB) void k_confine(
    uint32_t n, uint32_t nblocks, const float4 *__restrict__ pw,
    const float4 *__restrict__ omega, const uint32_t *__restrict__ list,
    const uint4 *__restrict__ head, DevParams P, const float4 *__restrict__ vin,
    float4 *__restrict__ vout, const uint32_t *__restrict__ id_s, float4 *__restrict__ x_can,
    float4 *__restrict__ v_can) {
  const uint32_t s = particle_index(nblocks);
  if (s >= n) return;
  const float4 p = pw[s];
  const NbList nl = nb_list(list, head, s);
  float ex = 0.0f, ey = 0.0f, ez = 0.0f;
  // only x_j and |omega_j| of the neighbour are needed
  const Records16 G(pw, n);
  for_neighbors<float4, NB>(nl, s, [&](uint32_t j) { return G[j]; }, [&](const float4 &q) {
    const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    const float r2c = fmaxf(sq3(dx, dy, dz), P.rmin2);
    const float ir = fast_rsq(r2c);
    const float hr = fmaf(-r2c, ir, P.h);
    const float hp = fmaxf(hr, 0.0f);
    const float g = q.w * hp * hp * ir;
    ex = fmaf(g, dx, ex);
    ey = fmaf(g, dy, ey);
    ez = fmaf(g, dz, ez);
  });
  ex *= P.DelW_norm;
  ey *= P.DelW_norm;
  ez *= P.DelW_norm;
  const float em = fmaxf(sqrtf(sq3(ex, ey, ez)), P.rmin);
  const float Nx = ex / em, Ny = ey / em, Nz = ez / em;
  const float4 wi = omega[s];
  float4 v = vin[s];
  v.x += P.conf_k * (Ny * wi.z - Nz * wi.y);
  v.y += P.conf_k * (Nz * wi.x - Nx * wi.z);
  v.z += P.conf_k * (Nx * wi.y - Ny * wi.x);
  v.w = 0.0f;
  if (COMMIT) {
    const uint32_t id = id_s[s];
    x_can[id] = make_float4(p.x, p.y, p.z, 0.0f);
    v_can[id] = v;
  } else {
    vout[s] = v;
  }
}

// UpdatePositions (particles.c:566-586) with the scatter to id order.
__global__ __launch_bounds__(kBlock) void k_commit(uint32_t n, const float4 *__restrict__ pos,
                                                   const float4 *__restrict__ vel,
                                                   const uint32_t *__restrict__ id_s,
                                                   float4 *__restrict__ x_can,
                                                   float4 *__restrict__ v_can) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const uint32_t id = id_s[s];
  const float4 p = pos[s];
  const float4 v = vel[s];
  x_can[id] = make_float4(p.x, p.y, p.z, 0.0f);
  v_can[id] = make_float4(v.x, v.y, v.z, 0.0f);
}

// Substep bookkeeping: clears per-iteration error slots and flags.  A kernel
// rather than hipMemsetAsync so that the substep can be captured in a graph.
__global__ void k_substep_init(uint32_t *stats) {
  const int i = threadIdx.x;
  if (i >= STAT_ITERS && i < STAT_WORDS) stats[i] = 0u;
}

// Standalone CalculateParticleCollisions on SoA float arrays.
__global__ __launch_bounds__(kBlock) void k_collide_soa(uint32_t n, float *__restrict__ x,
                                                        float *__restrict__ y,
                                                        float *__restrict__ z, SdfDev o) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float a = x[i], b = y[i], c = z[i];
  collide(a, b, c, o);
  x[i] = a;
  y[i] = b;
  z[i] = c;
}

// CalculateParticleCollisions_kernel on the reference's own FP64 arrays
// (obstacle_cuda.cu:56-116): the bounds test and the normalised coordinates
// in double against the double world box, the SDF arithmetic in FP32, and
// the FP32 push added to the double coordinate.
__global__ __launch_bounds__(kBlock) void k_collide_soa_f64(uint32_t n, double *__restrict__ x,
                                                            double *__restrict__ y,
                                                            double *__restrict__ z, SdfDev o,
                                                            SdfBox64 b) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double xw = x[i], yw = y[i], zw = z[i];
  if (!(xw > b.lo[0] && xw < b.hi[0] && yw > b.lo[1] && yw < b.hi[1] && zw > b.lo[2] &&
        zw < b.hi[2]))
    return;
  const float u = (float)((xw - b.lo[0]) / (double)o.len[0]);
  const float v = (float)((yw - b.lo[1]) / (double)o.len[1]);
  const float w = (float)((zw - b.lo[2]) / (double)o.len[2]);
  float px, py, pz;
  if (!sdf_push(o, u, v, w, px, py, pz)) return;
  y[i] = yw + (double)py;
  x[i] = xw + (double)px;
  z[i] = zw + (double)pz;
}

// CalculateParticleCollisions over the sorted positions (after a k_dp_update
// without the fused push).  Only particles inside the obstacle's box are
// written.
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_collide_sorted(uint32_t n, float4 *__restrict__ pos,
                                                           SdfDev o0, BatchDev B) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= n) return;
  const SdfDev &o = sdf_of<BATCH>(s, o0, B);
  if (!o.data) return;
  const float4 p = pos[s];
  if (!(p.x > o.wmin[0] && p.x < o.wmax[0] && p.y > o.wmin[1] && p.y < o.wmax[1] &&
        p.z > o.wmin[2] && p.z < o.wmax[2]))
    return;
  float x = p.x, y = p.y, z = p.z;
  collide(x, y, z, o);
  if (x != p.x || y != p.y || z != p.z) pos[s] = make_float4(x, y, z, p.w);
}

// Max density error per scene of a batch, from the last density pass (rho in
// sorted order; a scene's particles are one index range there).  One block
// per chunk of at most kSceneChunk particles of one scene (the chunk table is
// built with the batch): the block reduces its chunk and issues one atomic,
// at most a few dozen per scene.  With one atomic per wave, as first written,
// the ~440 waves of a 28k-particle scene serialised on its word (0.96 ms for
// 512 scenes, 14.3M particles, against ~0.01 ms of traffic).
__global__ __launch_bounds__(kBlock) void k_scene_errors(const uint4 *__restrict__ chunks,
                                                         const float *__restrict__ rho,
                                                         DevParams P, uint32_t *err_bits) {
  __shared__ float wmax[kBlock / kWave];
  const uint4 ch = chunks[blockIdx.x];  // scene, first, end
  float m = 0.0f;
  for (uint32_t i = ch.y + threadIdx.x; i < ch.z; i += kBlock)
    m = fmaxf(m, (rho[i] - P.rest_density) * P.inv_rest_density);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = wmax[0];
    for (int w = 1; w < kBlock / kWave; ++w) b = fmaxf(b, wmax[w]);
    if (b > 0.0f) atomicMax(&err_bits[ch.x], __float_as_uint(b));
  }
}

__global__ __launch_bounds__(kBlock) void k_fill_words(uint32_t n, uint32_t *p, uint32_t v) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

// FP64 kinetic energy: one partial sum per block into out[blockIdx.x] (the
// caller adds them in block order, so the result does not depend on the order
// in which blocks finish, as it did with one atomic per block).
__global__ __launch_bounds__(kBlock) void k_kinetic_energy(uint32_t n,
                                                           const float4 *__restrict__ v,
                                                           double half_m, double *out) {
  __shared__ double part[kBlock / 64];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  double e = 0.0;
  if (i < n) {
    const float4 u = v[i];
    e = half_m * ((double)u.x * u.x + (double)u.y * u.y + (double)u.z * u.z);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) e += __shfl_xor(e, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) t += part[w];
    out[blockIdx.x] = t;  // summed in block order on the host
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
// Gathers in flight per wave of a gather kernel: the autotuner's choice or 16.
static inline int tuned_depth(const sph_ctx *c, int kernel) {
  const int d = c->tune.depth[kernel];
  return d ? d : 16;
}

template <typename F>
static inline void with_depth(int depth, F &&f) {
  if (depth == 32)
    f(std::integral_constant<int, 32>{});
  else if (depth == 8)
    f(std::integral_constant<int, 8>{});
  else
    f(std::integral_constant<int, 16>{});
}

// Threads per block of a gather kernel: the autotuner's choice or kBlock.
static inline uint32_t tuned_block(const sph_ctx *c, int kernel) {
  const uint32_t b = c->tune.block[kernel];
  return b ? b : (uint32_t)kBlock;
}

#define LAUNCH_TIMED(ctx, kid, grid, ...)                               \
  do {                                                                  \
    hipEvent_t ev_ = nullptr;                                           \
    prof_begin(ctx, kid, &ev_);                                         \
    __VA_ARGS__;                                                        \
    prof_end(ctx, kid, ev_);                                            \
  } while (0)

int launch_gravity(sph_ctx *c) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  LAUNCH_TIMED(c, K_GRAVITY, 0,
               k_gravity<<<blocks_for(n), kBlock, 0, c->stream>>>(n, c->v_can, c->dev));
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_predict(sph_ctx *c) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  const BatchDev B = c->batch();
  if (c->nscenes)
    LAUNCH_TIMED(c, K_PREDICT, 0,
                 k_predict<true><<<blocks_for(n), kBlock, 0, c->stream>>>(
                     n, c->x_can, c->v_can, c->xs_can, c->dev, B));
  else
    LAUNCH_TIMED(c, K_PREDICT, 0,
                 k_predict<false><<<blocks_for(n), kBlock, 0, c->stream>>>(
                     n, c->x_can, c->v_can, c->xs_can, c->dev, B));
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

// Hash (fused with gravity+predict when `fused`), sort, cell tables, reorder,
// lists: FindAllNeighbors (neighbors.c:68-75) in sorted space.
int launch_neighbor_search(sph_ctx *c, bool fused) {
  const uint32_t n = c->n;
  hipStream_t s = c->stream;
  const BatchDev B = c->batch();
  if (n) {
    if (fused && c->nscenes) {
      LAUNCH_TIMED(c, K_PREDICT_HASH, 0,
                   k_predict_hash<true><<<blocks_for(n), kBlock, 0, s>>>(
                       n, c->x_can, c->v_can, c->xs_can, c->keys, c->dev,
                       c->stats, B));
    } else if (fused) {
      LAUNCH_TIMED(c, K_PREDICT_HASH, 0,
                   k_predict_hash<false><<<blocks_for(n), kBlock, 0, s>>>(
                       n, c->x_can, c->v_can, c->xs_can, c->keys, c->dev,
                       c->stats, B));
    } else if (c->nscenes) {
      LAUNCH_TIMED(c, K_HASH, 0,
                   k_hash<true><<<blocks_for(n), kBlock, 0, s>>>(
                       n, c->xs_can, c->keys, c->dev, c->stats, B));
    } else {
      LAUNCH_TIMED(c, K_HASH, 0,
                   k_hash<false><<<blocks_for(n), kBlock, 0, s>>>(
                       n, c->xs_can, c->keys, c->dev, c->stats, B));
    }
    SPH_HIP(hipGetLastError());
  }
  // keys -> (key, id) sorted; ids are implicit 0..n-1 on the first pass
  const uint32_t *ko = nullptr, *vo = nullptr;
  int rc = radix_sort_pairs(c->keys, nullptr, c->keys_alt, c->ids_alt, c->keys, c->ids, n,
                            c->key_bits, &c->hist, &c->hist_cap, &c->scan, s, c, &ko, &vo);
  if (rc) return rc;
  if (n == 0) {
    ko = c->keys;
    vo = c->ids;
  }
  c->key_s = ko;
  c->id_s = vo;
  if (n) {
    c->pos = c->pos_a;
    LAUNCH_TIMED(c, K_REORDER, 0,
                 k_reorder<<<blocks_for(n), kBlock, 0, s>>>(n, c->id_s, c->key_s, c->xs_can,
                                                            c->x_can, c->pos, c->xold,
                                                            c->cell_end, c->sx,
                                                            c->soa_stride));
    SPH_HIP(hipGetLastError());
  }
  // cell_start[c] = #keys < c (lower bound) = max end of the cells before c,
  // cell_start[cells] = n; re-zeroes the ends
  rc = exclusive_scan(c->cell_end, c->cell_start, c->cells, true, true, &c->scan, s, c, true);
  if (rc) return rc;
  if (n) {
    const uint32_t bs = tuned_block(c, K_BUILD_LISTS);
    const uint32_t nb = blocks_for(n, bs);
    if (c->nscenes)
      LAUNCH_TIMED(c, K_BUILD_LISTS, 0,
                   k_build_lists<true><<<nb, bs, kLdsHits * bs * sizeof(uint16_t), s>>>(
                       n, nb, c->pos, c->sx, c->soa_stride, c->key_s, c->cell_start, c->dev,
                       c->list, c->list_head, c->stats, c->pair_partial,
                       c->tune.wide_lists ? 1 : 0, B));
    else
      LAUNCH_TIMED(c, K_BUILD_LISTS, 0,
                   k_build_lists<false><<<nb, bs, kLdsHits * bs * sizeof(uint16_t), s>>>(
                       n, nb, c->pos, c->sx, c->soa_stride, c->key_s, c->cell_start, c->dev,
                       c->list, c->list_head, c->stats, c->pair_partial,
                       c->tune.wide_lists ? 1 : 0, B));
    c->pair_blocks = nb;
    SPH_HIP(hipGetLastError());
  } else {
    c->pos = c->pos_a;
  }
  c->vel = c->vel_a;
  c->sorted_valid = true;
  c->lists_valid = true;
  return SPH_OK;
}

static float stiffness_scale(const sph_ctx *c, int substep) {
  // particles.c:693-694, 742-744: stiffness * m * DelW_norm / rho0
  const double st = 1.0 - pow(1.0 - c->P.stiffness_k, 1.0 / (double)substep);
  return (float)(st * c->P.rest_mass * (-45.0 / (M_PI * pow(c->P.smoothing_radius, 6.0))) /
                 c->P.rest_density);
}

int launch_density(sph_ctx *c, int mode, int err_slot) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  const uint32_t bs = tuned_block(c, K_DENSITY_LAMBDA);
  const uint32_t nb = blocks_for(n, bs);
  hipStream_t s = c->stream;
  if (mode == 0)
    with_depth(tuned_depth(c, K_DENSITY_LAMBDA), [&](auto D_) {
      constexpr int NBv = decltype(D_)::value;
      LAUNCH_TIMED(c, K_DENSITY_LAMBDA, 0,
                   k_density_lambda<0, NBv><<<nb, bs, 0, s>>>(
                       n, nb, c->pos, c->list, c->list_head, c->dev, c->rho, nullptr, c->stats,
                       err_slot, c->P.density_tolerance > 0, 0, nullptr));
    });
  else if (mode == 1)
    with_depth(tuned_depth(c, K_DENSITY), [&](auto D_) {
      constexpr int NBv = decltype(D_)::value;
      LAUNCH_TIMED(c, K_DENSITY, 0,
                   k_density_lambda<1, NBv><<<nb, bs, 0, s>>>(
                       n, nb, c->pos, c->list, c->list_head, c->dev, c->rho, c->cstage, c->stats,
                       err_slot, 0, 0, nullptr));
    });
  else if (mode == 4)  // density and lambda in one gather: rho into w, lambda aside
    with_depth(tuned_depth(c, K_DENSITY_LAMBDA), [&](auto D_) {
      constexpr int NBv = decltype(D_)::value;
      LAUNCH_TIMED(c, K_DENSITY_LAMBDA, 0,
                   k_density_lambda<0, NBv><<<nb, bs, 0, s>>>(
                       n, nb, c->pos, c->list, c->list_head, c->dev, c->rho, c->cstage, c->stats,
                       err_slot, 0, 1, c->lam_stage));
    });
  else if (mode == 3)  // final density of a capped tolerance loop: rho also into w
    with_depth(tuned_depth(c, K_DENSITY), [&](auto D_) {
      constexpr int NBv = decltype(D_)::value;
      LAUNCH_TIMED(c, K_DENSITY, 0,
                   k_density_lambda<1, NBv><<<nb, bs, 0, s>>>(
                       n, nb, c->pos, c->list, c->list_head, c->dev, c->rho, c->cstage, c->stats,
                       err_slot, 1, 1, nullptr));
    });
  else
    with_depth(tuned_depth(c, K_LAMBDA), [&](auto D_) {
      constexpr int NBv = decltype(D_)::value;
      LAUNCH_TIMED(c, K_LAMBDA, 0,
                   k_density_lambda<2, NBv><<<nb, bs, 0, s>>>(
                       n, nb, c->pos, c->list, c->list_head, c->dev, c->rho, c->cstage, c->stats,
                       -1, 0, 0, nullptr));
    });
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

// pos.w = the lambda a mode-4 density pass put aside.
__global__ __launch_bounds__(kBlock) void k_lambda_to_w(uint32_t n, const float *__restrict__ lam,
                                                        float4 *__restrict__ pos) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s < n) pos[s].w = lam[s];
}

int launch_lambda_from_density(sph_ctx *c) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  LAUNCH_TIMED(c, K_LAMBDA_COPY, 0,
               k_lambda_to_w<<<blocks_for(n), kBlock, 0, c->stream>>>(n, c->lam_stage, c->pos));
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_dp(sph_ctx *c, int substep, float4 *xnext) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  const uint32_t bs = tuned_block(c, K_DP);
  const uint32_t nb = blocks_for(n, bs);
  const float scale = stiffness_scale(c, substep);
  const BatchDev B = c->batch();
  with_depth(tuned_depth(c, K_DP), [&](auto D_) {
    constexpr int NBv = decltype(D_)::value;
    if (c->nscenes)
      LAUNCH_TIMED(c, K_DP, 0,
                   (k_dp<true, NBv><<<nb, bs, 0, c->stream>>>(n, nb, c->pos, c->list,
                                                             c->list_head, c->dev, scale,
                                                             c->dpbuf, xnext, c->rho, c->sdf, B)));
    else
      LAUNCH_TIMED(c, K_DP, 0,
                   (k_dp<false, NBv><<<nb, bs, 0, c->stream>>>(n, nb, c->pos, c->list,
                                                              c->list_head, c->dev, scale,
                                                              c->dpbuf, xnext, c->rho, c->sdf,
                                                              B)));
  });
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_position_stars(sph_ctx *c) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  if (c->nscenes)
    LAUNCH_TIMED(c, K_POS_STARS, 0,
                 k_position_stars<true><<<blocks_for(n), kBlock, 0, c->stream>>>(
                     n, c->pos, c->dpbuf, c->rho, c->dev, c->sdf, c->batch()));
  else
    LAUNCH_TIMED(c, K_POS_STARS, 0,
                 k_position_stars<false><<<blocks_for(n), kBlock, 0, c->stream>>>(
                     n, c->pos, c->dpbuf, c->rho, c->dev, c->sdf, c->batch()));
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

static float stiffness_unitless(const sph_ctx *c, int substep) {
  return (float)(1.0 - pow(1.0 - c->P.stiffness_k, 1.0 / (double)substep));
}

int launch_dp_update(sph_ctx *c, int substep, int err_slot, bool accumulate) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  const uint32_t bs = tuned_block(c, K_DP_UPDATE);
  const uint32_t nb = blocks_for(n, bs);
  float4 *out = (c->pos == c->pos_a) ? c->pos_b : c->pos_a;
  float *acc = accumulate ? c->lam_acc : nullptr;
  const float scale = stiffness_scale(c, substep), stiff = stiffness_unitless(c, substep);
  const float tol = (float)c->P.density_tolerance;
  const BatchDev B = c->batch();
  const bool obstacle = c->nscenes ? c->batch_sdf != nullptr : c->sdf.data != nullptr;
  const bool split = obstacle && c->tune.split_collide;
#define DP_UPDATE(COL, BAT)                                                                 \
  with_depth(tuned_depth(c, K_DP_UPDATE), [&](auto D_) {                                    \
    constexpr int NBv = decltype(D_)::value;                                                \
    LAUNCH_TIMED(c, K_DP_UPDATE, 0,                                                         \
                 (k_dp_update<COL, BAT, NBv><<<nb, bs, 0, c->stream>>>(                      \
                     n, nb, c->pos, out, c->rho, c->list, c->list_head, c->dev, scale, c->sdf, \
                     c->stats, err_slot, tol, acc, stiff, B)));                             \
  })
  if (c->nscenes) {
    if (split) DP_UPDATE(false, true); else DP_UPDATE(true, true);
  } else {
    if (split) DP_UPDATE(false, false); else DP_UPDATE(true, false);
  }
#undef DP_UPDATE
  SPH_HIP(hipGetLastError());
  if (split) {
    if (c->nscenes)
      LAUNCH_TIMED(c, K_COLLIDE, 0,
                   k_collide_sorted<true><<<blocks_for(n), kBlock, 0, c->stream>>>(
                       n, out, c->sdf, B));
    else
      LAUNCH_TIMED(c, K_COLLIDE, 0,
                   k_collide_sorted<false><<<blocks_for(n), kBlock, 0, c->stream>>>(
                       n, out, c->sdf, B));
    SPH_HIP(hipGetLastError());
  }
  c->pos = out;
  return SPH_OK;
}

int launch_warm_lambda(sph_ctx *c, bool log) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  const uint32_t bs = tuned_block(c, K_WARM_LAMBDA);
  const uint32_t nb = blocks_for(n, bs);
  const BatchDev B = c->batch();
  float2 *plain = reinterpret_cast<float2 *>(c->dpbuf);
#define WARM(LOG, BAT)                                                                      \
  LAUNCH_TIMED(c, K_WARM_LAMBDA, 0,                                                         \
               (k_warm_lambda<LOG, BAT><<<nb, bs, 0, c->stream>>>(                       \
                   n, nb, c->pos, c->xold, c->list, c->list_head, c->dev, c->rho, c->stats,   \
                   c->id_s, c->lam_prev, c->sdf, LOG ? nullptr : c->warm_net, c->lam_acc,   \
                   LOG ? c->feat : nullptr, LOG ? nullptr : plain, B)))
  if (c->nscenes) {
    if (log) WARM(true, true); else WARM(false, true);
  } else {
    if (log) WARM(true, false); else WARM(false, false);
  }
#undef WARM
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_warm_prev(sph_ctx *c) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  int rc = launch_density(c, 0, 0);
  if (rc) return rc;
  LAUNCH_TIMED(c, K_AUX, 0,
               k_warm_prev<<<blocks_for(n), kBlock, 0, c->stream>>>(
                   n, c->pos, c->rho, c->id_s, c->lam_prev, c->warm_prev_scale,
                   reinterpret_cast<float2 *>(c->dpbuf), c->lam_acc));
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_warm_fallback(sph_ctx *c, const float4 *pre_warm) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  if (c->nscenes) {
    // per scene; scene_err0 / scene_err hold the errors before and after the step
    LAUNCH_TIMED(c, K_AUX, 0,
                 k_warm_fallback_scenes<<<blocks_for(n), kBlock, 0, c->stream>>>(
                     n, c->stats, c->scene_of, c->scene_err0, c->scene_err, pre_warm, c->pos,
                     reinterpret_cast<const float2 *>(c->dpbuf), c->rho, c->lam_acc));
    SPH_HIP(hipGetLastError());
    k_warm_fallback_stats<<<1, kBlock, 0, c->stream>>>((uint32_t)c->nscenes, c->stats,
                                                       c->scene_err0, c->scene_err);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
  }
  LAUNCH_TIMED(c, K_AUX, 0,
               k_warm_fallback<<<blocks_for(n), kBlock, 0, c->stream>>>(
                   n, c->stats, pre_warm, c->pos, reinterpret_cast<const float2 *>(c->dpbuf),
                   c->rho, c->lam_acc));
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_scene_errors(sph_ctx *c, uint32_t *dst) {
  const uint32_t ns = (uint32_t)c->nscenes;
  if (!ns) return SPH_OK;
  if (!dst) dst = c->scene_err;
  k_fill_words<<<blocks_for(ns), kBlock, 0, c->stream>>>(ns, dst, 0u);
  SPH_HIP(hipGetLastError());
  if (c->n && c->nchunks) {
    k_scene_errors<<<c->nchunks, kBlock, 0, c->stream>>>(c->scene_chunks, c->rho, c->dev, dst);
    SPH_HIP(hipGetLastError());
  }
  return SPH_OK;
}

int launch_scatter_lambda(sph_ctx *c) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  LAUNCH_TIMED(c, K_AUX, 0,
               k_scatter_lambda<<<blocks_for(n), kBlock, 0, c->stream>>>(n, c->id_s,
                                                                         c->lam_acc,
                                                                         c->lam_prev));
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_warm_samples(sph_ctx *c, float *x, float *t, uint32_t count, int every) {
  if (!count) return SPH_OK;
  k_warm_samples<<<blocks_for(count), kBlock, 0, c->stream>>>(count, (uint32_t)every, c->feat,
                                                              c->lam_acc, x, t);
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}


int launch_velocity(sph_ctx *c) {
  const uint32_t n = c->n;
  c->vel = c->vel_a;
  if (!n) return SPH_OK;
  LAUNCH_TIMED(c, K_VELOCITY, 0,
               k_velocity<<<blocks_for(n), kBlock, 0, c->stream>>>(n, c->pos, c->xold, c->vel,
                                                                   c->dev));
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_color(sph_ctx *c, bool with_velocity) {
  const uint32_t n = c->n;
  if (with_velocity) c->vel = c->vel_a;
  if (!n) return SPH_OK;
  const uint32_t bs = tuned_block(c, K_COLOR);
  const uint32_t nb = blocks_for(n, bs);
  if (with_velocity)
    with_depth(tuned_depth(c, K_COLOR), [&](auto D_) {
      constexpr int NBv = decltype(D_)::value;
      LAUNCH_TIMED(c, K_COLOR, 0,
                   k_color<true, NBv><<<nb, bs, 0, c->stream>>>(n, nb, c->pos, c->xold, c->list,
                                                               c->list_head, c->dev, c->vel,
                                                               c->aux));
    });
  else
    with_depth(tuned_depth(c, K_COLOR), [&](auto D_) {
      constexpr int NBv = decltype(D_)::value;
      LAUNCH_TIMED(c, K_COLOR, 0,
                   k_color<false, NBv><<<nb, bs, 0, c->stream>>>(n, nb, c->pos, c->xold, c->list,
                                                                c->list_head, c->dev, c->vel,
                                                                c->aux));
    });
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_surface(sph_ctx *c) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  const uint32_t bs = tuned_block(c, K_SURFACE);
  const uint32_t nb = blocks_for(n, bs);
  with_depth(tuned_depth(c, K_SURFACE), [&](auto D_) {
    constexpr int NBv = decltype(D_)::value;
    LAUNCH_TIMED(c, K_SURFACE, 0,
                 k_surface<NBv><<<nb, bs, 0, c->stream>>>(n, nb, c->pos, c->aux, c->list,
                                                         c->list_head, c->dev, c->vel));
  });
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_xsph(sph_ctx *c) {
  const uint32_t n = c->n;
  float4 *out = (c->vel == c->vel_a) ? c->vel_b : c->vel_a;
  if (n) {
    const uint32_t bs = tuned_block(c, K_XSPH);
    const uint32_t nb = blocks_for(n, bs);
    with_depth(tuned_depth(c, K_XSPH), [&](auto D_) {
      constexpr int NBv = decltype(D_)::value;
      LAUNCH_TIMED(c, K_XSPH, 0,
                   k_xsph<NBv><<<nb, bs, 0, c->stream>>>(n, nb, c->pos, c->vel, c->list,
                                                        c->list_head, c->dev, out));
    });
    SPH_HIP(hipGetLastError());
  }
  c->vel = out;
  return SPH_OK;
}

int launch_vorticity(sph_ctx *c) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  const uint32_t bs = tuned_block(c, K_VORTICITY);
  const uint32_t nb = blocks_for(n, bs);
  with_depth(tuned_depth(c, K_VORTICITY), [&](auto D_) {
    constexpr int NBv = decltype(D_)::value;
    LAUNCH_TIMED(c, K_VORTICITY, 0,
                 k_vorticity<NBv><<<nb, bs, 0, c->stream>>>(n, nb, c->pos, c->vel, c->list,
                                                           c->list_head, c->dev, c->aux,
                                                           c->pw));
  });
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_confine(sph_ctx *c, bool commit) {
  const uint32_t n = c->n;
  float4 *out = (c->vel == c->vel_a) ? c->vel_b : c->vel_a;
  if (n) {
    const uint32_t bs = tuned_block(c, K_CONFINE);
    const uint32_t nb = blocks_for(n, bs);
    if (commit)
      with_depth(tuned_depth(c, K_CONFINE), [&](auto D_) {
        constexpr int NBv = decltype(D_)::value;
        LAUNCH_TIMED(c, K_CONFINE, 0,
                       k_confine<true, NBv><<<nb, bs, 0, c->stream>>>(
                           n, nb, c->pw, c->aux, c->list, c->list_head, c->dev, c->vel, out,
                           c->id_s, c->x_can, c->v_can));
      });
    else
      with_depth(tuned_depth(c, K_CONFINE), [&](auto D_) {
        constexpr int NBv = decltype(D_)::value;
        LAUNCH_TIMED(c, K_CONFINE, 0,
                       k_confine<false, NBv><<<nb, bs, 0, c->stream>>>(
                           n, nb, c->pw, c->aux, c->list, c->list_head, c->dev, c->vel, out,
                           c->id_s, c->x_can, c->v_can));
      });
    SPH_HIP(hipGetLastError());
  }
  if (!commit) c->vel = out;
  return SPH_OK;
}

int launch_commit(sph_ctx *c) {
  const uint32_t n = c->n;
  if (!n) return SPH_OK;
  LAUNCH_TIMED(c, K_COMMIT, 0,
               k_commit<<<blocks_for(n), kBlock, 0, c->stream>>>(n, c->pos, c->vel, c->id_s,
                                                                 c->x_can, c->v_can));
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_substep_init(sph_ctx *c) {
  k_substep_init<<<1, STAT_WORDS, 0, c->stream>>>(c->stats);
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_collide_soa(float *x, float *y, float *z, uint32_t n, const SdfDev &o,
                       hipStream_t s) {
  if (!n) return SPH_OK;
  k_collide_soa<<<blocks_for(n), kBlock, 0, s>>>(n, x, y, z, o);
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_collide_soa_f64(double *x, double *y, double *z, uint32_t n, const SdfDev &o,
                           const SdfBox64 &b, hipStream_t s) {
  if (!n) return SPH_OK;
  k_collide_soa_f64<<<blocks_for(n), kBlock, 0, s>>>(n, x, y, z, o, b);
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

int launch_kinetic_energy(sph_ctx *c, double *out_dev) {
  if (c->n)
    k_kinetic_energy<<<blocks_for(c->n), kBlock, 0, c->stream>>>(
        c->n, c->v_can, 0.5 * c->P.rest_mass, out_dev);
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

}  // namespace sph
