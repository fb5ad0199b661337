// Stable key/value radix sort, exclusive scan and cell bounds for gfx950.
//
// Replaces the Thrust wrappers of the reference (thrust_c.cu:86-120):
//  - SortByKey: LSD radix sort, 8-bit digits.  Each pass is an upsweep
//    (per-tile digit histogram), a scan of the histogram in digit-major
//    order, and a downsweep that ranks keys inside the tile with wavefront
//    ballots (peers with the same digit), stages the tile in LDS in digit
//    order and writes each digit run out contiguously.  Ranks follow input
//    position (wave, round, lane), so the sort is stable.
//  - FindLowerBounds/FindUpperBounds: a key histogram followed by an
//    exclusive scan gives lower_bound(c) for every cell; upper_bound(c) is
//    lower_bound(c+1).
#include <mutex>

#include "sph_internal.h"

namespace sph {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;  // 4096 keys per block
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 16;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;

// Scan operators (identity 0 for both: the values are unsigned).
struct OpAdd {
  static __device__ __forceinline__ uint32_t f(uint32_t a, uint32_t b) { return a + b; }
};
struct OpMax {
  static __device__ __forceinline__ uint32_t f(uint32_t a, uint32_t b) { return max(a, b); }
};

template <class Op = OpAdd>
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x = Op::f(x, y);
  }
  return x;
}

// Exclusive scan over a 256-thread block; *tot receives the block total.
template <class Op = OpAdd>
__device__ __forceinline__ uint32_t block_excl_scan_256(uint32_t x, uint32_t *wtot,
                                                        uint32_t *tot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan<Op>(x);
  if (lane == 63) wtot[w] = inc;
  __syncthreads();
  uint32_t pre = 0, all = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint32_t t = wtot[i];
    if (i < w) pre = Op::f(pre, t);
    all = Op::f(all, t);
  }
  __syncthreads();
  *tot = all;
  const uint32_t before = __shfl_up(inc, 1, 64);
  return Op::f(pre, lane ? before : 0u);
}

// Lanes of the wave holding the same 8-bit digit (valid lanes only).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bb = __ballot(bit);
    peers &= bit ? bb : ~bb;
  }
  return peers;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// --------------------------------------------------------------------------
// Scan
// --------------------------------------------------------------------------
// Thread t of a tile owns entries [t*SC_ITEMS, (t+1)*SC_ITEMS): four 16-byte
// loads and stores per thread, one block scan per tile.
__device__ __forceinline__ void load_items(const uint32_t *in, uint32_t i0, uint32_t len,
                                           uint32_t (&x)[SC_ITEMS]) {
  if (i0 + SC_ITEMS <= len) {
    const uint4 *v = reinterpret_cast<const uint4 *>(in + i0);
#pragma unroll
    for (int q = 0; q < SC_ITEMS / 4; ++q) {
      const uint4 a = v[q];
      x[4 * q] = a.x;
      x[4 * q + 1] = a.y;
      x[4 * q + 2] = a.z;
      x[4 * q + 3] = a.w;
    }
  } else {
#pragma unroll
    for (int r = 0; r < SC_ITEMS; ++r) x[r] = i0 + r < len ? in[i0 + r] : 0u;
  }
}

template <class Op>
__global__ __launch_bounds__(SC_THREADS) void k_scan_tile_sums(const uint32_t *__restrict__ in,
                                                               uint32_t len,
                                                               uint32_t *__restrict__ sums) {
  __shared__ uint32_t wtot[4];
  uint32_t x[SC_ITEMS];
  load_items(in, blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS, len, x);
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < SC_ITEMS; ++r) s = Op::f(s, x[r]);
  uint32_t tot;
  block_excl_scan_256<Op>(s, wtot, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <class Op>
__global__ __launch_bounds__(SC_THREADS) void k_scan_spine(uint32_t *sums, uint32_t tiles) {
  __shared__ uint32_t wtot[4];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < tiles; base += SC_THREADS) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t x = i < tiles ? sums[i] : 0u;
    uint32_t tot;
    const uint32_t e = block_excl_scan_256<Op>(x, wtot, &tot);
    if (i < tiles) sums[i] = Op::f(carry, e);
    carry = Op::f(carry, tot);
  }
}

// in and out may alias (in-place scan).
template <class Op>
__global__ __launch_bounds__(SC_THREADS) void k_scan_apply(uint32_t *in, uint32_t *out,
                                                           uint32_t len,
                                                           const uint32_t *__restrict__ sums,
                                                           int zero_in, int write_total) {
  __shared__ uint32_t wtot[4];
  const uint32_t i0 = blockIdx.x * SC_TILE + threadIdx.x * SC_ITEMS;
  uint32_t x[SC_ITEMS];
  load_items(in, i0, len, x);
  uint32_t own = 0;
#pragma unroll
  for (int r = 0; r < SC_ITEMS; ++r) own = Op::f(own, x[r]);
  uint32_t tot;
  uint32_t run = Op::f(sums[blockIdx.x], block_excl_scan_256<Op>(own, wtot, &tot));
#pragma unroll
  for (int r = 0; r < SC_ITEMS; ++r) {
    const uint32_t e = run;
    run = Op::f(run, x[r]);
    x[r] = e;
  }
  if (i0 + SC_ITEMS <= len) {
    uint4 *o = reinterpret_cast<uint4 *>(out + i0);
#pragma unroll
    for (int q = 0; q < SC_ITEMS / 4; ++q)
      o[q] = make_uint4(x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3]);
    if (zero_in) {
      uint4 *z = reinterpret_cast<uint4 *>(in + i0);
#pragma unroll
      for (int q = 0; q < SC_ITEMS / 4; ++q) z[q] = make_uint4(0u, 0u, 0u, 0u);
    }
  } else {
#pragma unroll
    for (int r = 0; r < SC_ITEMS; ++r)
      if (i0 + r < len) {
        out[i0 + r] = x[r];
        if (zero_in) in[i0 + r] = 0u;
      }
  }
  // the thread holding entry len-1 writes the total after it
  if (write_total && i0 < len && len <= i0 + SC_ITEMS) out[len] = run;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint32_t len) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) p[i] = v;
}

int exclusive_scan(uint32_t *in, uint32_t *out, uint32_t len, bool zero_in, bool write_total,
                   ScanScratch *sc, hipStream_t s, sph_ctx *prof, bool max_op) {
  if (len == 0) {
    if (write_total) {
      k_fill_u32<<<1, 64, 0, s>>>(out, 0u, 1u);
      SPH_HIP(hipGetLastError());
    }
    return SPH_OK;
  }
  if (((uintptr_t)in | (uintptr_t)out) & 15u) {
    set_error("exclusive_scan: buffers must be 16-byte aligned");
    return SPH_ERR_ARG;
  }
  const uint32_t tiles = (len + SC_TILE - 1) / SC_TILE;
  if (sc->tile_cap < tiles) {
    if (sc->tile_sums) SPH_HIP(hipFree(sc->tile_sums));
    sc->tile_sums = nullptr;
    SPH_HIP(hipMalloc(&sc->tile_sums, sizeof(uint32_t) * tiles));
    sc->tile_cap = tiles;
  }
  hipEvent_t ev = nullptr;
  if (prof) prof_begin(prof, K_SCAN, &ev);
  if (max_op) {
    k_scan_tile_sums<OpMax><<<tiles, SC_THREADS, 0, s>>>(in, len, sc->tile_sums);
    k_scan_spine<OpMax><<<1, SC_THREADS, 0, s>>>(sc->tile_sums, tiles);
    k_scan_apply<OpMax><<<tiles, SC_THREADS, 0, s>>>(in, out, len, sc->tile_sums,
                                                     zero_in ? 1 : 0, write_total ? 1 : 0);
  } else {
    k_scan_tile_sums<OpAdd><<<tiles, SC_THREADS, 0, s>>>(in, len, sc->tile_sums);
    k_scan_spine<OpAdd><<<1, SC_THREADS, 0, s>>>(sc->tile_sums, tiles);
    k_scan_apply<OpAdd><<<tiles, SC_THREADS, 0, s>>>(in, out, len, sc->tile_sums,
                                                     zero_in ? 1 : 0, write_total ? 1 : 0);
  }
  if (prof) prof_end(prof, K_SCAN, ev);
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

// --------------------------------------------------------------------------
// Radix sort
// --------------------------------------------------------------------------
__global__ __launch_bounds__(RS_THREADS) void k_radix_upsweep(const uint32_t *__restrict__ keys,
                                                              uint32_t n, int shift,
                                                              uint32_t *__restrict__ hist,
                                                              uint32_t tiles) {
  __shared__ uint32_t h[RS_WAVES][256];
  for (int i = threadIdx.x; i < RS_WAVES * 256; i += RS_THREADS) (&h[0][0])[i] = 0u;
  __syncthreads();
  const int w = threadIdx.x >> 6;
  const uint64_t lt = lanemask_lt();
  const uint32_t base = blockIdx.x * RS_TILE;
  uint32_t kk[RS_ITEMS];
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const uint32_t i = base + r * RS_THREADS + threadIdx.x;
    kk[r] = i < n ? keys[i] : 0u;
  }
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const uint32_t i = base + r * RS_THREADS + threadIdx.x;
    const bool valid = i < n;
    const uint32_t d = (kk[r] >> shift) & 255u;
    const uint64_t peers = match_digit(d, valid);
    // one leader per digit per wave adds the digit's count (a non-returning
    // LDS add: no wait between rounds)
    if (valid && (peers & lt) == 0) atomicAdd(&h[w][d], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  uint32_t s = 0;
#pragma unroll
  for (int ww = 0; ww < RS_WAVES; ++ww) s += h[ww][threadIdx.x];
  hist[(size_t)threadIdx.x * tiles + blockIdx.x] = s;
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_downsweep(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
    uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, uint32_t n, int shift,
    const uint32_t *__restrict__ offs, uint32_t tiles) {
  __shared__ uint32_t wcnt[RS_WAVES][256];
  __shared__ uint32_t dstart[256];
  __shared__ uint32_t goff[256];
  __shared__ uint32_t wtot[4];
  __shared__ uint32_t sk[RS_TILE];
  __shared__ uint32_t sv[RS_TILE];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  for (int i = tid; i < RS_WAVES * 256; i += RS_THREADS) (&wcnt[0][0])[i] = 0u;
  __syncthreads();
  const uint64_t lt = lanemask_lt();
  const uint32_t tile0 = blockIdx.x * RS_TILE;
  const uint32_t wbase = tile0 + (uint32_t)w * 64u * RS_ITEMS;
  uint32_t k[RS_ITEMS], v[RS_ITEMS], rk[RS_ITEMS];
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const uint32_t i = wbase + (uint32_t)r * 64u + (uint32_t)lane;
    const bool valid = i < n;
    k[r] = valid ? kin[i] : 0u;
    v[r] = valid ? (vin ? vin[i] : i) : 0u;
  }
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const uint32_t i = wbase + (uint32_t)r * 64u + (uint32_t)lane;
    const bool valid = i < n;
    const uint32_t d = (k[r] >> shift) & 255u;
    const uint64_t peers = match_digit(d, valid);
    const uint32_t below = (uint32_t)__popcll(peers & lt);
    // the digit's leader (lowest peer) takes the wave's running count and adds
    // the peers; LDS operations of a wave complete in order, so later rounds
    // see this one's add without waiting for its result
    uint32_t old = 0u;
    if (valid && below == 0) old = atomicAdd(&wcnt[w][d], (uint32_t)__popcll(peers));
    const int leader = peers ? (int)__builtin_ctzll(peers) : lane;
    rk[r] = __shfl(old, leader, 64) + below;
  }
  __syncthreads();
  // Digit `tid`: prefix over waves, then tile-level digit starts.
  uint32_t run = 0;
#pragma unroll
  for (int ww = 0; ww < RS_WAVES; ++ww) {
    const uint32_t c = wcnt[ww][tid];
    wcnt[ww][tid] = run;
    run += c;
  }
  uint32_t tot;
  const uint32_t ex = block_excl_scan_256(run, wtot, &tot);
  dstart[tid] = ex;
  goff[tid] = offs[(size_t)tid * tiles + blockIdx.x];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ITEMS; ++r) {
    const uint32_t i = wbase + (uint32_t)r * 64u + (uint32_t)lane;
    if (i < n) {
      const uint32_t d = (k[r] >> shift) & 255u;
      const uint32_t lp = dstart[d] + wcnt[w][d] + rk[r];
      sk[lp] = k[r];
      sv[lp] = v[r];
    }
  }
  __syncthreads();
  const uint32_t tile_n = min((uint32_t)RS_TILE, n - tile0);
  for (uint32_t i = tid; i < tile_n; i += RS_THREADS) {
    const uint32_t kk = sk[i];
    const uint32_t d = (kk >> shift) & 255u;
    const uint32_t o = goff[d] + (i - dstart[d]);
    kout[o] = kk;
    vout[o] = sv[i];
  }
}

int radix_sort_pairs(const uint32_t *kin, const uint32_t *vin, uint32_t *k0, uint32_t *v0,
                     uint32_t *k1, uint32_t *v1, uint32_t n, int key_bits, uint32_t **hist,
                     size_t *hist_cap, ScanScratch *sc, hipStream_t s, sph_ctx *prof,
                     const uint32_t **kout, const uint32_t **vout) {
  *kout = kin;
  *vout = vin;
  if (n == 0) return SPH_OK;
  const int passes = key_bits <= 0 ? 1 : (key_bits + 7) / 8;
  const uint32_t tiles = (n + RS_TILE - 1) / RS_TILE;
  const size_t need = (size_t)256 * tiles;
  if (*hist_cap < need) {
    if (*hist) SPH_HIP(hipFree(*hist));
    *hist = nullptr;
    SPH_HIP(hipMalloc(hist, sizeof(uint32_t) * need));
    *hist_cap = need;
  }
  const uint32_t *ck = kin, *cv = vin;
  for (int p = 0; p < passes; ++p) {
    uint32_t *ok = (p & 1) ? k1 : k0;
    uint32_t *ov = (p & 1) ? v1 : v0;
    const int shift = 8 * p;
    hipEvent_t ev = nullptr;
    if (prof) prof_begin(prof, K_RADIX_UP, &ev);
    k_radix_upsweep<<<tiles, RS_THREADS, 0, s>>>(ck, n, shift, *hist, tiles);
    if (prof) prof_end(prof, K_RADIX_UP, ev);
    SPH_HIP(hipGetLastError());
    int rc = exclusive_scan(*hist, *hist, (uint32_t)need, false, false, sc, s, prof);
    if (rc) return rc;
    if (prof) prof_begin(prof, K_RADIX_DOWN, &ev);
    k_radix_downsweep<<<tiles, RS_THREADS, 0, s>>>(ck, cv, ok, ov, n, shift, *hist, tiles);
    if (prof) prof_end(prof, K_RADIX_DOWN, ev);
    SPH_HIP(hipGetLastError());
    ck = ok;
    cv = ov;
  }
  *kout = ck;
  *vout = cv;
  return SPH_OK;
}

__global__ void k_count_keys(const uint32_t *__restrict__ keys, uint32_t n,
                             uint32_t *__restrict__ count, uint32_t cells) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const uint32_t k = keys[i];
    if (k < cells) atomicAdd(&count[k], 1u);
  }
}

int count_keys(const uint32_t *keys, uint32_t n, uint32_t *count, uint32_t cells,
               hipStream_t s) {
  const uint32_t len = cells + 1;
  k_fill_u32<<<blocks_for(len), kBlock, 0, s>>>(count, 0u, len);
  if (n) k_count_keys<<<blocks_for(n), kBlock, 0, s>>>(keys, n, count, cells);
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

}  // namespace sph

using namespace sph;

// --------------------------------------------------------------------------
// C ABI: native sort / bounds (replacing thrust_c.cu:86-120)
// --------------------------------------------------------------------------
// Scratch of the native sort / bounds entry points, kept per device between
// calls (grow-only, held until the process ends) so that a caller that sorts
// every substep does not allocate each time.  Calls on one device take turns;
// each ends with a synchronize of the caller's stream.
namespace {
struct AbiScratch {
  std::mutex m;
  uint32_t *a = nullptr, *b = nullptr;  // `words` each
  size_t words = 0;
  uint32_t *hist = nullptr;
  size_t hist_cap = 0;
  ScanScratch sc;
  uint32_t *word = nullptr;  // one word for the key OR-reduction
};
constexpr int kAbiDevices = 64;
AbiScratch g_abi[kAbiDevices];

int abi_scratch(AbiScratch **out) {
  int dev = 0;
  SPH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kAbiDevices) {
    set_error("device index beyond the native entry points' scratch table");
    return SPH_ERR_ARG;
  }
  *out = &g_abi[dev];
  return SPH_OK;
}

// a and b hold at least `words` words each (contents not kept on growth)
int abi_words(AbiScratch *w, size_t words) {
  if (w->words >= words) return SPH_OK;
  if (w->a) SPH_HIP(hipFree(w->a));
  if (w->b) SPH_HIP(hipFree(w->b));
  w->a = w->b = nullptr;
  w->words = 0;
  hipError_t e = hipMalloc(&w->a, sizeof(uint32_t) * words);
  if (e == hipSuccess) e = hipMalloc(&w->b, sizeof(uint32_t) * words);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc (native entry point scratch)");
  w->words = words;
  return SPH_OK;
}
}  // namespace

// OR of all keys (one device-scope atomic per wave): the highest set bit gives
// the radix passes a 32-bit sort needs.
__global__ __launch_bounds__(kBlock) void k_or_keys(const uint32_t *__restrict__ keys,
                                                    uint32_t n, uint32_t *__restrict__ out) {
  uint32_t v = 0;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    v |= keys[i];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v |= __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicOr(out, v);
}

static int sort_impl(uint32_t *keys, uint32_t *values, int n, int key_bits, void *stream) {
  if (n < 0 || (n > 0 && (!keys || !values)) || key_bits < 0 || key_bits > 32) {
    set_error("sph_sort_by_key: invalid argument");
    return SPH_ERR_ARG;
  }
  if (n <= 1) return SPH_OK;
  hipStream_t s = (hipStream_t)stream;
  AbiScratch *w = nullptr;
  int rc = abi_scratch(&w);
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(w->m);
  if ((rc = abi_words(w, (size_t)n))) return rc;
  if (key_bits == 0) {  // unknown key range: from the keys' OR
    if (!w->word) {
      hipError_t e = hipMalloc(&w->word, sizeof(uint32_t));
      if (e != hipSuccess) {
        w->word = nullptr;
        return hip_fail(e, "hipMalloc (key range)");
      }
    }
    k_fill_u32<<<1, 64, 0, s>>>(w->word, 0u, 1u);
    k_or_keys<<<std::min(blocks_for((uint32_t)n), 1024u), kBlock, 0, s>>>(keys, (uint32_t)n,
                                                                          w->word);
    SPH_HIP(hipGetLastError());
    uint32_t all = 0;
    SPH_HIP(hipMemcpyAsync(&all, w->word, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    SPH_HIP(hipStreamSynchronize(s));
    key_bits = all ? 32 - __builtin_clz(all) : 1;
  }
  const uint32_t *kout = nullptr, *vout = nullptr;
  rc = radix_sort_pairs(keys, values, w->a, w->b, keys, values, (uint32_t)n, key_bits, &w->hist,
                        &w->hist_cap, &w->sc, s, nullptr, &kout, &vout);
  if (rc == SPH_OK && kout != keys) {
    hipError_t e =
        hipMemcpyAsync(keys, kout, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(values, vout, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) rc = hip_fail(e, "hipMemcpyAsync (sort result)");
  }
  const hipError_t e = hipStreamSynchronize(s);
  if (rc == SPH_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize (sort)");
  return rc;
}

extern "C" int sph_sort_by_key_bits(uint32_t *keys, uint32_t *values, int n, int key_bits,
                                    void *stream) {
  if (key_bits < 1) {
    set_error("sph_sort_by_key: invalid argument");
    return SPH_ERR_ARG;
  }
  return sort_impl(keys, values, n, key_bits, stream);
}

// Keys of any range: the passes follow the keys' highest set bit.
extern "C" int sph_sort_by_key(uint32_t *keys, uint32_t *values, int n, void *stream) {
  return sort_impl(keys, values, n, 0, stream);
}

// --------------------------------------------------------------------------
// Sort without allocation: scratch in a file-scope device array
// --------------------------------------------------------------------------
// For callers that cannot allocate device memory (code under stream capture,
// serving loops, harnesses that forbid hipMalloc).  Layout, in words: the two
// ping-pong buffers (kStaticPairs each), the digit histogram (256 per 4096-key
// tile) and the scan's tile sums; each part starts on a 16-byte boundary.
namespace {
constexpr size_t kStaticPairs = size_t(1) << 21;
constexpr size_t kStaticTiles = (kStaticPairs + RS_TILE - 1) / RS_TILE;
constexpr size_t kStaticHist = 256 * kStaticTiles;
constexpr size_t kStaticScanTiles = (kStaticHist + SC_TILE - 1) / SC_TILE;
constexpr size_t kStaticWords = 2 * kStaticPairs + kStaticHist + ((kStaticScanTiles + 3) & ~3ull);
static_assert(kStaticPairs % 4 == 0 && kStaticHist % 4 == 0, "16-byte aligned parts");
static_assert(kStaticPairs == SPH_SORT_NOALLOC_MAX_PAIRS, "header limit");
}  // namespace

// Each GPU holds its own copy until the process ends (about 17 MB).
static __device__ __align__(256) uint32_t scratch[kStaticWords];

namespace {
// Device address of `scratch` on each device, looked up once per device with
// that device current (the host-side name is not a device address).
std::mutex g_static_m;
uint32_t *g_static_ptr[kAbiDevices] = {};

int static_scratch(uint32_t **out) {
  int dev = 0;
  SPH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kAbiDevices) {
    set_error("device index beyond the native entry points' scratch table");
    return SPH_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_static_m);
  if (!g_static_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress (sort scratch)");
    g_static_ptr[dev] = static_cast<uint32_t *>(p);
  }
  *out = g_static_ptr[dev];
  return SPH_OK;
}
}  // namespace

extern "C" int sph_sort_by_key_bits_noalloc(uint32_t *keys, uint32_t *values, int n,
                                            int key_bits, void *stream) {
  if (n < 0 || (n > 0 && (!keys || !values)) || key_bits < 1 || key_bits > 32) {
    set_error("sph_sort_by_key_bits_noalloc: invalid argument");
    return SPH_ERR_ARG;
  }
  if ((size_t)n > kStaticPairs) {
    set_error("sph_sort_by_key_bits_noalloc: n exceeds SPH_SORT_NOALLOC_MAX_PAIRS (2^21)");
    return SPH_ERR_ARG;
  }
  if (n <= 1) return SPH_OK;
  uint32_t *base = nullptr;
  int rc = static_scratch(&base);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  uint32_t *a = base, *b = base + kStaticPairs;
  uint32_t *hist = base + 2 * kStaticPairs;
  size_t hist_cap = kStaticHist;  // large enough: radix_sort_pairs allocates nothing
  ScanScratch sc;
  sc.tile_sums = hist + kStaticHist;
  sc.tile_cap = kStaticScanTiles;
  const uint32_t *kout = nullptr, *vout = nullptr;
  rc = radix_sort_pairs(keys, values, a, b, keys, values, (uint32_t)n, key_bits, &hist,
                        &hist_cap, &sc, s, nullptr, &kout, &vout);
  if (rc) return rc;
  if (kout != keys) {  // an odd number of passes ends in the scratch buffers
    SPH_HIP(hipMemcpyAsync(keys, kout, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice,
                           s));
    SPH_HIP(hipMemcpyAsync(values, vout, sizeof(uint32_t) * (size_t)n,
                           hipMemcpyDeviceToDevice, s));
  }
  return SPH_OK;
}

static int bounds_impl(const uint32_t *keys, int n, int cells, uint32_t *out, void *stream,
                       bool upper) {
  if (n < 0 || cells < 0 || (cells > 0 && !out) || (n > 0 && !keys)) {
    set_error("sph_find_bounds: invalid argument");
    return SPH_ERR_ARG;
  }
  if (cells == 0) return SPH_OK;
  hipStream_t s = (hipStream_t)stream;
  AbiScratch *w = nullptr;
  int rc = abi_scratch(&w);
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(w->m);
  if ((rc = abi_words(w, (size_t)cells + 1))) return rc;
  uint32_t *count = w->a, *start = w->b;
  rc = count_keys(keys, (uint32_t)n, count, (uint32_t)cells, s);
  if (rc == SPH_OK)
    rc = exclusive_scan(count, start, (uint32_t)cells, false, true, &w->sc, s, nullptr);
  if (rc == SPH_OK) {
    const hipError_t e = hipMemcpyAsync(out, start + (upper ? 1 : 0),
                                        sizeof(uint32_t) * (size_t)cells,
                                        hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) rc = hip_fail(e, "hipMemcpyAsync (bounds)");
  }
  const hipError_t e = hipStreamSynchronize(s);
  if (rc == SPH_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize (bounds)");
  return rc;
}

extern "C" int sph_find_lower_bounds(const uint32_t *keys, int n, int cells, uint32_t *out,
                                     void *stream) {
  return bounds_impl(keys, n, cells, out, stream, false);
}

extern "C" int sph_find_upper_bounds(const uint32_t *keys, int n, int cells, uint32_t *out,
                                     void *stream) {
  return bounds_impl(keys, n, cells, out, stream, true);
}
