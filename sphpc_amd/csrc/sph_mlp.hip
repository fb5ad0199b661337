// Lambda warm-start network: FP32 training kernels (forward, backward, Adam)
// and the BF16 MFMA inference path, behind the sph_mlp_* C ABI.
//
// The network has no counterpart in the reference; its inputs are reference
// quantities of the constraint solve (neighbour count, neighbors.h:27; density,
// particles.c:512; previous lambda, particles.c:664; |grad C|^2, particles.c:
// 654-661; SDF distance and gradient, obstacle_cuda.cu:33-53,88) plus the
// velocity divergence and first density moments.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "sph_internal.h"
#include "sph_mlp_dev.h"

namespace sph {

constexpr int kGradThreads = 256;
constexpr int kChunk = 64;          // samples per block iteration
constexpr int kLd = 65;             // padded LDS row stride (conflict-free columns)
constexpr int kXd = 17;
constexpr int kMaxGradBlocks = 1024;
constexpr int kMomentCols = 2 * (kMlpIn + 1);  // sums and squares of 12 features + target
constexpr int kMomentBlocks = 512;

// splitmix64 (Steele, Lea, Flood 2014) for counter-based batch sampling.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

__device__ __forceinline__ uint64_t batch_row(uint64_t k, uint64_t count, uint64_t seed,
                                              int random) {
  return random ? mix64(seed ^ mix64(k)) % count : k;
}

constexpr size_t kGradLdsFloats =
    kMlpParams + 3 + kMlpH * kLd /* W2 transposed */ + kChunk * kXd + 4 * kChunk * kLd + kChunk;

// Per-block partial gradient of L = 0.5/B sum (y - t_n)^2 over a batch, with
// x normalised by (x - mu) * inv_sigma and t_n = (t - y_shift) / y_scale.
// 256 threads work on 64 samples at a time: in the forward and backward
// passes thread (s = tid / 4, g = tid % 4) owns hidden units g + 4m of sample s.
__global__ __launch_bounds__(kGradThreads) void k_mlp_grad(
    uint64_t batch, uint64_t count, const float *__restrict__ x, const float *__restrict__ t,
    uint64_t seed, int random, const float *__restrict__ theta, const float *__restrict__ norm,
    float inv_batch, float *__restrict__ partial, float *__restrict__ loss_partial) {
  extern __shared__ float sm[];
  float *sW = sm;
  float *sW2T = sW + kMlpParams + 3;
  float *sX = sW2T + kMlpH * kLd;
  float *sH1 = sX + kChunk * kXd;
  float *sD1 = sH1 + kChunk * kLd;
  float *sH2 = sD1 + kChunk * kLd;
  float *sD2 = sH2 + kChunk * kLd;
  float *sdy = sD2 + kChunk * kLd;
  __shared__ float snorm[kNormWords];
  __shared__ float sloss[kGradThreads / 64];
  const int tid = threadIdx.x;
  for (int i = tid; i < kMlpParams; i += kGradThreads) sW[i] = theta[i];
  for (int i = tid; i < kMlpH * kMlpH; i += kGradThreads)
    sW2T[(i & 63) * kLd + (i >> 6)] = theta[kMlpW2 + i];
  if (tid < kNormWords) snorm[tid] = norm[tid];
  __syncthreads();
  const float *W1 = sW + kMlpW1, *b1 = sW + kMlpB1, *W2 = sW + kMlpW2, *b2 = sW + kMlpB2,
              *w3 = sW + kMlpW3;
  const float b3 = sW[kMlpB3];
  const float y_shift = snorm[kNormYShift], inv_scale = 1.0f / snorm[kNormYScale];
  const int sf = tid >> 2, gf = tid & 3;
  float gW2[16], gW1[4], gv = 0.0f, loss = 0.0f;
#pragma unroll
  for (int m = 0; m < 16; ++m) gW2[m] = 0.0f;
#pragma unroll
  for (int m = 0; m < 4; ++m) gW1[m] = 0.0f;

  const uint64_t nchunks = (batch + kChunk - 1) / kChunk;
  for (uint64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    for (int e = tid; e < kChunk * kMlpPad; e += kGradThreads) {
      const int s = e >> 4, i = e & 15;
      const uint64_t k = ch * kChunk + s;
      float val = 0.0f;
      if (k < batch && i < kMlpIn) {
        const uint64_t row = batch_row(k, count, seed, random);
        val = (x[row * kMlpPad + i] - snorm[kNormMu + i]) * snorm[kNormIsd + i];
      }
      sX[s * kXd + i] = val;
    }
    __syncthreads();
    {  // layer 1
      float z[16];
#pragma unroll
      for (int m = 0; m < 16; ++m) z[m] = b1[gf + 4 * m];
      for (int i = 0; i < kMlpIn; ++i) {
        const float xi = sX[sf * kXd + i];
#pragma unroll
        for (int m = 0; m < 16; ++m) z[m] += xi * W1[i * kMlpH + gf + 4 * m];
      }
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        sH1[sf * kLd + gf + 4 * m] = silu(z[m]);
        sD1[sf * kLd + gf + 4 * m] = dsilu(z[m]);
      }
    }
    __syncthreads();
    float yp = 0.0f;
    {  // layer 2 and the output unit
      float z[16];
#pragma unroll
      for (int m = 0; m < 16; ++m) z[m] = b2[gf + 4 * m];
      for (int k = 0; k < kMlpH; ++k) {
        const float hk = sH1[sf * kLd + k];
#pragma unroll
        for (int m = 0; m < 16; ++m) z[m] += hk * W2[k * kMlpH + gf + 4 * m];
      }
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const float h = silu(z[m]);
        sH2[sf * kLd + gf + 4 * m] = h;
        sD2[sf * kLd + gf + 4 * m] = dsilu(z[m]);
        yp += w3[gf + 4 * m] * h;
      }
    }
    yp += __shfl_xor(yp, 1, 64);
    yp += __shfl_xor(yp, 2, 64);
    if (gf == 0) {
      const uint64_t k = ch * kChunk + sf;
      float dy = 0.0f;
      if (k < batch) {
        const uint64_t row = batch_row(k, count, seed, random);
        const float r = yp + b3 - (t[row] - y_shift) * inv_scale;
        loss += r * r;
        dy = r * inv_batch;
      }
      sdy[sf] = dy;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 16; ++m) sD2[sf * kLd + gf + 4 * m] *= sdy[sf] * w3[gf + 4 * m];  // dZ2
    __syncthreads();
    {  // dZ1 = (dZ2 W2^T) * silu'(Z1)
      float a[16];
#pragma unroll
      for (int m = 0; m < 16; ++m) a[m] = 0.0f;
      for (int k = 0; k < kMlpH; ++k) {
        const float dk = sD2[sf * kLd + k];
#pragma unroll
        for (int m = 0; m < 16; ++m) a[m] += dk * sW2T[k * kLd + gf + 4 * m];
      }
#pragma unroll
      for (int m = 0; m < 16; ++m) sD1[sf * kLd + gf + 4 * m] *= a[m];
    }
    __syncthreads();
    {  // dW2[j][g + 4m] += sum_s H1[s][j] dZ2[s][g + 4m]
      const int j = tid >> 2, g = tid & 3;
      for (int s = 0; s < kChunk; ++s) {
        const float hj = sH1[s * kLd + j];
#pragma unroll
        for (int m = 0; m < 16; ++m) gW2[m] += hj * sD2[s * kLd + g + 4 * m];
      }
    }
    {  // dW1[i][jj + 16m] += sum_s X[s][i] dZ1[s][jj + 16m]
      const int i = tid >> 4, jj = tid & 15;
      for (int s = 0; s < kChunk; ++s) {
        const float xi = sX[s * kXd + i];
#pragma unroll
        for (int m = 0; m < 4; ++m) gW1[m] += xi * sD1[s * kLd + jj + 16 * m];
      }
    }
    if (tid < 64) {
      for (int s = 0; s < kChunk; ++s) gv += sD1[s * kLd + tid];
    } else if (tid < 128) {
      for (int s = 0; s < kChunk; ++s) gv += sD2[s * kLd + tid - 64];
    } else if (tid < 192) {
      for (int s = 0; s < kChunk; ++s) gv += sH2[s * kLd + tid - 128] * sdy[s];
    } else if (tid == 192) {
      for (int s = 0; s < kChunk; ++s) gv += sdy[s];
    }
    __syncthreads();
  }
  float *P = partial + (size_t)blockIdx.x * kMlpParams;
  {
    const int j = tid >> 2, g = tid & 3;
#pragma unroll
    for (int m = 0; m < 16; ++m) P[kMlpW2 + j * kMlpH + g + 4 * m] = gW2[m];
  }
  {
    const int i = tid >> 4, jj = tid & 15;
#pragma unroll
    for (int m = 0; m < 4; ++m) P[kMlpW1 + i * kMlpH + jj + 16 * m] = gW1[m];
  }
  if (tid < 64) P[kMlpB1 + tid] = gv;
  else if (tid < 128) P[kMlpB2 + tid - 64] = gv;
  else if (tid < 192) P[kMlpW3 + tid - 128] = gv;
  else if (tid == 192) P[kMlpB3] = gv;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) loss += __shfl_xor(loss, o, 64);
  if ((tid & 63) == 0) sloss[tid >> 6] = loss;
  __syncthreads();
  if (tid == 0) {
    float l = 0.0f;
    for (int w = 0; w < kGradThreads / 64; ++w) l += sloss[w];
    loss_partial[blockIdx.x] = 0.5f * l;
  }
}

// Sums the block partials in a fixed order (deterministic), zeroes the
// gradient of the padding inputs and optionally applies one Adam step
// (Kingma & Ba, beta1 0.9, beta2 0.999, eps 1e-8; bc = bias corrections).
__global__ __launch_bounds__(kBlock) void k_mlp_adam(int nblocks, const float *__restrict__ partial,
                                                     const float *__restrict__ loss_partial,
                                                     float *__restrict__ theta, float *__restrict__ mom,
                                                     float *__restrict__ vel, float *__restrict__ grad,
                                                     float *__restrict__ loss_out, float inv_batch,
                                                     float lr, float bc1, float bc2, int apply) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p < kMlpParams) {
    float g = 0.0f;
    for (int b = 0; b < nblocks; ++b) g += partial[(size_t)b * kMlpParams + p];
    if (p >= kMlpW1 + kMlpIn * kMlpH && p < kMlpB1) g = 0.0f;
    grad[p] = g;
    if (apply) {
      const float m = 0.9f * mom[p] + 0.1f * g;
      const float v = 0.999f * vel[p] + 0.001f * g * g;
      mom[p] = m;
      vel[p] = v;
      theta[p] -= lr * (m / bc1) / (sqrtf(v / bc2) + 1e-8f);
    }
  }
  if (p == 0) {
    float l = 0.0f;
    for (int b = 0; b < nblocks; ++b) l += loss_partial[b];
    *loss_out = l * inv_batch;
  }
}

// FP32 forward pass, one thread per sample (weights broadcast from LDS).
__global__ __launch_bounds__(kBlock) void k_mlp_forward(uint64_t count, const float *__restrict__ x,
                                                        const float *__restrict__ theta,
                                                        const float *__restrict__ norm,
                                                        float *__restrict__ y) {
  __shared__ float sW[kMlpParams];
  __shared__ float snorm[kNormWords];
  for (int i = threadIdx.x; i < kMlpParams; i += kBlock) sW[i] = theta[i];
  if (threadIdx.x < kNormWords) snorm[threadIdx.x] = norm[threadIdx.x];
  __syncthreads();
  const uint64_t s = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (s >= count) return;
  float xn[kMlpIn];
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i)
    xn[i] = (x[s * kMlpPad + i] - snorm[kNormMu + i]) * snorm[kNormIsd + i];
  float h1[kMlpH];
#pragma unroll
  for (int j = 0; j < kMlpH; ++j) {
    float z = sW[kMlpB1 + j];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) z += xn[i] * sW[kMlpW1 + i * kMlpH + j];
    h1[j] = silu(z);
  }
  float out = sW[kMlpB3];
  for (int k = 0; k < kMlpH; ++k) {
    float z = sW[kMlpB2 + k];
#pragma unroll
    for (int j = 0; j < kMlpH; ++j) z += h1[j] * sW[kMlpW2 + j * kMlpH + k];
    out += sW[kMlpW3 + k] * silu(z);
  }
  y[s] = out * snorm[kNormYScale] + snorm[kNormYShift];
}

// BF16 MFMA inference of a standalone batch (the path fused into the solver).
__global__ __launch_bounds__(kBlock) void k_mlp_infer(uint64_t count, const float *__restrict__ x,
                                                      const MlpPacked *__restrict__ M,
                                                      float *__restrict__ y) {
  const uint64_t s = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  float f[kMlpPad];
#pragma unroll
  for (int i = 0; i < kMlpPad; ++i) f[i] = 0.0f;
  if (s < count) {
    const float4 *r = reinterpret_cast<const float4 *>(x + s * kMlpPad);
    const float4 a = r[0], b = r[1], c = r[2];
    f[0] = a.x; f[1] = a.y; f[2] = a.z; f[3] = a.w;
    f[4] = b.x; f[5] = b.y; f[6] = b.z; f[7] = b.w;
    f[8] = c.x; f[9] = c.y; f[10] = c.z; f[11] = c.w;
  }
  const float out = mlp_wave(f, M);  // whole wave, including lanes past the end
  if (s < count) y[s] = out;
}

// Packs FP32 master weights into MFMA operand order (BF16) for inference.
__global__ __launch_bounds__(kBlock) void k_mlp_pack(const float *__restrict__ theta,
                                                     const float *__restrict__ norm,
                                                     MlpPacked *__restrict__ M) {
  const int tid = threadIdx.x;
  for (int e = tid; e < 2 * 64; e += kBlock) {  // a1[rb][lane]
    const int rb = e >> 6, l = e & 63, half = l >> 5, col = 32 * rb + (l & 31);
    uint32_t w[4];
    for (int q = 0; q < 4; ++q) {
      const int k0 = 8 * half + 2 * q;
      w[q] = pack_bf16x2(theta[kMlpW1 + k0 * kMlpH + col], theta[kMlpW1 + (k0 + 1) * kMlpH + col]);
    }
    M->a1[e] = make_uint4(w[0], w[1], w[2], w[3]);
  }
  for (int e = tid; e < 2 * 4 * 64; e += kBlock) {  // a2[ob][kc][lane]
    const int ob = e >> 8, kc = (e >> 6) & 3, l = e & 63, half = l >> 5, col = 32 * ob + (l & 31);
    uint32_t w[4];
    for (int q = 0; q < 4; ++q) {
      const int r0 = mlp_l2_hidden(kc, half, 2 * q), r1 = mlp_l2_hidden(kc, half, 2 * q + 1);
      w[q] = pack_bf16x2(theta[kMlpW2 + r0 * kMlpH + col], theta[kMlpW2 + r1 * kMlpH + col]);
    }
    M->a2[e] = make_uint4(w[0], w[1], w[2], w[3]);
  }
  if (tid < 64) {  // [rb][half][v]
    const int rb = tid >> 5, half = (tid >> 4) & 1, v = tid & 15;
    const int u = 32 * rb + mfma_drow(v, half);
    reinterpret_cast<float *>(M->b1)[tid] = theta[kMlpB1 + u];
    reinterpret_cast<float *>(M->b2)[tid] = theta[kMlpB2 + u];
    reinterpret_cast<float *>(M->w3)[tid] = bf16_round(theta[kMlpW3 + u]);
  }
  if (tid < kNormWords) M->norm[tid] = norm[tid];
  if (tid == 0) M->b3 = theta[kMlpB3];
}

// Per-block sums and squares (FP64) of the 12 features and the target.
__global__ __launch_bounds__(kBlock) void k_mlp_moments(uint64_t count, const float *__restrict__ x,
                                                        const float *__restrict__ t,
                                                        double *__restrict__ partial) {
  __shared__ double red[kBlock / 64][kMomentCols];
  double acc[kMomentCols];
#pragma unroll
  for (int c = 0; c < kMomentCols; ++c) acc[c] = 0.0;
  for (uint64_t s = (uint64_t)blockIdx.x * kBlock + threadIdx.x; s < count;
       s += (uint64_t)gridDim.x * kBlock) {
#pragma unroll
    for (int i = 0; i <= kMlpIn; ++i) {
      const double v = i < kMlpIn ? (double)x[s * kMlpPad + i] : (double)t[s];
      acc[2 * i] += v;
      acc[2 * i + 1] += v * v;
    }
  }
#pragma unroll
  for (int c = 0; c < kMomentCols; ++c) {
    double v = acc[c];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < kMomentCols) {
    double v = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) v += red[w][threadIdx.x];
    partial[(size_t)blockIdx.x * kMomentCols + threadIdx.x] = v;
  }
}

static int mlp_configure() {
  static int done = 0;
  if (done) return SPH_OK;
  SPH_HIP(hipFuncSetAttribute((const void *)k_mlp_grad,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(kGradLdsFloats * sizeof(float))));
  done = 1;
  return SPH_OK;
}

}  // namespace sph

using namespace sph;

struct sph_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float *theta = nullptr, *mom = nullptr, *vel = nullptr, *grad = nullptr;
  float *norm = nullptr;
  float *partial = nullptr, *loss_partial = nullptr, *loss = nullptr;
  double *moments = nullptr;
  MlpPacked *packed = nullptr;
  float host_norm[kNormWords] = {};
  long long step = 0;
};

static int mlp_arg(const char *msg) {
  set_error(msg);
  return SPH_ERR_ARG;
}

static int upload_norm(sph_mlp *m) {
  SPH_HIP(hipMemcpyAsync(m->norm, m->host_norm, sizeof(m->host_norm), hipMemcpyHostToDevice,
                         m->stream));
  SPH_HIP(hipStreamSynchronize(m->stream));
  return SPH_OK;
}

static void mlp_free(sph_mlp *m) {
  void *bufs[] = {m->theta, m->mom, m->vel, m->grad, m->norm, m->partial,
                  m->loss_partial, m->loss, m->moments, m->packed};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (m->stream) (void)hipStreamDestroy(m->stream);
}

extern "C" int sph_mlp_create(sph_mlp **out) {
  if (!out) return mlp_arg("NULL argument");
  *out = nullptr;
  int rc = mlp_configure();
  if (rc) return rc;
  sph_mlp *m = new sph_mlp();
  hipError_t e = hipGetDevice(&m->device);
  if (e == hipSuccess) e = hipStreamCreate(&m->stream);
  struct { void **p; size_t bytes; } allocs[] = {
      {(void **)&m->theta, sizeof(float) * kMlpParams},
      {(void **)&m->mom, sizeof(float) * kMlpParams},
      {(void **)&m->vel, sizeof(float) * kMlpParams},
      {(void **)&m->grad, sizeof(float) * kMlpParams},
      {(void **)&m->norm, sizeof(float) * kNormWords},
      {(void **)&m->partial, sizeof(float) * kMlpParams * kMaxGradBlocks},
      {(void **)&m->loss_partial, sizeof(float) * kMaxGradBlocks},
      {(void **)&m->loss, sizeof(float)},
      {(void **)&m->moments, sizeof(double) * kMomentCols * kMomentBlocks},
      {(void **)&m->packed, sizeof(MlpPacked)},
  };
  for (auto &a : allocs)
    if (e == hipSuccess) e = hipMalloc(a.p, a.bytes);
  for (int i = 0; i < 3 && e == hipSuccess; ++i)
    e = hipMemsetAsync(i == 0 ? m->theta : i == 1 ? m->mom : m->vel, 0,
                       sizeof(float) * kMlpParams, m->stream);
  if (e != hipSuccess) {
    mlp_free(m);
    delete m;
    return hip_fail(e, "sph_mlp_create");
  }
  for (int i = 0; i < kMlpIn; ++i) m->host_norm[kNormIsd + i] = 1.0f;
  m->host_norm[kNormYScale] = 1.0f;
  rc = upload_norm(m);
  if (rc) {
    mlp_free(m);
    delete m;
    return rc;
  }
  *out = m;
  return SPH_OK;
}

extern "C" int sph_mlp_destroy(sph_mlp *m) {
  if (!m) return SPH_OK;
  (void)hipStreamSynchronize(m->stream);
  mlp_free(m);
  delete m;
  return SPH_OK;
}

extern "C" int sph_mlp_set_weights(sph_mlp *m, const float *theta) {
  if (!m || !theta) return mlp_arg("NULL argument");
  SPH_HIP(hipMemcpyAsync(m->theta, theta, sizeof(float) * kMlpParams, hipMemcpyHostToDevice,
                         m->stream));
  SPH_HIP(hipStreamSynchronize(m->stream));
  return SPH_OK;
}

extern "C" int sph_mlp_get_weights(sph_mlp *m, float *theta) {
  if (!m || !theta) return mlp_arg("NULL argument");
  SPH_HIP(hipMemcpyAsync(theta, m->theta, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost,
                         m->stream));
  SPH_HIP(hipStreamSynchronize(m->stream));
  return SPH_OK;
}

// He-style uniform initialisation: W ~ U(-1, 1) * sqrt(3 / fan_in), zero biases,
// zero rows for the padding inputs.  Also resets the optimiser.
extern "C" int sph_mlp_init(sph_mlp *m, unsigned long long seed) {
  if (!m) return mlp_arg("NULL argument");
  std::vector<float> th(kMlpParams, 0.0f);
  uint64_t ctr = seed * 0x2545f4914f6cdd1dull;
  auto uni = [&]() { return (double)(mix64(ctr++) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0; };
  for (int i = 0; i < kMlpIn; ++i)
    for (int j = 0; j < kMlpH; ++j) th[kMlpW1 + i * kMlpH + j] = (float)(uni() * sqrt(3.0 / kMlpIn));
  for (int i = 0; i < kMlpH * kMlpH; ++i) th[kMlpW2 + i] = (float)(uni() * sqrt(3.0 / kMlpH));
  for (int i = 0; i < kMlpH; ++i) th[kMlpW3 + i] = (float)(uni() * sqrt(3.0 / kMlpH));
  int rc = sph_mlp_set_weights(m, th.data());
  if (rc) return rc;
  SPH_HIP(hipMemsetAsync(m->mom, 0, sizeof(float) * kMlpParams, m->stream));
  SPH_HIP(hipMemsetAsync(m->vel, 0, sizeof(float) * kMlpParams, m->stream));
  SPH_HIP(hipStreamSynchronize(m->stream));
  m->step = 0;
  return SPH_OK;
}

extern "C" int sph_mlp_reset_optimizer(sph_mlp *m) {
  if (!m) return mlp_arg("NULL argument");
  SPH_HIP(hipMemsetAsync(m->mom, 0, sizeof(float) * kMlpParams, m->stream));
  SPH_HIP(hipMemsetAsync(m->vel, 0, sizeof(float) * kMlpParams, m->stream));
  SPH_HIP(hipStreamSynchronize(m->stream));
  m->step = 0;
  return SPH_OK;
}

extern "C" int sph_mlp_set_normalization(sph_mlp *m, const float *mean, const float *stddev,
                                         float target_mean, float target_std) {
  if (!m || !mean || !stddev) return mlp_arg("NULL argument");
  if (!(target_std > 0.0f)) return mlp_arg("target_std must be positive");
  for (int i = 0; i < kMlpIn; ++i) {
    if (!(stddev[i] >= 0.0f)) return mlp_arg("feature stddev must be >= 0");
    m->host_norm[kNormMu + i] = mean[i];
    m->host_norm[kNormIsd + i] = stddev[i] > 0.0f ? 1.0f / stddev[i] : 0.0f;
  }
  m->host_norm[kNormYShift] = target_mean;
  m->host_norm[kNormYScale] = target_std;
  return upload_norm(m);
}

extern "C" int sph_mlp_set_output_clip(sph_mlp *m, float max_target) {
  if (!m) return mlp_arg("NULL argument");
  if (!(max_target >= 0.0f)) return mlp_arg("max_target must be >= 0 (0: no clip)");
  m->host_norm[kNormYMax] = max_target;
  return upload_norm(m);
}

extern "C" float sph_mlp_get_output_clip(sph_mlp *m) { return m ? m->host_norm[kNormYMax] : 0.0f; }

extern "C" int sph_mlp_set_step_scale(sph_mlp *m, float scale) {
  if (!m) return mlp_arg("NULL argument");
  if (!(scale > 0.0f)) return mlp_arg("scale must be positive");
  m->host_norm[kNormStep] = scale;
  return upload_norm(m);
}

extern "C" float sph_mlp_get_step_scale(sph_mlp *m) {
  return (m && m->host_norm[kNormStep] > 0.0f) ? m->host_norm[kNormStep] : 1.0f;
}

extern "C" int sph_mlp_get_normalization(sph_mlp *m, float *mean, float *stddev,
                                         float *target_mean, float *target_std) {
  if (!m) return mlp_arg("NULL argument");
  for (int i = 0; i < kMlpIn; ++i) {
    if (mean) mean[i] = m->host_norm[kNormMu + i];
    const float isd = m->host_norm[kNormIsd + i];
    if (stddev) stddev[i] = isd > 0.0f ? 1.0f / isd : 0.0f;
  }
  if (target_mean) *target_mean = m->host_norm[kNormYShift];
  if (target_std) *target_std = m->host_norm[kNormYScale];
  return SPH_OK;
}

// Mean and standard deviation of the features and the target over a device
// dataset (FP64 sums), installed as the network's normalisation.
extern "C" int sph_mlp_fit_normalization(sph_mlp *m, const float *x, const float *t,
                                         size_t count) {
  if (!m || !x || !t) return mlp_arg("NULL argument");
  if (count < 2) return mlp_arg("need at least two samples");
  const int nb = (int)std::min<size_t>(kMomentBlocks, (count + kBlock - 1) / kBlock);
  k_mlp_moments<<<nb, kBlock, 0, m->stream>>>(count, x, t, m->moments);
  SPH_HIP(hipGetLastError());
  std::vector<double> part((size_t)nb * kMomentCols);
  SPH_HIP(hipMemcpyAsync(part.data(), m->moments, sizeof(double) * part.size(),
                         hipMemcpyDeviceToHost, m->stream));
  SPH_HIP(hipStreamSynchronize(m->stream));
  double sum[kMomentCols] = {};
  for (int b = 0; b < nb; ++b)
    for (int c = 0; c < kMomentCols; ++c) sum[c] += part[(size_t)b * kMomentCols + c];
  float mean[kMlpIn], sd[kMlpIn], tm = 0.0f, ts = 1.0f;
  for (int i = 0; i <= kMlpIn; ++i) {
    const double mu = sum[2 * i] / (double)count;
    const double var = std::max(0.0, sum[2 * i + 1] / (double)count - mu * mu);
    double s = sqrt(var);
    if (s < 1e-12 * std::max(1.0, fabs(mu))) s = 0.0;
    if (i < kMlpIn) {
      mean[i] = (float)mu;
      sd[i] = (float)s;
    } else {
      tm = (float)mu;
      ts = s > 0.0 ? (float)s : 1.0f;
    }
  }
  return sph_mlp_set_normalization(m, mean, sd, tm, ts);
}

static int grad_pass(sph_mlp *m, const float *x, const float *t, size_t count, size_t batch,
                     uint64_t seed, int random, float lr, int apply) {
  const uint64_t nchunks = (batch + kChunk - 1) / kChunk;
  const int nb = (int)std::min<uint64_t>(kMaxGradBlocks, nchunks);
  const float inv_batch = (float)(1.0 / (double)batch);
  k_mlp_grad<<<nb, kGradThreads, kGradLdsFloats * sizeof(float), m->stream>>>(
      batch, count, x, t, seed, random, m->theta, m->norm, inv_batch, m->partial,
      m->loss_partial);
  SPH_HIP(hipGetLastError());
  float bc1 = 1.0f, bc2 = 1.0f;
  if (apply) {
    m->step += 1;
    bc1 = (float)(1.0 - pow(0.9, (double)m->step));
    bc2 = (float)(1.0 - pow(0.999, (double)m->step));
  }
  k_mlp_adam<<<blocks_for(kMlpParams), kBlock, 0, m->stream>>>(
      nb, m->partial, m->loss_partial, m->theta, m->mom, m->vel, m->grad, m->loss, inv_batch, lr,
      bc1, bc2, apply);
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}

extern "C" int sph_mlp_gradient(sph_mlp *m, const float *x, const float *t, size_t count,
                                float *grad, float *loss) {
  if (!m || !x || !t) return mlp_arg("NULL argument");
  if (count == 0) return mlp_arg("empty batch");
  int rc = grad_pass(m, x, t, count, count, 0, 0, 0.0f, 0);
  if (rc) return rc;
  if (grad)
    SPH_HIP(hipMemcpyAsync(grad, m->grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost,
                           m->stream));
  if (loss)
    SPH_HIP(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  SPH_HIP(hipStreamSynchronize(m->stream));
  return SPH_OK;
}

extern "C" int sph_mlp_train_step(sph_mlp *m, const float *x, const float *t, size_t count,
                                  size_t batch, float lr, unsigned long long seed, float *loss) {
  if (!m || !x || !t) return mlp_arg("NULL argument");
  if (count == 0) return mlp_arg("empty dataset");
  if (!(lr > 0.0f)) return mlp_arg("learning rate must be positive");
  const int random = batch != 0;
  if (!random) batch = count;
  int rc = grad_pass(m, x, t, count, batch, seed, random, lr, 1);
  if (rc) return rc;
  if (loss) {
    SPH_HIP(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    SPH_HIP(hipStreamSynchronize(m->stream));
  }
  return SPH_OK;
}

extern "C" int sph_mlp_forward(sph_mlp *m, const float *x, size_t count, float *y) {
  if (!m || (count && (!x || !y))) return mlp_arg("NULL argument");
  if (!count) return SPH_OK;
  k_mlp_forward<<<(unsigned)((count + kBlock - 1) / kBlock), kBlock, 0, m->stream>>>(
      count, x, m->theta, m->norm, y);
  SPH_HIP(hipGetLastError());
  SPH_HIP(hipStreamSynchronize(m->stream));
  return SPH_OK;
}

namespace sph {
int mlp_pack(sph_mlp *m, MlpPacked *dst, hipStream_t s) {
  SPH_HIP(hipStreamSynchronize(m->stream));
  k_mlp_pack<<<1, kBlock, 0, s>>>(m->theta, m->norm, dst);
  SPH_HIP(hipGetLastError());
  return SPH_OK;
}
}  // namespace sph

extern "C" int sph_mlp_infer_bf16(sph_mlp *m, const float *x, size_t count, float *y) {
  if (!m || (count && (!x || !y))) return mlp_arg("NULL argument");
  if (((uintptr_t)x & 15u) != 0) return mlp_arg("features must be 16-byte aligned");
  int rc = mlp_pack(m, m->packed, m->stream);
  if (rc || !count) return rc;
  k_mlp_infer<<<(unsigned)((count + kBlock - 1) / kBlock), kBlock, 0, m->stream>>>(
      count, x, m->packed, y);
  SPH_HIP(hipGetLastError());
  SPH_HIP(hipStreamSynchronize(m->stream));
  return SPH_OK;
}

extern "C" int sph_mlp_synchronize(sph_mlp *m) {
  if (!m) return mlp_arg("NULL argument");
  SPH_HIP(hipStreamSynchronize(m->stream));
  return SPH_OK;
}

// File format: 8-byte magic "SPHMLP01", int32 parameter count, int32 norm
// word count, FP32 parameters, FP32 normalisation words (little endian).
static const char kMagic[8] = {'S', 'P', 'H', 'M', 'L', 'P', '0', '1'};

extern "C" int sph_mlp_save(sph_mlp *m, const char *path) {
  if (!m || !path) return mlp_arg("NULL argument");
  std::vector<float> th(kMlpParams);
  int rc = sph_mlp_get_weights(m, th.data());
  if (rc) return rc;
  FILE *f = fopen(path, "wb");
  if (!f) return mlp_arg("cannot open file for writing");
  const int32_t hdr[2] = {kMlpParams, kNormWords};
  bool ok = fwrite(kMagic, 1, 8, f) == 8 && fwrite(hdr, 4, 2, f) == 2 &&
            fwrite(th.data(), 4, kMlpParams, f) == (size_t)kMlpParams &&
            fwrite(m->host_norm, 4, kNormWords, f) == (size_t)kNormWords;
  ok = (fclose(f) == 0) && ok;
  return ok ? SPH_OK : mlp_arg("write failed");
}

extern "C" int sph_mlp_load(sph_mlp *m, const char *path) {
  if (!m || !path) return mlp_arg("NULL argument");
  FILE *f = fopen(path, "rb");
  if (!f) return mlp_arg("cannot open network file");
  char magic[8];
  int32_t hdr[2];
  std::vector<float> th(kMlpParams);
  float nrm[kNormWords];
  bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, kMagic, 8) == 0 &&
            fread(hdr, 4, 2, f) == 2 && hdr[0] == kMlpParams && hdr[1] == kNormWords &&
            fread(th.data(), 4, kMlpParams, f) == (size_t)kMlpParams &&
            fread(nrm, 4, kNormWords, f) == (size_t)kNormWords;
  fclose(f);
  if (!ok) return mlp_arg("not a network file of this layout");
  memcpy(m->host_norm, nrm, sizeof(nrm));
  int rc = upload_norm(m);
  if (rc) return rc;
  rc = sph_mlp_set_weights(m, th.data());
  if (rc) return rc;
  return sph_mlp_reset_optimizer(m);
}
