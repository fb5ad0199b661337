// Device side of the lambda warm-start network: BF16 conversion, SiLU and the
// MFMA inference of one wavefront's 64 particles.  Included by the solver
// kernels (fused into the warm lambda pass) and by sph_mlp.hip.
#pragma once

#include "sph_internal.h"

namespace sph {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));

// Round-to-nearest-even FP32 -> BF16 bits (finite inputs).
__host__ __device__ __forceinline__ uint32_t bf16_bits(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}

__host__ __device__ __forceinline__ float bf16_round(float f) {
  const uint32_t u = bf16_bits(f) << 16;
  float r;
  __builtin_memcpy(&r, &u, 4);
  return r;
}

__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
  return bf16_bits(lo) | (bf16_bits(hi) << 16);
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
// SiLU of the BF16 inference from the hardware exp2 and reciprocal (v_exp_f32,
// v_rcp_f32, about 1 ulp each): a few VALU instead of the ~20 of expf and an
// IEEE division, 128 times per lane per wave.  Layer 1 rounds it to BF16
// (2^-8 relative) and layer 2 feeds an FP32 dot product checked to 2e-3, far
// above the few 1e-7 this changes.  Saturates to 0 for large negative z
// (exp2 -> inf, rcp -> 0) and to z for large positive z.
__device__ __forceinline__ float silu_infer(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// v_mfma_f32_32x32x16_bf16 operand layout, wave64:
//   A (32 x 16): lane l holds row l % 32, k = 8 * (l / 32) + i, i = 0..7
//   B (16 x 32): lane l holds column l % 32, k = 8 * (l / 32) + i
//   D (32 x 32): lane l, accumulator v holds column l % 32 and row mfma_drow(v, l / 32)
__host__ __device__ constexpr int mfma_drow(int v, int half) {
  return 8 * (v >> 2) + 4 * half + (v & 3);
}

// The network computes Z^T = W^T X^T, so hidden units are MFMA rows and the
// wave's particles are columns (two 32-particle column blocks).  Layer 1's D
// accumulators are re-used as layer 2's B operand without moving data: the
// k index of layer 2 is permuted to the hidden unit each lane already holds,
// k chunk c = 2 * row_block + vh, element (half, i) <-> hidden unit
// 32 * row_block + mfma_drow(8 * vh + i, half).  The packed A operands of
// layer 2 follow that permutation (k_mlp_pack).
__host__ __device__ constexpr int mlp_l2_hidden(int chunk, int half, int i) {
  return 32 * (chunk >> 1) + mfma_drow(8 * (chunk & 1) + i, half);
}

__device__ __forceinline__ bf16x8_t as_bf16x8(uint4 u) { return __builtin_bit_cast(bf16x8_t, u); }

// Runs the network for the 64 particles of the calling wave.  f holds the raw
// (unnormalised) features of this lane's particle; every lane of the wave must
// call it (inactive lanes pass zeros).  Returns the network output in target
// units (y * y_scale + y_shift).
__device__ __forceinline__ float mlp_wave(const float (&f)[kMlpPad], const MlpPacked *__restrict__ M) {
  const int lane = threadIdx.x & 63;
  const int half = lane >> 5;
  float xn[kMlpPad];
#pragma unroll
  for (int i = 0; i < kMlpPad; ++i) xn[i] = (f[i] - M->norm[kNormMu + i]) * M->norm[kNormIsd + i];
  uint32_t lo[4], hi[4], olo[4], ohi[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    lo[k] = pack_bf16x2(xn[2 * k], xn[2 * k + 1]);
    hi[k] = pack_bf16x2(xn[8 + 2 * k], xn[9 + 2 * k]);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    olo[k] = (uint32_t)__shfl_xor((int)lo[k], 32, 64);
    ohi[k] = (uint32_t)__shfl_xor((int)hi[k], 32, 64);
  }
  // column block 0 = particles 0..31, block 1 = particles 32..63
  const uint4 xb0 = half ? make_uint4(ohi[0], ohi[1], ohi[2], ohi[3])
                         : make_uint4(lo[0], lo[1], lo[2], lo[3]);
  const uint4 xb1 = half ? make_uint4(hi[0], hi[1], hi[2], hi[3])
                         : make_uint4(olo[0], olo[1], olo[2], olo[3]);
  const bf16x8_t xcol[2] = {as_bf16x8(xb0), as_bf16x8(xb1)};

  uint4 h1[2][4];  // [column block][k chunk] layer 2 B operands
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
    const bf16x8_t a = as_bf16x8(M->a1[rb * 64 + lane]);
    float bias[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 b = M->b1[(rb * 2 + half) * 4 + q];
      bias[4 * q] = b.x; bias[4 * q + 1] = b.y; bias[4 * q + 2] = b.z; bias[4 * q + 3] = b.w;
    }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      floatx16_t acc = {};
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, xcol[cb], acc, 0, 0, 0);
      float h[16];
#pragma unroll
      for (int v = 0; v < 16; ++v) h[v] = silu_infer(acc[v] + bias[v]);
      h1[cb][2 * rb] = make_uint4(pack_bf16x2(h[0], h[1]), pack_bf16x2(h[2], h[3]),
                                  pack_bf16x2(h[4], h[5]), pack_bf16x2(h[6], h[7]));
      h1[cb][2 * rb + 1] = make_uint4(pack_bf16x2(h[8], h[9]), pack_bf16x2(h[10], h[11]),
                                      pack_bf16x2(h[12], h[13]), pack_bf16x2(h[14], h[15]));
    }
  }
  float part[2] = {0.0f, 0.0f};
#pragma unroll
  for (int ob = 0; ob < 2; ++ob) {
    float bias[16], w[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 b = M->b2[(ob * 2 + half) * 4 + q];
      const float4 u = M->w3[(ob * 2 + half) * 4 + q];
      bias[4 * q] = b.x; bias[4 * q + 1] = b.y; bias[4 * q + 2] = b.z; bias[4 * q + 3] = b.w;
      w[4 * q] = u.x; w[4 * q + 1] = u.y; w[4 * q + 2] = u.z; w[4 * q + 3] = u.w;
    }
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      floatx16_t acc = {};
#pragma unroll
      for (int kc = 0; kc < 4; ++kc)
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(as_bf16x8(M->a2[(ob * 4 + kc) * 64 + lane]),
                                                      as_bf16x8(h1[cb][kc]), acc, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 16; ++v) part[cb] += w[v] * silu_infer(acc[v] + bias[v]);
    }
  }
#pragma unroll
  for (int cb = 0; cb < 2; ++cb) part[cb] += __shfl_xor(part[cb], 32, 64);
  const float y = (half ? part[1] : part[0]) + M->b3;
  return y * M->norm[kNormYScale] + M->norm[kNormYShift];
}

}  // namespace sph
