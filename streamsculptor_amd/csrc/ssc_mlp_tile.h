// Register-chained MLP forward for 32 points per wave (shared by ssc_mlp_infer and the
// surrogate integrator).  Lane l evaluates point l % 32; the upper half of the wave (l >= 32)
// carries the same point and supplies the other half of every MFMA operand.  The weight
// source is a policy: GlobalWeights reads fragments straight from the packed buffer, the
// integrator's policy stages each layer in LDS first (a block-wide operation with barriers,
// so every lane of the block must call the forward together).
#pragma once

#include "ssc_mlp.h"

namespace ssc {

// Eight fp32 values to one BF16 fragment (hardware round-to-nearest-even conversion).
__device__ __forceinline__ mlp::bf16x8 pack8(const float* v) {
  mlp::bf16x8 r;
#pragma unroll
  for (int e = 0; e < 8; ++e) r[e] = (__bf16)v[e];
  return r;
}

struct GlobalWeights {
  const uint16_t* pk16;
  // fragments of chained layer 0 (W2), 1 (W3) or 2 (W4)
  __device__ const mlp::bf16x8* layer(int which) const {
    const int64_t off = which == 0 ? mlp::PK_W2 : (which == 1 ? mlp::PK_W3 : mlp::PK_W4);
    return reinterpret_cast<const mlp::bf16x8*>(pk16 + off);
  }
};

// pk32: W1 (fp32, BF16-rounded values) and the biases, in the packed fp32 layout.
template <class Weights>
__device__ __forceinline__ void mlp_tile_forward(Weights& wsrc, const float* __restrict__ pk32,
                                                 float x0, float x1, float x2, float* y) {
  using namespace mlp;
  const int l = threadIdx.x & 63, hh = l >> 5;
  const float* w1 = pk32 + PK_W1;
  bf16x8 bin[H / 16];
  // layer 1 (VALU): k-slot 16q + 8hh + e of the layer-2 B operand is neuron kperm(.)
#pragma unroll
  for (int q = 0; q < H / 16; ++q) {
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int n = kperm(16 * q + 8 * hh + e);
      float z = pk32[PK_B1 + n];
      z = fmaf(x0, w1[3 * n], z);
      z = fmaf(x1, w1[3 * n + 1], z);
      z = fmaf(x2, w1[3 * n + 2], z);
      v[e] = silu(z);
    }
    bin[q] = pack8(v);
    __builtin_amdgcn_sched_barrier(0);
  }
  // two chained 256x256 layers; output tiles are produced two at a time so that only two
  // accumulators, the input fragments and the output fragments are live
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const bf16x8* W = wsrc.layer(layer);
    const float* bias = pk32 + (layer == 0 ? PK_B2 : PK_B3);
    bf16x8 bout[H / 16];
#pragma unroll
    for (int t = 0; t < H / 32; t += 2) {
      f32x16 a0, a1;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = 8 * (r >> 2) + 4 * hh + (r & 3);
        a0[r] = bias[32 * t + row];
        a1[r] = bias[32 * (t + 1) + row];
      }
#pragma unroll
      for (int q = 0; q < H / 16; ++q) {
        a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W[(t * (H / 16) + q) * 64 + l], bin[q], a0, 0, 0, 0);
        a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W[((t + 1) * (H / 16) + q) * 64 + l], bin[q], a1,
                                                    0, 0, 0);
        // bound how far the scheduler hoists fragment loads (register pressure)
        if ((q & 3) == 3) __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        float v0[8], v1[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          v0[e] = silu(a0[8 * u + e]);
          v1[e] = silu(a1[8 * u + e]);
        }
        bout[2 * t + u] = pack8(v0);
        bout[2 * (t + 1) + u] = pack8(v1);
      }
    }
#pragma unroll
    for (int q = 0; q < H / 16; ++q) bin[q] = bout[q];
  }
  // output layer (4 rows of one padded 32-row tile)
  const bf16x8* W4 = wsrc.layer(2);
  f32x16 out;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = 8 * (r >> 2) + 4 * hh + (r & 3);
    out[r] = row < OUT ? pk32[PK_B4 + row] : 0.0f;
  }
#pragma unroll
  for (int q = 0; q < H / 16; ++q)
    out = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W4[q * 64 + l], bin[q], out, 0, 0, 0);
#pragma unroll
  for (int j = 0; j < OUT; ++j) y[j] = out[j];
}

}  // namespace ssc
