// Shared definitions of the clump-force MLP (3 -> 256 -> 256 -> 256 -> 4, SiLU) used by the
// training kernels (ssc_mlp.hip) and the fused surrogate integrator (ssc_surrogate.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ssc {
namespace mlp {

constexpr int IN = 3;
constexpr int H = 256;
constexpr int OUT = 4;
constexpr int OUTP = 16;  // output rows padded to one 16-row MFMA tile

// Offsets (floats) of the fp32 master parameter vector:
//   W1[H][IN] b1[H] W2[H][H] b2[H] W3[H][H] b3[H] W4[OUT][H] b4[OUT]
constexpr int64_t OW1 = 0;
constexpr int64_t OB1 = OW1 + H * IN;
constexpr int64_t OW2 = OB1 + H;
constexpr int64_t OB2 = OW2 + H * H;
constexpr int64_t OW3 = OB2 + H;
constexpr int64_t OB3 = OW3 + H * H;
constexpr int64_t OW4 = OB3 + H;
constexpr int64_t OB4 = OW4 + OUT * H;
constexpr int64_t NPARAM = OB4 + OUT;

// Packed inference form (bytes), laid out for v_mfma_f32_16x16x32_bf16 with the points on the
// MFMA N axis (16 per wave):
//   fragments of W2 and W3: [tile m (H/16)][k-step j (H/32)][lane 64][8 bf16], lane l holding
//   A[16 m + l % 16][32 j + 8 (l / 16) + e] with the k order permuted by kperm (below); W4 as
//   one 16-row tile (rows >= OUT zero): [j][lane][8];
//   then fp32 W1 rows (in natural neuron order) and fp32 biases b1, b2, b3, b4 (padded to 16).
constexpr int64_t FRAG_L = (int64_t)(H / 16) * (H / 32) * 64 * 8;  // bf16 elements per H x H layer
constexpr int64_t FRAG_4 = (int64_t)(H / 32) * 64 * 8;
constexpr int64_t PK_W2 = 0;                            // bf16 element offsets
constexpr int64_t PK_W3 = PK_W2 + FRAG_L;
constexpr int64_t PK_W4 = PK_W3 + FRAG_L;
constexpr int64_t PK_BF16_END = PK_W4 + FRAG_4;
constexpr int64_t PK_F32_BASE_BYTES = PK_BF16_END * 2;  // fp32 section starts here
constexpr int64_t PK_W1 = 0;                            // float offsets inside the fp32 section
constexpr int64_t PK_B1 = PK_W1 + H * IN;
constexpr int64_t PK_B2 = PK_B1 + H;
constexpr int64_t PK_B3 = PK_B2 + H;
constexpr int64_t PK_B4 = PK_B3 + H;
constexpr int64_t PK_F32_END = PK_B4 + OUTP;
constexpr int64_t PACKED_BYTES = PK_F32_BASE_BYTES + PK_F32_END * 4;

// Packed transposes for the input vector-Jacobian product: W3^T and W2^T fragments in the
// forward fragment order (rows = the layer's input neurons, k = its output neurons in kperm
// order; BF16 element offsets), then W4^T as fp32 [H][OUT] (BF16-rounded values; this small
// product runs on the VALU so the output cotangent stays fp32).
constexpr int64_t PKV_W3T = 0;
constexpr int64_t PKV_W2T = PKV_W3T + FRAG_L;
constexpr int64_t PKV_BF16_END = PKV_W2T + FRAG_L;
constexpr int64_t PKV_W4T_BYTES = PKV_BF16_END * 2;  // byte offset of the fp32 W4^T table
constexpr int64_t PACKED_VJP_BYTES = PKV_W4T_BYTES + (int64_t)H * OUT * 4;

// K order of every chained layer: within each block of 32 inputs, k-slot 8g + e of a B fragment
// (lane group g = l / 16, element e) is the neuron that lane group already holds from the
// previous layer's two 16x16 output tiles 2j and 2j + 1 (rows 4g + r of each):
//   e < 4: neuron 4g + e;  e >= 4: neuron 16 + 4g + (e - 4).
__host__ __device__ inline int kperm(int s) {
  const int blk = s & ~31, w = s & 31;
  const int g = w >> 3, e = w & 7;
  return blk + (e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4));
}

// SiLU z * sigmoid(z) with the hardware exp2 and reciprocal (about 1 ulp each; the result is
// rounded to BF16 by every consumer).  Training and inference use this same formula.
__device__ __forceinline__ float sigmoid_fast(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

__device__ __forceinline__ float silu(float z) { return z * sigmoid_fast(z); }

__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoid_fast(z);
  return s * (1.0f + z * (1.0f - s));
}

// fp32 -> bf16 bits, round to nearest even (NaN kept quiet).
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short i16x8 __attribute__((ext_vector_type(8)));

}  // namespace mlp
}  // namespace ssc
