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
constexpr int OUTP = 32;  // output rows padded to one 32-row MFMA tile

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

// Packed inference form (bytes):
//   fragments of W2, W3 (H/32 tiles x H/16 k-steps x 64 lanes x 8 bf16) and of W4 padded to
//   32 rows (1 tile), all with the k order permuted by kperm (below);
//   then fp32 W1 rows (in natural neuron order) and fp32 biases b1, b2, b3, b4.
constexpr int64_t FRAG_L = (int64_t)(H / 32) * (H / 16) * 64 * 8;  // bf16 elements per H x H layer
constexpr int64_t FRAG_4 = (int64_t)(H / 16) * 64 * 8;
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

// K order of every chained layer: within each block of 16 inputs the middle two quads swap,
// so that k-slot 8h + e of a B fragment (lane half h, element e) is exactly the neuron the
// lane already holds in accumulator register r = 8u + e of the previous layer's 32x32 output
// tile (rows 8(r/4) + 4h + r%4).  The permutation is its own inverse.
__host__ __device__ inline int kperm(int s) {
  const int blk = s & ~15, w = s & 15;
  const int h = w >> 3, e = w & 7;
  return blk + 8 * (e >> 2) + 4 * h + (e & 3);
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
typedef short i16x8 __attribute__((ext_vector_type(8)));

}  // namespace mlp
}  // namespace ssc
