// Register-chained MLP forward for 16 points per wave (shared by ssc_mlp_infer and the
// surrogate integrator), on v_mfma_f32_16x16x32_bf16 with the points on the MFMA N axis.
// Lane l evaluates point l % 16 together with the three other lanes of its column
// (l % 16 + 16 g, g = 0..3); lane group g = l / 16 holds a quarter of every activation vector:
// 64 of the 256 hidden neurons, i.e. 8 BF16 fragments (32 VGPRs) per layer input and output.
// A layer's output tile pair (2j, 2j + 1) leaves rows 4g + r of each in lane group g, which is
// exactly k-block j of the next layer's B operand once the weights' k order is permuted
// (kperm), so activations never leave the registers.  The weight source is a policy:
// GlobalWeights reads fragments straight from the packed buffer, the integrator's policy stages
// each layer in LDS first (a block-wide operation with barriers, so every lane of the block
// must call the forward together).
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

// An empty volatile statement that reads a finished fragment: the fragment must be computed
// here, in program order.  Without it LLVM sinks a layer's MFMAs and activations to their use
// in the next layer -- below the barrier that stages the next layer's weights in LDS -- while
// the fragment loads have to stay above it, so all 128 weight fragments of a layer end up live
// at once (and spilled).
__device__ __forceinline__ void pin(mlp::bf16x8& v) { asm volatile("" : "+v"(v)); }

// Weight policies.  A chained 256x256 layer is consumed in NQ chunks of 16 / NQ tiles each:
// chunk(c), c = 0 .. 2 NQ - 1, returns the fragments of chunk c of the two chained layers (W2
// first, then W3), and out() the output layer's fragments.  The vector-Jacobian product also
// uses chunkT(c) -- the same split of W3^T, then W2^T -- and w4t(), the fp32 W4^T table
// [H][OUT].
struct GlobalWeights {
  static constexpr int NQ = 2;
  const uint16_t* pk16;
  const uint16_t* pkv = nullptr;  // packed transposes (input VJP only)
  __device__ const mlp::bf16x8* chunk(int c) const {
    const int64_t off = (c / NQ == 0 ? mlp::PK_W2 : mlp::PK_W3) + (c % NQ) * (mlp::FRAG_L / NQ);
    return reinterpret_cast<const mlp::bf16x8*>(pk16 + off);
  }
  __device__ const mlp::bf16x8* out() const {
    return reinterpret_cast<const mlp::bf16x8*>(pk16 + mlp::PK_W4);
  }
  __device__ const mlp::bf16x8* chunkT(int c) const {
    const int64_t off = (c / NQ == 0 ? mlp::PKV_W3T : mlp::PKV_W2T) + (c % NQ) * (mlp::FRAG_L / NQ);
    return reinterpret_cast<const mlp::bf16x8*>(pkv + off);
  }
  __device__ const float* w4t() const {
    return reinterpret_cast<const float*>(reinterpret_cast<const char*>(pkv) + mlp::PKV_W4T_BYTES);
  }
};

// pk32: W1 (fp32, BF16-rounded values) and the biases, in the packed fp32 layout.
// y[0..3]: the OUT outputs of point l % 16, valid in lane group 0 (lanes 0..15).
template <class Weights>
__device__ __forceinline__ void mlp_tile_forward(Weights& wsrc, const float* __restrict__ pk32,
                                                 float x0, float x1, float x2, float* y) {
  using namespace mlp;
  constexpr int KB = H / 32;  // k-blocks of a 256-wide input
  // The lane index is made opaque to the optimiser: every weight, bias and fragment address
  // below is a lane-dependent base plus a constant, and inside the integrator's step loop the
  // compiler would otherwise hoist hundreds of such addresses out of the loop and keep them
  // live (then spill them).  Recomputing them from one base per call costs a few adds.
  int l = threadIdx.x & 63;
  asm volatile("" : "+v"(l));
  const int g = l >> 4;
  const float* w1 = pk32 + PK_W1;
  bf16x8 bin[KB];
  // layer 1 (fp32 VALU): k-slot 32j + 8g + e of the layer-2 B operand is neuron kperm(.)
#pragma unroll
  for (int j = 0; j < KB; ++j) {
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int n = kperm(32 * j + 8 * g + e);
      float z = pk32[PK_B1 + n];
      z = fmaf(x0, w1[3 * n], z);
      z = fmaf(x1, w1[3 * n + 1], z);
      z = fmaf(x2, w1[3 * n + 2], z);
      v[e] = silu(z);
    }
    bin[j] = pack8(v);
    pin(bin[j]);
  }
  // two chained 256x256 layers, each in NQ chunks of 16 / NQ tiles; output k-block j = tiles
  // 2j and 2j + 1 (tiles 2jj, 2jj + 1 of chunk j / (8 / NQ))
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const float* bias = pk32 + (layer == 0 ? PK_B2 : PK_B3);
    bf16x8 bout[KB];
#pragma unroll
    for (int qc = 0; qc < Weights::NQ; ++qc) {
      const bf16x8* W = wsrc.chunk(Weights::NQ * layer + qc);
#pragma unroll
      for (int jj = 0; jj < KB / Weights::NQ; ++jj) {
        const int j = (KB / Weights::NQ) * qc + jj;
        f32x4 a0, a1;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          a0[r] = bias[32 * j + 4 * g + r];
          a1[r] = bias[32 * j + 16 + 4 * g + r];
        }
#pragma unroll
        for (int q = 0; q < KB; ++q) {
          a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W[((2 * jj) * KB + q) * 64 + l], bin[q], a0,
                                                       0, 0, 0);
          a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W[((2 * jj + 1) * KB + q) * 64 + l], bin[q],
                                                       a1, 0, 0, 0);
        }
        float v[8];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          v[r] = silu(a0[r]);
          v[4 + r] = silu(a1[r]);
        }
        bout[j] = pack8(v);
        pin(bout[j]);
      }
    }
#pragma unroll
    for (int q = 0; q < KB; ++q) bin[q] = bout[q];
  }
  // output layer: rows 0..OUT-1 of one padded 16-row tile land in lane group 0
  const bf16x8* W4 = wsrc.out();
  f32x4 out;
#pragma unroll
  for (int r = 0; r < 4; ++r) out[r] = pk32[PK_B4 + 4 * g + r];
#pragma unroll
  for (int q = 0; q < KB; ++q)
    out = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W4[q * 64 + l], bin[q], out, 0, 0, 0);
#pragma unroll
  for (int j = 0; j < OUT; ++j) y[j] = out[j];
}

// Layer-1 table for mlp_tiles_forward: slot (4 j + g) 8 + e holds (W1[n][0], W1[n][1], W1[n][2],
// b1[n]) of neuron n = kperm(32 j + 8 g + e), i.e. lane group g reads its 64 neurons as 64
// consecutive float4 (one broadcast LDS read per neuron instead of four scattered ones).
// Filled cooperatively by the threads of a block (tid, nthreads).
__device__ __forceinline__ void build_l1_table(const float* __restrict__ pk32, float4* tab, int tid,
                                               int nthreads) {
  using namespace mlp;
  for (int s = tid; s < H; s += nthreads) {
    const int n = kperm(s);  // s = 32 j + 8 g + e
    tab[s] = make_float4(pk32[PK_W1 + 3 * n], pk32[PK_W1 + 3 * n + 1], pk32[PK_W1 + 3 * n + 2],
                         pk32[PK_B1 + n]);
  }
}

// NT tiles of 16 points per wave (point t of lane l is point l % 16 of tile t), every weight
// fragment read once from the weight source and applied to all NT tiles: the matrix-core
// operand traffic per point drops NT-fold while each point's arithmetic -- the order of every
// FMA and MFMA accumulation -- is that of mlp_tile_forward, so the outputs are bit for bit the
// same.  Layer 1 reads the table of build_l1_table.  xin[t]: inputs of point t (every lane of
// the column); y[t][0..3]: outputs of point t, valid in lane group 0.
template <int NT, class Weights>
__device__ __forceinline__ void mlp_tiles_forward(Weights& wsrc, const float* __restrict__ pk32,
                                                  const float4* __restrict__ l1tab,
                                                  const float (&xin)[NT][3], float (&y)[NT][4]) {
  using namespace mlp;
  constexpr int KB = H / 32;
  int l = threadIdx.x & 63;
  asm volatile("" : "+v"(l));
  const int g = l >> 4;
  const float4* tab = l1tab + 8 * g;
  bf16x8 bin[NT][KB];
#pragma unroll
  for (int j = 0; j < KB; ++j) {
    float v[NT][8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float4 w = tab[32 * j + e];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        float z = w.w;
        z = fmaf(xin[t][0], w.x, z);
        z = fmaf(xin[t][1], w.y, z);
        z = fmaf(xin[t][2], w.z, z);
        v[t][e] = silu(z);
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      bin[t][j] = pack8(v[t]);
      pin(bin[t][j]);
    }
  }
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const float* bias = pk32 + (layer == 0 ? PK_B2 : PK_B3);
    bf16x8 bout[NT][KB];
#pragma unroll
    for (int qc = 0; qc < Weights::NQ; ++qc) {
      const bf16x8* W = wsrc.chunk(Weights::NQ * layer + qc);
#pragma unroll
      for (int jj = 0; jj < KB / Weights::NQ; ++jj) {
        const int j = (KB / Weights::NQ) * qc + jj;
        f32x4 a0[NT], a1[NT];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float b0 = bias[32 * j + 4 * g + r], b1 = bias[32 * j + 16 + 4 * g + r];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            a0[t][r] = b0;
            a1[t][r] = b1;
          }
        }
#pragma unroll
        for (int q = 0; q < KB; ++q) {
          const bf16x8 A0 = W[((2 * jj) * KB + q) * 64 + l];
          const bf16x8 A1 = W[((2 * jj + 1) * KB + q) * 64 + l];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            a0[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, bin[t][q], a0[t], 0, 0, 0);
            a1[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, bin[t][q], a1[t], 0, 0, 0);
          }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          float v[8];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            v[r] = silu(a0[t][r]);
            v[4 + r] = silu(a1[t][r]);
          }
          bout[t][j] = pack8(v);
          pin(bout[t][j]);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int q = 0; q < KB; ++q) bin[t][q] = bout[t][q];
  }
  const bf16x8* W4 = wsrc.out();
  f32x4 out[NT];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const float b = pk32[PK_B4 + 4 * g + r];
#pragma unroll
    for (int t = 0; t < NT; ++t) out[t][r] = b;
  }
#pragma unroll
  for (int q = 0; q < KB; ++q) {
    const bf16x8 A = W4[q * 64 + l];
#pragma unroll
    for (int t = 0; t < NT; ++t)
      out[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, bin[t][q], out[t], 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int j = 0; j < OUT; ++j) y[t][j] = out[t][j];
}

// Input vector-Jacobian product dx = d(dy . MLP(x)) / dx for the 16 points of a wave, together
// with the forward output y (same lane layout as mlp_tile_forward: dy[0..3] of point l % 16 in
// every lane of its column; y valid in lane group 0, dx in every lane of the column).  The
// forward keeps the SiLU derivatives of layers 2 and 3 (64 neurons per lane) as BF16 fragments
// in k-slot order; layer 1's are recomputed in the last backward step (three FMAs a neuron).
// Backward: W4^T dy on the VALU in fp32 (each lane needs only its own 64 neurons), then W3^T and
// W2^T on the MFMA chain in the forward's layout (an output tile pair of one product is the next
// product's B fragment), and the last product, with the fp32 W1, on the VALU.
template <class Weights>
__device__ __forceinline__ void mlp_tile_input_vjp(Weights& wsrc, const float* __restrict__ pk32,
                                                   float x0, float x1, float x2, const float* dy,
                                                   float* y, float* dx) {
  using namespace mlp;
  constexpr int KB = H / 32;
  int l = threadIdx.x & 63;
  asm volatile("" : "+v"(l));
  const int g = l >> 4;
  const float* w1 = pk32 + PK_W1;
  bf16x8 bin[KB], sg2[KB], sg3[KB];
#pragma unroll
  for (int j = 0; j < KB; ++j) {
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int n = kperm(32 * j + 8 * g + e);
      float z = pk32[PK_B1 + n];
      z = fmaf(x0, w1[3 * n], z);
      z = fmaf(x1, w1[3 * n + 1], z);
      z = fmaf(x2, w1[3 * n + 2], z);
      v[e] = silu(z);
    }
    bin[j] = pack8(v);
    pin(bin[j]);
  }
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const float* bias = pk32 + (layer == 0 ? PK_B2 : PK_B3);
    bf16x8* sgo = layer == 0 ? sg2 : sg3;
    bf16x8 bout[KB];
#pragma unroll
    for (int qc = 0; qc < Weights::NQ; ++qc) {
      const bf16x8* W = wsrc.chunk(Weights::NQ * layer + qc);
#pragma unroll
      for (int jj = 0; jj < KB / Weights::NQ; ++jj) {
        const int j = (KB / Weights::NQ) * qc + jj;
        f32x4 a0, a1;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          a0[r] = bias[32 * j + 4 * g + r];
          a1[r] = bias[32 * j + 16 + 4 * g + r];
        }
#pragma unroll
        for (int q = 0; q < KB; ++q) {
          a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W[((2 * jj) * KB + q) * 64 + l], bin[q], a0,
                                                       0, 0, 0);
          a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W[((2 * jj + 1) * KB + q) * 64 + l], bin[q],
                                                       a1, 0, 0, 0);
        }
        float v[8], d[8];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          v[r] = silu(a0[r]);
          v[4 + r] = silu(a1[r]);
          d[r] = silu_grad(a0[r]);
          d[4 + r] = silu_grad(a1[r]);
        }
        bout[j] = pack8(v);
        sgo[j] = pack8(d);
        pin(bout[j]);
        pin(sgo[j]);
      }
    }
#pragma unroll
    for (int q = 0; q < KB; ++q) bin[q] = bout[q];
  }
  {  // forward output layer (rows 0..OUT-1 land in lane group 0)
    const bf16x8* W4 = wsrc.out();
    f32x4 out;
#pragma unroll
    for (int r = 0; r < 4; ++r) out[r] = pk32[PK_B4 + 4 * g + r];
#pragma unroll
    for (int q = 0; q < KB; ++q)
      out = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W4[q * 64 + l], bin[q], out, 0, 0, 0);
#pragma unroll
    for (int j = 0; j < OUT; ++j) y[j] = out[j];
  }
  // backward: dZ3 = (W4^T dy) * silu'(Z3) for this lane's 64 neurons (fp32 VALU)
  bf16x8 dcur[KB];
  {
    const float4* wt = reinterpret_cast<const float4*>(wsrc.w4t());
#pragma unroll
    for (int j = 0; j < KB; ++j) {
      float v[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float4 w = wt[kperm(32 * j + 8 * g + e)];
        float a = w.x * dy[0];
        a = fmaf(w.y, dy[1], a);
        a = fmaf(w.z, dy[2], a);
        a = fmaf(w.w, dy[3], a);
        v[e] = a * (float)sg3[j][e];
      }
      dcur[j] = pack8(v);
      pin(dcur[j]);
    }
  }
  float gx = 0.0f, gy = 0.0f, gz = 0.0f;
  // opaque copies of the inputs: the layer-1 pre-activations are recomputed below rather than
  // kept live (64 registers) from the forward
  asm volatile("" : "+v"(x0), "+v"(x1), "+v"(x2));
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    bf16x8 dnext[KB];
#pragma unroll
    for (int qc = 0; qc < Weights::NQ; ++qc) {
      const bf16x8* W = wsrc.chunkT(Weights::NQ * layer + qc);
#pragma unroll
      for (int jj = 0; jj < KB / Weights::NQ; ++jj) {
        const int j = (KB / Weights::NQ) * qc + jj;
        f32x4 a0 = {0.0f, 0.0f, 0.0f, 0.0f}, a1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int q = 0; q < KB; ++q) {
          a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W[((2 * jj) * KB + q) * 64 + l], dcur[q], a0,
                                                       0, 0, 0);
          a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W[((2 * jj + 1) * KB + q) * 64 + l], dcur[q],
                                                       a1, 0, 0, 0);
        }
        if (layer == 0) {
          float v[8];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            v[r] = a0[r] * (float)sg2[j][r];
            v[4 + r] = a1[r] * (float)sg2[j][4 + r];
          }
          dnext[j] = pack8(v);
          pin(dnext[j]);
        } else {
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int n = kperm(32 * j + 8 * g + e);
            float z = pk32[PK_B1 + n];
            z = fmaf(x0, w1[3 * n], z);
            z = fmaf(x1, w1[3 * n + 1], z);
            z = fmaf(x2, w1[3 * n + 2], z);
            const float sgr = bf2f(f2bf(silu_grad(z)));  // stored-BF16 semantics
            const float v = (e < 4 ? a0[e] : a1[e - 4]) * sgr;
            gx = fmaf(v, w1[3 * n], gx);
            gy = fmaf(v, w1[3 * n + 1], gy);
            gz = fmaf(v, w1[3 * n + 2], gz);
          }
        }
      }
    }
    if (layer == 0) {
#pragma unroll
      for (int q = 0; q < KB; ++q) dcur[q] = dnext[q];
    }
  }
  // sum over the four lane groups of the column
  gx += __shfl_xor(gx, 16, 64);
  gy += __shfl_xor(gy, 16, 64);
  gz += __shfl_xor(gz, 16, 64);
  gx += __shfl_xor(gx, 32, 64);
  gy += __shfl_xor(gy, 32, 64);
  gz += __shfl_xor(gz, 32, 64);
  dx[0] = gx;
  dx[1] = gy;
  dx[2] = gz;
}

}  // namespace ssc
