// The clump-force MLP on gfx950: BF16 MFMA training (forward, backward, Adam) and the packed,
// register-chained inference form.
//
// Network: 3 -> 256 -> 256 -> 256 -> 4, SiLU after the three hidden layers.  Master parameters
// are fp32; every matrix product runs on v_mfma_f32_32x32x16_bf16 with BF16 operands (weights
// and activations rounded to nearest-even) and fp32 accumulation.  The 3-wide first layer runs
// on the fp32 VALU (BF16-rounded weights, fp32 inputs) so the input coordinates keep fp32
// resolution.
//
// Training path: batch-major activations, generic strided GEMM C = A . B (64x64 block tile,
// four 32x32 wave tiles, split-K for the batch reductions of the weight gradients).
// Inference path: weights pre-packed in MFMA fragment order; a wave evaluates 16 points per
// tile with the activations kept in registers between layers (points on the MFMA N axis).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "ssc_device.h"
#include "ssc_mlp.h"
#include "ssc_mlp_tile.h"

namespace ssc {
namespace mlp {

// ------------------------------------------------------------------------------------------
// generic GEMM: Cz[m][n] = sum_{k in chunk z} A(m,k) B(k,n) (+ bias[n] when bias != NULL)
// A(m,k) = A[m*sAm + k*sAk], B(k,n) = B[k*sBk + n*sBn]  (BF16 bits)
// ------------------------------------------------------------------------------------------
constexpr int GT = 64;   // block tile
constexpr int GK = 32;   // k per LDS stage
constexpr int GPAD = 40; // LDS row pitch (elements): 80 B keeps 16-byte fragment reads aligned

// Operand layouts of the tile loader.  GL_K: k contiguous (stride 1 in k, row stride a multiple
// of 8, 16-byte aligned base): a thread moves 8 k's of one row with one 16-byte load and one
// 16-byte LDS store.  GL_R: rows contiguous (stride 1 in m or n, k stride a multiple of 8,
// aligned base): a thread loads 8 rows of one k with one 16-byte load and stores them across
// the 8 LDS rows.  GL_ANY: element loads at any strides.  Every layout leaves the same values
// (zeros past the edges) in LDS, so the products do not depend on the choice.
enum { GL_ANY = 0, GL_K = 1, GL_R = 2 };

// One operand's 64 x 32 stage of the tile in registers: 8 BF16 values per thread.
template <int L>
struct TileLoad {
  uint4 v;
  // X(r, k) = X[r * sr + k * sk]; rows r0 .. r0 + 63 of R, k's k0 .. k0 + 31 of [.., ke)
  __device__ void load(const uint16_t* __restrict__ X, int64_t sr, int64_t sk, int r0, int R,
                       int k0, int ke, int t) {
    uint16_t e8[8];
    if (L == GL_K) {
      const int r = r0 + (t >> 2), k = k0 + (t & 3) * 8;
      if (r < R && k + 8 <= ke) {
        v = *reinterpret_cast<const uint4*>(X + (int64_t)r * sr + k);
        return;
      }
#pragma unroll
      for (int e = 0; e < 8; ++e)
        e8[e] = (r < R && k + e < ke) ? X[(int64_t)r * sr + (k + e)] : (uint16_t)0;
    } else if (L == GL_R) {
      const int k = k0 + (t >> 3), r = r0 + (t & 7) * 8;
      if (k < ke && r + 8 <= R) {
        v = *reinterpret_cast<const uint4*>(X + (int64_t)k * sk + r);
        return;
      }
#pragma unroll
      for (int e = 0; e < 8; ++e)
        e8[e] = (k < ke && r + e < R) ? X[(int64_t)k * sk + (r + e)] : (uint16_t)0;
    } else {
      const int r = r0 + (t >> 2), k = k0 + (t & 3) * 8;
#pragma unroll
      for (int e = 0; e < 8; ++e)
        e8[e] = (r < R && k + e < ke) ? X[(int64_t)r * sr + (int64_t)(k + e) * sk] : (uint16_t)0;
    }
    memcpy(&v, e8, 16);
  }
  __device__ void store(uint16_t (*Xs)[GPAD], int t) const {
    if (L == GL_R) {
      const int kk = t >> 3, rr = (t & 7) * 8;
      uint16_t e8[8];
      memcpy(e8, &v, 16);
#pragma unroll
      for (int e = 0; e < 8; ++e) Xs[rr + e][kk] = e8[e];
    } else {
      *reinterpret_cast<uint4*>(&Xs[t >> 2][(t & 3) * 8]) = v;
    }
  }
};

// The next stage's operands are loaded into registers while the current stage's MFMAs run.
template <int LA, int LB>
__global__ __launch_bounds__(256) void gemm_kernel(const uint16_t* __restrict__ A, int64_t sAm,
                                                   int64_t sAk, const uint16_t* __restrict__ B,
                                                   int64_t sBk, int64_t sBn, float* __restrict__ C,
                                                   int64_t ldc, int64_t cz, int M, int N, int K,
                                                   int kchunk, const float* __restrict__ bias) {
  __shared__ __align__(16) uint16_t As[GT][GPAD];
  __shared__ __align__(16) uint16_t Bs[GT][GPAD];
  const int m0 = blockIdx.x * GT, n0 = blockIdx.y * GT;
  const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
  const int t = threadIdx.x, l = t & 63, w = t >> 6;
  const int wm = w & 1, wn = w >> 1;
  f32x16 acc = {};
  TileLoad<LA> la;
  TileLoad<LB> lb;
  if (kb < ke) {
    la.load(A, sAm, sAk, m0, M, kb, ke, t);
    lb.load(B, sBn, sBk, n0, N, kb, ke, t);
  }
  for (int k0 = kb; k0 < ke; k0 += GK) {
    la.store(As, t);
    lb.store(Bs, t);
    __syncthreads();
    if (k0 + GK < ke) {
      la.load(A, sAm, sAk, m0, M, k0 + GK, ke, t);
      lb.load(B, sBn, sBk, n0, N, k0 + GK, ke, t);
    }
#pragma unroll
    for (int ks = 0; ks < GK; ks += 16) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[32 * wm + (l & 31)][ks + 8 * (l >> 5)]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[32 * wn + (l & 31)][ks + 8 * (l >> 5)]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  const int col = n0 + 32 * wn + (l & 31);
  if (col >= N) return;
  const float bv = bias ? bias[col] : 0.0f;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + 32 * wm + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
    if (row < M) C[(int64_t)blockIdx.z * cz + (int64_t)row * ldc + col] = acc[r] + bv;
  }
}

__global__ void splitk_reduce_kernel(const float* __restrict__ P, int nz, int64_t mn,
                                     float* __restrict__ C) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mn) return;
  float s = 0.0f;
  for (int z = 0; z < nz; ++z) s += P[(int64_t)z * mn + i];
  C[i] = s;
}

// ------------------------------------------------------------------------------------------
// elementwise / reduction kernels
// ------------------------------------------------------------------------------------------
__global__ void round_kernel(const float* __restrict__ src, uint16_t* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = f2bf(src[i]);
}

// The three matrix-core weights W2, W3, W4 to BF16 in one launch.
__global__ void round_weights_kernel(const float* __restrict__ p, uint16_t* __restrict__ w2,
                                     uint16_t* __restrict__ w3, uint16_t* __restrict__ w4) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < H * H) {
    w2[i] = f2bf(p[OW2 + i]);
  } else if (i < 2 * H * H) {
    w3[i - H * H] = f2bf(p[OW3 + i - H * H]);
  } else if (i < 2 * H * H + OUT * H) {
    w4[i - 2 * H * H] = f2bf(p[OW4 + i - 2 * H * H]);
  }
}

// Z1[b][o] = sum_i x[b][i] bf16(W1[o][i]) + b1[o];  A1 = bf16(silu(Z1))
__global__ void layer1_fwd_kernel(const float* __restrict__ p, const float* __restrict__ x,
                                  int64_t nb, float* __restrict__ Z, uint16_t* __restrict__ Aout) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nb * H) return;
  const int64_t b = idx / H;
  const int o = (int)(idx - b * H);
  float z = p[OB1 + o];
#pragma unroll
  for (int i = 0; i < IN; ++i) z = fmaf(x[b * IN + i], bf2f(f2bf(p[OW1 + o * IN + i])), z);
  Z[idx] = z;
  Aout[idx] = f2bf(silu(z));
}

__global__ void act_fwd_kernel(const float* __restrict__ Z, uint16_t* __restrict__ Aout, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) Aout[i] = f2bf(silu(Z[i]));
}

// dZ = dA * silu'(Z): fp32 in place in dA, BF16 copy in dZb
__global__ void act_bwd_kernel(float* __restrict__ dA, const float* __restrict__ Z,
                               uint16_t* __restrict__ dZb, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = dA[i] * silu_grad(Z[i]);
  dA[i] = d;
  if (dZb) dZb[i] = f2bf(d);
}

// Weighted squared error of the 4 outputs: loss = sum_b w_b |Y_b - T_b|^2 / sum_b w_b;
// dY = 2 w_b (Y_b - T_b) / sum_b w_b (fp32 and BF16); per-block partial losses.  The weight sum
// comes as 1/sum (inv_wsum) or, when wsum_dev is set, as the sum itself in device memory.
__global__ __launch_bounds__(256) void mse_kernel(const float* __restrict__ Y,
                                                  const float* __restrict__ T,
                                                  const float* __restrict__ wt, float inv_wsum,
                                                  const float* __restrict__ wsum_dev,
                                                  int64_t nb, float* __restrict__ dY,
                                                  uint16_t* __restrict__ dYb,
                                                  float* __restrict__ partial) {
  __shared__ float red[256];
  if (wsum_dev) inv_wsum = 1.0f / *wsum_dev;
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float acc = 0.0f;
  if (b < nb) {
    const float w = wt ? wt[b] : 1.0f;
#pragma unroll
    for (int j = 0; j < OUT; ++j) {
      const float d = Y[b * OUT + j] - T[b * OUT + j];
      acc += w * d * d;
      const float g = 2.0f * w * d * inv_wsum;
      dY[b * OUT + j] = g;
      dYb[b * OUT + j] = f2bf(g);
    }
  }
  red[threadIdx.x] = acc * inv_wsum;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void sum_kernel(const float* __restrict__ partial, int n,
                                                  float* __restrict__ out) {
  __shared__ float red[256];
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += 256) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// Column sums over the batch in two deterministic passes: blocks of (64 columns x 4 row
// phases) over batch chunks of CS_ROWS rows write partials, a second kernel adds the chunks.
constexpr int CS_ROWS = 512;

__global__ __launch_bounds__(256) void colsum_partial_kernel(const float* __restrict__ D,
                                                             int64_t ld, int64_t nb, int ncol,
                                                             float* __restrict__ part) {
  __shared__ float red[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), ph = threadIdx.x >> 6;
  const int64_t b0 = (int64_t)blockIdx.y * CS_ROWS, b1 = min(nb, b0 + CS_ROWS);
  float s = 0.0f;
  if (c < ncol)
    // rows b0 + ph, b0 + ph + 4, ... added in order; the loads of 32 of them are issued first
    for (int64_t b = b0 + ph; b < b1; b += 128) {
      float v[32];
#pragma unroll
      for (int u = 0; u < 32; ++u) v[u] = b + 4 * u < b1 ? D[(b + 4 * u) * ld + c] : 0.0f;
#pragma unroll
      for (int u = 0; u < 32; ++u)
        if (b + 4 * u < b1) s += v[u];
    }
  red[ph][threadIdx.x & 63] = s;
  __syncthreads();
  if (ph == 0 && c < ncol)
    part[(int64_t)blockIdx.y * ncol + c] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// dW1[o][i] = sum_b dZ1[b][o] x[b][i] and db1[o] = sum_b dZ1[b][o], partials per batch chunk:
// part[chunk][o][4] = (dW1[o][0..2], db1[o]).
__global__ __launch_bounds__(256) void layer1_grad_partial_kernel(const float* __restrict__ dZ,
                                                                  const float* __restrict__ x,
                                                                  int64_t nb,
                                                                  float* __restrict__ part) {
  __shared__ float red[4][64][IN + 1];
  const int o = blockIdx.x * 64 + (threadIdx.x & 63), ph = threadIdx.x >> 6;
  const int64_t b0 = (int64_t)blockIdx.y * CS_ROWS, b1 = min(nb, b0 + CS_ROWS);
  float s[IN + 1] = {0.0f, 0.0f, 0.0f, 0.0f};
  // rows b0 + ph, b0 + ph + 4, ... accumulated in order; the loads of 16 of them are issued first
  for (int64_t b = b0 + ph; b < b1; b += 64) {
    float d[16], xv[16][IN];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const bool in = b + 4 * u < b1;
      d[u] = in ? dZ[(b + 4 * u) * H + o] : 0.0f;
#pragma unroll
      for (int i = 0; i < IN; ++i) xv[u][i] = in ? x[(b + 4 * u) * IN + i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      if (b + 4 * u >= b1) break;
#pragma unroll
      for (int i = 0; i < IN; ++i) s[i] = fmaf(d[u], xv[u][i], s[i]);
      s[IN] += d[u];
    }
  }
#pragma unroll
  for (int i = 0; i <= IN; ++i) red[ph][threadIdx.x & 63][i] = s[i];
  __syncthreads();
  if (ph == 0)
#pragma unroll
    for (int i = 0; i <= IN; ++i)
      part[((int64_t)blockIdx.y * H + o) * (IN + 1) + i] =
          red[0][threadIdx.x][i] + red[1][threadIdx.x][i] + red[2][threadIdx.x][i] +
          red[3][threadIdx.x][i];
}

// out[e] = sum_z part[z][e] (e < ne); with stride_out = IN + 1 split for the layer-1 case
__global__ void chunk_sum_kernel(const float* __restrict__ part, int nz, int ne,
                                 float* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  float s = 0.0f;
  for (int z = 0; z < nz; ++z) s += part[(int64_t)z * ne + e];
  out[e] = s;
}

__global__ void layer1_grad_finish(const float* __restrict__ part, int nz, float* __restrict__ dW1,
                                   float* __restrict__ db1) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= H) return;
  float s[IN + 1] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int z = 0; z < nz; ++z)
#pragma unroll
    for (int i = 0; i <= IN; ++i) s[i] += part[((int64_t)z * H + o) * (IN + 1) + i];
#pragma unroll
  for (int i = 0; i < IN; ++i) dW1[o * IN + i] = s[i];
  db1[o] = s[IN];
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2, float wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * ((mi / c1) / (sqrtf(vi / c2) + eps) + wd * p[i]);
}

// Packed inference form (layout in ssc_mlp.h).
__global__ void pack_kernel(const float* __restrict__ p, uint16_t* __restrict__ pk16,
                            float* __restrict__ pk32) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2 * FRAG_L + FRAG_4) {
    int64_t r = i;
    const float* W;
    int rows;
    if (r < FRAG_L) { W = p + OW2; rows = H; }
    else if (r < 2 * FRAG_L) { r -= FRAG_L; W = p + OW3; rows = H; }
    else { r -= 2 * FRAG_L; W = p + OW4; rows = OUT; }
    // fragment order [tile m][k-step j][lane][e] of v_mfma_f32_16x16x32_bf16's A operand
    const int e = (int)(r & 7), l = (int)((r >> 3) & 63);
    const int64_t mj = r >> 9;
    const int j = (int)(mj % (H / 32)), m = (int)(mj / (H / 32));
    const int row = 16 * m + (l & 15);
    const int col = kperm(32 * j + 8 * (l >> 4) + e);
    pk16[i] = row < rows ? f2bf(W[(int64_t)row * H + col]) : (uint16_t)0;
  }
  if (i < PK_F32_END) {
    float v;
    if (i < PK_B1) v = bf2f(f2bf(p[OW1 + i]));
    else if (i < PK_B2) v = p[OB1 + (i - PK_B1)];
    else if (i < PK_B3) v = p[OB2 + (i - PK_B2)];
    else if (i < PK_B4) v = p[OB3 + (i - PK_B3)];
    else v = (i - PK_B4) < OUT ? p[OB4 + (i - PK_B4)] : 0.0f;
    pk32[i] = v;
  }
}

// Packed transposes of the input vector-Jacobian product (layout in ssc_mlp.h).
__global__ void pack_vjp_kernel(const float* __restrict__ p, uint16_t* __restrict__ pkv,
                                float* __restrict__ w4t) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < H * OUT) {  // W4^T [n][o] = W4[o][n], BF16-rounded values in fp32
    const int n = (int)(i / OUT), o = (int)(i % OUT);
    w4t[i] = bf2f(f2bf(p[OW4 + (int64_t)o * H + n]));
  }
  if (i >= PKV_BF16_END) return;
  const int e = (int)(i & 7), l = (int)((i >> 3) & 63);
  int64_t r = i - PKV_W3T;
  const float* W = p + OW3;
  if (r >= FRAG_L) { r -= FRAG_L; W = p + OW2; }
  const int64_t mj = r >> 9;
  const int j = (int)(mj % (H / 32)), m = (int)(mj / (H / 32));
  const int row = 16 * m + (l & 15);                    // input neuron of the layer
  const int col = kperm(32 * j + 8 * (l >> 4) + e);      // output neuron of the layer
  pkv[i] = f2bf(W[(int64_t)col * H + row]);
}

}  // namespace mlp

__global__ __launch_bounds__(256) void mlp_vjp_kernel(const uint16_t* __restrict__ pk16,
                                                      const float* __restrict__ pk32,
                                                      const uint16_t* __restrict__ pkv,
                                                      const float* __restrict__ x,
                                                      const float* __restrict__ dy, int64_t nb,
                                                      float* __restrict__ dx) {
  const int64_t p = ((int64_t)blockIdx.x * 256 + threadIdx.x) / 64 * 16 + (threadIdx.x & 15);
  const bool ok = p < nb;
  const float x0 = ok ? x[3 * p] : 0.0f, x1 = ok ? x[3 * p + 1] : 0.0f, x2 = ok ? x[3 * p + 2] : 0.0f;
  float d[4], yo[4], o[3];
#pragma unroll
  for (int j = 0; j < 4; ++j) d[j] = ok ? dy[4 * p + j] : 0.0f;
  GlobalWeights gw{pk16, pkv};
  mlp_tile_input_vjp(gw, pk32, x0, x1, x2, d, yo, o);
  if (ok && (threadIdx.x & 63) < 16)
#pragma unroll
    for (int j = 0; j < 3; ++j) dx[3 * p + j] = o[j];
}

__global__ __launch_bounds__(256) void mlp_infer_kernel(const uint16_t* __restrict__ pk16,
                                                        const float* __restrict__ pk32,
                                                        const float* __restrict__ x, int64_t nb,
                                                        float* __restrict__ y) {
  const int64_t p = ((int64_t)blockIdx.x * 256 + threadIdx.x) / 64 * 16 + (threadIdx.x & 15);
  const bool ok = p < nb;
  const float x0 = ok ? x[3 * p] : 0.0f, x1 = ok ? x[3 * p + 1] : 0.0f, x2 = ok ? x[3 * p + 2] : 0.0f;
  float o[4];
  GlobalWeights gw{pk16};
  mlp_tile_forward(gw, pk32, x0, x1, x2, o);
  if (ok && (threadIdx.x & 63) < 16)
#pragma unroll
    for (int j = 0; j < 4; ++j) y[4 * p + j] = o[j];
}

// ------------------------------------------------------------------------------------------
// host orchestration
// ------------------------------------------------------------------------------------------
namespace {

using namespace mlp;

struct Workspace {
  float *Z1, *Z2, *Z3, *dA, *dY, *part, *red;
  uint16_t *A1, *A2, *A3, *dZb, *dYb, *W2b, *W3b, *W4b;
};

int64_t splitk_for(int64_t nb) {
  int64_t z = nb / 512;
  return z < 1 ? 1 : (z > 64 ? 64 : z);
}

int64_t ws_bytes(int64_t nb) {
  const int64_t nz = splitk_for(nb);
  int64_t b = 0;
  b += 3 * nb * H * 4;            // Z1..Z3
  b += nb * H * 4;                // dA / dZ (fp32)
  b += 2 * nb * OUT * 4;          // dY, (spare)
  b += (nz * H * H + ((nb + CS_ROWS - 1) / CS_ROWS) * H * (IN + 1)) * 4;  // partials
  b += ((nb + 255) / 256 + 64) * 4;  // loss partials
  b += 3 * nb * H * 2;            // A1..A3
  b += nb * H * 2;                // dZ bf16
  b += nb * OUT * 2;              // dY bf16
  b += (2 * H * H + OUT * H) * 2; // BF16 weights
  return b + 16 * 64;             // alignment slack
}

Workspace carve(void* base, int64_t nb) {
  char* p = (char*)base;
  auto take = [&](int64_t bytes) {
    char* r = p;
    p += (bytes + 63) / 64 * 64;
    return r;
  };
  const int64_t nz = splitk_for(nb);
  Workspace w;
  w.Z1 = (float*)take(nb * H * 4);
  w.Z2 = (float*)take(nb * H * 4);
  w.Z3 = (float*)take(nb * H * 4);
  w.dA = (float*)take(nb * H * 4);
  w.dY = (float*)take(2 * nb * OUT * 4);
  w.part = (float*)take((nz * H * H + ((nb + CS_ROWS - 1) / CS_ROWS) * H * (IN + 1)) * 4);
  w.red = (float*)take(((nb + 255) / 256 + 64) * 4);
  w.A1 = (uint16_t*)take(nb * H * 2);
  w.A2 = (uint16_t*)take(nb * H * 2);
  w.A3 = (uint16_t*)take(nb * H * 2);
  w.dZb = (uint16_t*)take(nb * H * 2);
  w.dYb = (uint16_t*)take(nb * OUT * 2);
  w.W2b = (uint16_t*)take(H * H * 2);
  w.W3b = (uint16_t*)take(H * H * 2);
  w.W4b = (uint16_t*)take(OUT * H * 2);
  return w;
}

inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

void colsum(const float* D, int64_t ld, int64_t nb, int ncol, float* out, float* scratch,
            hipStream_t st) {
  const int nz = (int)((nb + CS_ROWS - 1) / CS_ROWS);
  hipLaunchKernelGGL(colsum_partial_kernel, dim3((ncol + 63) / 64, nz), dim3(256), 0, st, D, ld, nb,
                     ncol, scratch);
  hipLaunchKernelGGL(chunk_sum_kernel, dim3(nblk(ncol)), dim3(256), 0, st, scratch, nz, ncol, out);
}

// Loader layout of an operand X(r, k) = X[r * sr + k * sk] (see TileLoad).
int gemm_layout(const uint16_t* X, int64_t sr, int64_t sk) {
  if ((reinterpret_cast<uintptr_t>(X) & 15) != 0) return GL_ANY;
  if (sk == 1 && sr % 8 == 0) return GL_K;
  if (sr == 1 && sk % 8 == 0) return GL_R;
  return GL_ANY;
}

template <int LA, int LB>
void gemm_launch(dim3 grid, hipStream_t st, const uint16_t* A, int64_t sAm, int64_t sAk,
                 const uint16_t* B, int64_t sBk, int64_t sBn, float* C, int64_t ldc, int64_t cz,
                 int M, int N, int K, int kchunk, const float* bias) {
  hipLaunchKernelGGL((gemm_kernel<LA, LB>), grid, dim3(256), 0, st, A, sAm, sAk, B, sBk, sBn, C,
                     ldc, cz, M, N, K, kchunk, bias);
}

template <int LA>
void gemm_launch_b(int lb, dim3 grid, hipStream_t st, const uint16_t* A, int64_t sAm, int64_t sAk,
                   const uint16_t* B, int64_t sBk, int64_t sBn, float* C, int64_t ldc, int64_t cz,
                   int M, int N, int K, int kchunk, const float* bias) {
  if (lb == GL_K)
    gemm_launch<LA, GL_K>(grid, st, A, sAm, sAk, B, sBk, sBn, C, ldc, cz, M, N, K, kchunk, bias);
  else if (lb == GL_R)
    gemm_launch<LA, GL_R>(grid, st, A, sAm, sAk, B, sBk, sBn, C, ldc, cz, M, N, K, kchunk, bias);
  else
    gemm_launch<LA, GL_ANY>(grid, st, A, sAm, sAk, B, sBk, sBn, C, ldc, cz, M, N, K, kchunk, bias);
}

void gemm_dispatch(dim3 grid, hipStream_t st, const uint16_t* A, int64_t sAm, int64_t sAk,
                   const uint16_t* B, int64_t sBk, int64_t sBn, float* C, int64_t ldc, int64_t cz,
                   int M, int N, int K, int kchunk, const float* bias) {
  const int la = gemm_layout(A, sAm, sAk), lb = gemm_layout(B, sBn, sBk);
  if (la == GL_K)
    gemm_launch_b<GL_K>(lb, grid, st, A, sAm, sAk, B, sBk, sBn, C, ldc, cz, M, N, K, kchunk, bias);
  else if (la == GL_R)
    gemm_launch_b<GL_R>(lb, grid, st, A, sAm, sAk, B, sBk, sBn, C, ldc, cz, M, N, K, kchunk, bias);
  else
    gemm_launch_b<GL_ANY>(lb, grid, st, A, sAm, sAk, B, sBk, sBn, C, ldc, cz, M, N, K, kchunk,
                          bias);
}

void gemm(const uint16_t* A, int64_t sAm, int64_t sAk, const uint16_t* B, int64_t sBk,
          int64_t sBn, float* C, int M, int N, int K, const float* bias, hipStream_t st,
          float* part = nullptr, int nz = 1) {
  if (nz <= 1 || !part) {
    gemm_dispatch(dim3(nblk(M, GT), nblk(N, GT), 1), st, A, sAm, sAk, B, sBk, sBn, C, (int64_t)N,
                  (int64_t)0, M, N, K, K, bias);
    return;
  }
  int kchunk = (K + nz - 1) / nz;
  kchunk = (kchunk + GK - 1) / GK * GK;
  nz = (K + kchunk - 1) / kchunk;
  gemm_dispatch(dim3(nblk(M, GT), nblk(N, GT), nz), st, A, sAm, sAk, B, sBk, sBn, part, (int64_t)N,
                (int64_t)M * N, M, N, K, kchunk, nullptr);
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3(nblk((int64_t)M * N)), dim3(256), 0, st, part, nz,
                     (int64_t)M * N, C);
}

void round_weights(const float* p, Workspace& w, hipStream_t st) {
  hipLaunchKernelGGL(round_weights_kernel, dim3(nblk(2 * H * H + OUT * H)), dim3(256), 0, st, p,
                     w.W2b, w.W3b, w.W4b);
}

void forward(const float* p, const float* x, int64_t nb, float* y, Workspace& w, hipStream_t st) {
  round_weights(p, w, st);
  hipLaunchKernelGGL(layer1_fwd_kernel, dim3(nblk(nb * H)), dim3(256), 0, st, p, x, nb, w.Z1, w.A1);
  gemm(w.A1, H, 1, w.W2b, 1, H, w.Z2, (int)nb, H, H, p + OB2, st);
  hipLaunchKernelGGL(act_fwd_kernel, dim3(nblk(nb * H)), dim3(256), 0, st, w.Z2, w.A2, nb * H);
  gemm(w.A2, H, 1, w.W3b, 1, H, w.Z3, (int)nb, H, H, p + OB3, st);
  hipLaunchKernelGGL(act_fwd_kernel, dim3(nblk(nb * H)), dim3(256), 0, st, w.Z3, w.A3, nb * H);
  gemm(w.A3, H, 1, w.W4b, 1, H, y, (int)nb, OUT, H, p + OB4, st);
}

// Backward from dY (fp32 in w.dY, BF16 in w.dYb) given a forward with the same workspace.
void backward(const float* x, int64_t nb, float* g, Workspace& w, hipStream_t st) {
  const int nz = (int)splitk_for(nb);
  // output layer
  gemm(w.dYb, 1, OUT, w.A3, H, 1, g + OW4, OUT, H, (int)nb, nullptr, st, w.part, nz);
  colsum(w.dY, OUT, nb, OUT, g + OB4, w.part, st);
  gemm(w.dYb, OUT, 1, w.W4b, H, 1, w.dA, (int)nb, H, OUT, nullptr, st);
  // hidden layers 3 and 2
  const float* Zs[2] = {w.Z3, w.Z2};
  const uint16_t* Ain[2] = {w.A2, w.A1};
  const uint16_t* Wb[2] = {w.W3b, w.W2b};
  const int64_t oW[2] = {OW3, OW2}, oB[2] = {OB3, OB2};
  for (int s = 0; s < 2; ++s) {
    hipLaunchKernelGGL(act_bwd_kernel, dim3(nblk(nb * H)), dim3(256), 0, st, w.dA, Zs[s], w.dZb, nb * H);
    gemm(w.dZb, 1, H, Ain[s], H, 1, g + oW[s], H, H, (int)nb, nullptr, st, w.part, nz);
    colsum(w.dA, H, nb, H, g + oB[s], w.part, st);
    gemm(w.dZb, H, 1, Wb[s], H, 1, w.dA, (int)nb, H, H, nullptr, st);
  }
  // first layer (fp32)
  hipLaunchKernelGGL(act_bwd_kernel, dim3(nblk(nb * H)), dim3(256), 0, st, w.dA, w.Z1,
                     (uint16_t*)nullptr, nb * H);
  const int nz1 = (int)((nb + CS_ROWS - 1) / CS_ROWS);
  hipLaunchKernelGGL(layer1_grad_partial_kernel, dim3(H / 64, nz1), dim3(256), 0, st, w.dA, x, nb,
                     w.part);
  hipLaunchKernelGGL(layer1_grad_finish, dim3(1), dim3(H), 0, st, w.part, nz1, g + OW1, g + OB1);
}

}  // namespace
}  // namespace ssc

using namespace ssc;
using namespace ssc::mlp;

extern "C" {

int64_t ssc_mlp_num_params(void) { return NPARAM; }

int64_t ssc_mlp_packed_bytes(void) { return PACKED_BYTES; }

int64_t ssc_mlp_workspace_bytes(int64_t batch) { return batch > 0 ? ws_bytes(batch) : 0; }

int ssc_mlp_forward(const float* params, const float* x, int64_t batch, float* y, void* workspace,
                    int64_t workspace_bytes, void* stream) {
  if (!params || !x || !y || !workspace || batch <= 0)
    return fail(SSC_E_ARG, "ssc_mlp_forward: NULL pointer or empty batch");
  if (batch > (1 << 24)) return fail(SSC_E_LIMIT, "ssc_mlp_forward: batch above 16M");
  if (workspace_bytes < ws_bytes(batch)) return fail(SSC_E_ARG, "ssc_mlp_forward: workspace too small");
  Workspace w = carve(workspace, batch);
  forward(params, x, batch, y, w, (hipStream_t)stream);
  return check_launch("ssc_mlp_forward");
}

static int loss_backward(const char* name, const float* params, const float* x,
                         const float* target, const float* weight, float weight_sum,
                         const float* weight_sum_dev, int64_t batch, float* y, float* grads,
                         float* loss, void* workspace, int64_t workspace_bytes, void* stream) {
  if (!params || !x || !target || !y || !grads || !loss || !workspace || batch <= 0)
    return fail(SSC_E_ARG, std::string(name) + ": NULL pointer or empty batch");
  if (batch > (1 << 24)) return fail(SSC_E_LIMIT, std::string(name) + ": batch above 16M");
  if (workspace_bytes < ws_bytes(batch))
    return fail(SSC_E_ARG, std::string(name) + ": workspace too small");
  if (!weight_sum_dev && !(weight_sum > 0.0f))
    return fail(SSC_E_ARG, std::string(name) + ": weight_sum must be > 0");
  hipStream_t st = (hipStream_t)stream;
  Workspace w = carve(workspace, batch);
  forward(params, x, batch, y, w, st);
  const unsigned nbk = nblk(batch);
  hipLaunchKernelGGL(mse_kernel, dim3(nbk), dim3(256), 0, st, y, target, weight,
                     weight_sum_dev ? 0.0f : 1.0f / weight_sum, weight_sum_dev, batch, w.dY,
                     w.dYb, w.red);
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, st, w.red, (int)nbk, loss);
  backward(x, batch, grads, w, st);
  return check_launch(name);
}

int ssc_mlp_loss_backward(const float* params, const float* x, const float* target,
                          const float* weight, float weight_sum, int64_t batch, float* y,
                          float* grads, float* loss, void* workspace, int64_t workspace_bytes,
                          void* stream) {
  return loss_backward("ssc_mlp_loss_backward", params, x, target, weight, weight_sum, nullptr,
                       batch, y, grads, loss, workspace, workspace_bytes, stream);
}

int ssc_mlp_loss_backward_dev(const float* params, const float* x, const float* target,
                              const float* weight, const float* weight_sum, int64_t batch,
                              float* y, float* grads, float* loss, void* workspace,
                              int64_t workspace_bytes, void* stream) {
  if (!weight_sum) return fail(SSC_E_ARG, "ssc_mlp_loss_backward_dev: NULL weight_sum");
  return loss_backward("ssc_mlp_loss_backward_dev", params, x, target, weight, 0.0f,
                       weight_sum, batch, y, grads, loss, workspace, workspace_bytes, stream);
}

int ssc_mlp_backward(const float* params, const float* x, const float* dy, int64_t batch,
                     float* grads, void* workspace, int64_t workspace_bytes, void* stream) {
  if (!params || !x || !dy || !grads || !workspace || batch <= 0)
    return fail(SSC_E_ARG, "ssc_mlp_backward: NULL pointer or empty batch");
  if (workspace_bytes < ws_bytes(batch)) return fail(SSC_E_ARG, "ssc_mlp_backward: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  Workspace w = carve(workspace, batch);
  // recompute the forward into the workspace, then copy dy into the fp32/BF16 buffers
  float* ytmp = w.dY + batch * OUT;
  forward(params, x, batch, ytmp, w, st);
  hipMemcpyAsync(w.dY, dy, batch * OUT * 4, hipMemcpyDeviceToDevice, st);
  hipLaunchKernelGGL(round_kernel, dim3(nblk(batch * OUT)), dim3(256), 0, st, dy, w.dYb, batch * OUT);
  backward(x, batch, grads, w, st);
  return check_launch("ssc_mlp_backward");
}

int ssc_mlp_adam(float* params, const float* grads, float* m, float* v, int64_t n, float lr,
                 float beta1, float beta2, float eps, int32_t step, float weight_decay,
                 void* stream) {
  if (!params || !grads || !m || !v || n <= 0 || step <= 0)
    return fail(SSC_E_ARG, "ssc_mlp_adam: bad argument");
  const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, params, grads,
                     m, v, n, lr, beta1, beta2, eps, c1, c2, weight_decay);
  return check_launch("ssc_mlp_adam");
}

int ssc_mlp_pack(const float* params, void* packed, void* stream) {
  if (!params || !packed) return fail(SSC_E_ARG, "ssc_mlp_pack: NULL pointer");
  uint16_t* pk16 = (uint16_t*)packed;
  float* pk32 = (float*)((char*)packed + PK_F32_BASE_BYTES);
  hipLaunchKernelGGL(pack_kernel, dim3(nblk(2 * FRAG_L + FRAG_4)), dim3(256), 0, (hipStream_t)stream,
                     params, pk16, pk32);
  return check_launch("ssc_mlp_pack");
}

int64_t ssc_mlp_packed_vjp_bytes(void) { return PACKED_VJP_BYTES; }

int ssc_mlp_pack_vjp(const float* params, void* packed_vjp, void* stream) {
  if (!params || !packed_vjp) return fail(SSC_E_ARG, "ssc_mlp_pack_vjp: NULL pointer");
  hipLaunchKernelGGL(pack_vjp_kernel, dim3(nblk(PKV_BF16_END)), dim3(256), 0, (hipStream_t)stream,
                     params, (uint16_t*)packed_vjp, (float*)((char*)packed_vjp + PKV_W4T_BYTES));
  return check_launch("ssc_mlp_pack_vjp");
}

int ssc_mlp_input_vjp(const void* packed, const void* packed_vjp, const float* x, const float* dy,
                      int64_t batch, float* dx, void* stream) {
  if (!packed || !packed_vjp || !x || !dy || !dx || batch < 0)
    return fail(SSC_E_ARG, "ssc_mlp_input_vjp: bad argument");
  if (batch == 0) return SSC_OK;
  const uint16_t* pk16 = (const uint16_t*)packed;
  const float* pk32 = (const float*)((const char*)packed + PK_F32_BASE_BYTES);
  const int64_t blocks = (batch + 63) / 64;  // 4 waves x 16 points
  hipLaunchKernelGGL(mlp_vjp_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     pk16, pk32, (const uint16_t*)packed_vjp, x, dy, batch, dx);
  return check_launch("ssc_mlp_input_vjp");
}

int ssc_mlp_infer(const void* packed, const float* x, int64_t batch, float* y, void* stream) {
  if (!packed || !x || !y || batch < 0) return fail(SSC_E_ARG, "ssc_mlp_infer: bad argument");
  if (batch == 0) return SSC_OK;
  const uint16_t* pk16 = (const uint16_t*)packed;
  const float* pk32 = (const float*)((const char*)packed + PK_F32_BASE_BYTES);
  const int64_t blocks = (batch + 63) / 64;  // 4 waves x 16 points
  hipLaunchKernelGGL(mlp_infer_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     pk16, pk32, x, batch, y);
  return check_launch("ssc_mlp_infer");
}

}  // extern "C"
