// Shared pieces of the surrogate integrator (ssc_surrogate.hip) and its discrete adjoint
// (ssc_surrogate_adjoint.hip): the input-scaling / residual-grid descriptor, the fallback test
// and the LDS weight-streaming policies of the MLP tiles.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "ssc_device.h"
#include "ssc_mlp.h"

namespace ssc {

struct SurrogateDesc {
  double center[3];
  double inv_half[3];
  double phi_scale;
  double g_scale;
  double grid_lo[3];
  double inv_cell[3];
  int32_t grid_n[3];
  int32_t reserved;
};

// scaling[17]: input centre[3], input half-width[3], phi scale, grad-phi scale, residual-grid
// origin[3], cell size[3], cells per axis[3].
inline int parse_scaling(const double* scaling, SurrogateDesc& sd, const char* who) {
  for (int k = 0; k < 3; ++k) {
    sd.center[k] = scaling[k];
    sd.inv_half[k] = 1.0 / scaling[3 + k];
    sd.grid_lo[k] = scaling[8 + k];
    sd.inv_cell[k] = 1.0 / scaling[11 + k];
    sd.grid_n[k] = (int32_t)scaling[14 + k];
    if (!(scaling[3 + k] > 0) || !(scaling[11 + k] > 0) || sd.grid_n[k] < 1)
      return fail(SSC_E_ARG, std::string(who) + ": bad scaling / grid");
  }
  sd.phi_scale = scaling[6];
  sd.g_scale = scaling[7];
  sd.reserved = 0;
  return SSC_OK;
}

// True where the exact clump sum is used: outside the residual grid or in a flagged cell.
__device__ __forceinline__ bool surrogate_fallback(const SurrogateDesc& sd,
                                                   const uint8_t* __restrict__ flags, double x,
                                                   double y, double z) {
  const double cx = (x - sd.grid_lo[0]) * sd.inv_cell[0];
  const double cy = (y - sd.grid_lo[1]) * sd.inv_cell[1];
  const double cz = (z - sd.grid_lo[2]) * sd.inv_cell[2];
  if (!(cx >= 0.0 && cy >= 0.0 && cz >= 0.0 && cx < sd.grid_n[0] && cy < sd.grid_n[1] &&
        cz < sd.grid_n[2]))
    return true;
  const int64_t cell = ((int64_t)cx * sd.grid_n[1] + (int64_t)cy) * sd.grid_n[2] + (int64_t)cz;
  return flags[cell] != 0;
}

// The chained 256x256 layers stream through two LDS buffers in chunks of 16 / NQ tiles
// (128 / NQ KB each).  Each thread holds its share (PIECES x 16 B) of the next chunk in
// registers while the current chunk is computed, so the L2 latency hides behind a chunk of
// MFMAs.  One barrier per chunk suffices: chunk c overwrites the buffer that chunk c - 2 was read
// from, and every wave finished that read before it reached the barrier of chunk c - 1, which the
// writer has passed (the number of chunks per call is even, so the buffer parity carries over
// between calls).
//
// VJP = false: the forward's chunks, W2 then W3 (gw = W2 | W3), 2 NQ per call.
// VJP = true: the input VJP's sequence, the same followed by W3^T then W2^T (gwt = W3^T | W2^T),
// 4 NQ per call; chunkT(c) is chunk 2 NQ + c.
template <int BLOCK, bool VJP, int NQ_ = 2>
struct LdsWeights {
  static constexpr int NQ = NQ_;
  static constexpr int CH = (int)(mlp::FRAG_L / NQ / 8);  // 16-byte fragments per chunk
  static constexpr int NCH = (VJP ? 4 : 2) * NQ;          // chunks per network call
  static constexpr int PIECES = (CH + BLOCK - 1) / BLOCK;
  static constexpr bool EVEN = CH % BLOCK == 0;           // else the last piece is partial
  mlp::bf16x8* buf;        // two chunk buffers of CH fragments
  const mlp::bf16x8* w4;   // output-layer fragments (resident)
  const float* w4t_;       // fp32 W4^T table (resident; VJP only)
  const mlp::bf16x8* gw;   // global fragments of W2 followed by W3
  const mlp::bf16x8* gwt;  // global fragments of W3^T followed by W2^T (VJP only)
  mlp::bf16x8 pre[PIECES];
  // The thread index is made opaque so the PIECES x NCH source addresses (loop-invariant in an
  // integrator's step loop) are recomputed per call instead of being hoisted and kept live.
  __device__ void prefetch(int c) {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    const mlp::bf16x8* src = c < 2 * NQ ? gw + (int64_t)c * CH : gwt + (int64_t)(c - 2 * NQ) * CH;
#pragma unroll
    for (int k = 0; k < PIECES; ++k)
      if (EVEN || k + 1 < PIECES || t + k * BLOCK < CH) pre[k] = src[t + k * BLOCK];
  }
  __device__ const mlp::bf16x8* chunk(int c) {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    mlp::bf16x8* dst = buf + (c & 1) * CH;
#pragma unroll
    for (int k = 0; k < PIECES; ++k)
      if (EVEN || k + 1 < PIECES || t + k * BLOCK < CH) dst[t + k * BLOCK] = pre[k];
    __syncthreads();
    prefetch(c + 1 == NCH ? 0 : c + 1);
    return dst;
  }
  __device__ const mlp::bf16x8* chunkT(int c) { return chunk(2 * NQ + c); }
  __device__ const mlp::bf16x8* out() const { return w4; }
  __device__ const float* w4t() const { return w4t_; }
};

}  // namespace ssc
