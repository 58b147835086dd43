// Leapfrog with the learned clump-force surrogate (gfx950).
//
// A block of 8 waves (two per SIMD, so one wave's MFMAs overlap the other's activations)
// integrates 128 particles for the whole grid.  The FP64 state of particle p lives in the four
// lanes p + 16 g of its wave (all four compute the same update, so the MLP's 64-lane MFMA
// operands need no exchange).  Every step:
//   kick + drift (FP64) -> MLP(x) on the matrix cores: the two 256x256 layers and the output
//   layer run as BF16 v_mfma_f32_16x16x32 with fp32 accumulation, the activations chained in
//   VGPRs and each layer's weight fragments staged once per block in LDS -> clump force =
//   g_scale * MLP
//   -> fallback: particles in flagged cells (or outside the training box) are compacted into an
//   LDS list and get the exact clump sum, one wave per particle with the clumps split over
//   lanes -> host force (FP64, analytic) -> kick.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdlib>

#include "ssc_device.h"
#include "ssc_mlp.h"
#include "ssc_mlp_tile.h"

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

// Block shape: 8 waves (two per SIMD, 256-register budget each), 16 particles per wave.
constexpr int SG_BLOCK_THREADS = 512;

// The two chained layers stream through two 64 KB LDS buffers in four chunks per step (W2 tiles
// 0-7, W2 tiles 8-15, W3 tiles 0-7, W3 tiles 8-15).  Each thread holds its share (PIECES x
// 16 B) of the next chunk in registers while the current chunk is computed, so the L2 latency
// hides behind a chunk of MFMAs.  One barrier per chunk suffices: chunk c overwrites the buffer
// that chunk c - 2 was read from, and every wave finished that read before it reached the
// barrier of chunk c - 1, which the writer has passed.
template <int SG_BLOCK>
struct LdsWeights {
  static constexpr int CH = (int)(mlp::FRAG_L / 2 / 8);  // 16-byte fragments per chunk
  static constexpr int PIECES = CH / SG_BLOCK;
  static_assert(CH % SG_BLOCK == 0, "chunk must split evenly over the block");
  mlp::bf16x8* buf;        // two chunk buffers of CH fragments
  const mlp::bf16x8* w4;   // output-layer fragments (resident)
  const mlp::bf16x8* gw;   // global fragments of W2 followed by W3
  mlp::bf16x8 pre[PIECES];
  __device__ void prefetch(int c) {
    const mlp::bf16x8* src = gw + (int64_t)c * CH;
#pragma unroll
    for (int k = 0; k < PIECES; ++k) pre[k] = src[threadIdx.x + k * SG_BLOCK];
  }
  __device__ const mlp::bf16x8* chunk(int c) {
    mlp::bf16x8* dst = buf + (c & 1) * CH;
#pragma unroll
    for (int k = 0; k < PIECES; ++k) dst[threadIdx.x + k * SG_BLOCK] = pre[k];
    __syncthreads();
    prefetch((c + 1) & 3);
    return dst;
  }
  __device__ const mlp::bf16x8* out() const { return w4; }
};


template <int SG_BLOCK>
__global__ __launch_bounds__(SG_BLOCK) void surrogate_leapfrog_kernel(
    ssc_host_potential hp, ssc_sources clumps, SurrogateDesc sd, const uint16_t* __restrict__ pk16,
    const float* __restrict__ pk32g, const uint8_t* __restrict__ flags,
    const double* __restrict__ w0, const int32_t* __restrict__ start, int64_t n, double t0,
    double dt, int nsteps, double* __restrict__ w_out, unsigned long long* __restrict__ counters) {
  using namespace mlp;
  constexpr int SG_PARTS = SG_BLOCK / 4;
  extern __shared__ __align__(16) char lds_raw[];
  __shared__ int kmin_sh;
  __shared__ int nfb;
  __shared__ double fb_buf[SG_PARTS][3];  // fallback positions, overwritten by their forces

  bf16x8* wbuf = reinterpret_cast<bf16x8*>(lds_raw);
  bf16x8* w4 = reinterpret_cast<bf16x8*>(lds_raw + FRAG_L * 2);
  float* pk32 = reinterpret_cast<float*>(lds_raw + FRAG_L * 2 + FRAG_4 * 2);

  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int pl = wv * 16 + (l & 15);
  const int64_t i = (int64_t)blockIdx.x * SG_PARTS + pl;
  const bool valid = i < n;
  const bool owner = (l < 16) && valid;

  double x = 10.0, y = 0.0, z = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
  int s = nsteps;
  if (valid) {
    const double* w = w0 + 6 * i;
    x = w[0]; y = w[1]; z = w[2]; vx = w[3]; vy = w[4]; vz = w[5];
    s = start ? start[i] : 0;
    s = s < 0 ? 0 : (s > nsteps ? nsteps : s);
  }
  if (threadIdx.x == 0) {
    kmin_sh = INT_MAX;
    nfb = 0;
  }
  // resident pieces: output-layer fragments, W1 and biases
  {
    const bf16x8* g4 = reinterpret_cast<const bf16x8*>(pk16 + PK_W4);
    for (int e = threadIdx.x; e < (int)(FRAG_4 / 8); e += SG_BLOCK) w4[e] = g4[e];
    for (int e = threadIdx.x; e < (int)PK_F32_END; e += SG_BLOCK) pk32[e] = pk32g[e];
  }
  __syncthreads();
  atomicMin(&kmin_sh, s);
  __syncthreads();
  const int kmin = kmin_sh;

  LdsWeights<SG_BLOCK> lw{wbuf, w4, reinterpret_cast<const bf16x8*>(pk16 + PK_W2), {}};
  lw.prefetch(0);
  unsigned long long nfall = 0;

  // total acceleration at (x, y, z); block-uniform call (barriers inside)
  auto force = [&](double& ax, double& ay, double& az) {
    const float u0 = (float)((x - sd.center[0]) * sd.inv_half[0]);
    const float u1 = (float)((y - sd.center[1]) * sd.inv_half[1]);
    const float u2 = (float)((z - sd.center[2]) * sd.inv_half[2]);
    float o[4];
    mlp_tile_forward(lw, pk32, u0, u1, u2, o);
    // outputs live in lane group 0; copy them to the particle's other three lanes
    const float o1 = __shfl(o[1], l & 15, 64), o2 = __shfl(o[2], l & 15, 64),
                o3 = __shfl(o[3], l & 15, 64);
    double gx = sd.g_scale * (double)o1, gy = sd.g_scale * (double)o2, gz = sd.g_scale * (double)o3;
    // fallback test: outside the training box or a flagged cell
    const double cx = (x - sd.grid_lo[0]) * sd.inv_cell[0];
    const double cy = (y - sd.grid_lo[1]) * sd.inv_cell[1];
    const double cz = (z - sd.grid_lo[2]) * sd.inv_cell[2];
    bool fb = !(cx >= 0.0 && cy >= 0.0 && cz >= 0.0 && cx < sd.grid_n[0] && cy < sd.grid_n[1] &&
                cz < sd.grid_n[2]);
    if (!fb) {
      const int64_t cell = ((int64_t)cx * sd.grid_n[1] + (int64_t)cy) * sd.grid_n[2] + (int64_t)cz;
      fb = flags[cell] != 0;
    }
    fb = fb && valid;
    if (__syncthreads_or(fb)) {
      int slot = -1;
      if (fb && l < 16) {
        slot = atomicAdd(&nfb, 1);
        fb_buf[slot][0] = x;
        fb_buf[slot][1] = y;
        fb_buf[slot][2] = z;
        ++nfall;
      }
      slot = __shfl(slot, l & 15, 64);
      __syncthreads();
      const int cnt = nfb;
      for (int f = wv; f < cnt; f += SG_BLOCK / 64) {
        const double px = fb_buf[f][0], py = fb_buf[f][1], pz = fb_buf[f][2];
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int j = l; j < clumps.n; j += 64) {
          const double rs = clumps.rs[j];
          plummer_grad(px - clumps.pos[3 * j], py - clumps.pos[3 * j + 1], pz - clumps.pos[3 * j + 2],
                       clumps.gm[j], rs * rs, sx, sy, sz);
        }
        sx = group_sum<64>(sx);
        sy = group_sum<64>(sy);
        sz = group_sum<64>(sz);
        if (l == 0) {
          fb_buf[f][0] = sx;
          fb_buf[f][1] = sy;
          fb_buf[f][2] = sz;
        }
      }
      __syncthreads();
      if (fb) {
        gx = fb_buf[slot][0];
        gy = fb_buf[slot][1];
        gz = fb_buf[slot][2];
      }
      __syncthreads();
      if (threadIdx.x == 0) nfb = 0;
    }
    double hx, hy, hz;
    host_grad(hp, x, y, z, hx, hy, hz);
    ax = -(hx + gx);
    ay = -(hy + gy);
    az = -(hz + gz);
  };

  const double h = 0.5 * dt;
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (kmin < nsteps) force(ax, ay, az);
  for (int k = kmin; k < nsteps; ++k) {
    const bool act = k >= s;
    if (act) {
      vx += h * ax; vy += h * ay; vz += h * az;
      x += dt * vx; y += dt * vy; z += dt * vz;
    }
    force(ax, ay, az);
    if (act) {
      vx += h * ax; vy += h * ay; vz += h * az;
    }
  }
  if (owner) {
    double* o = w_out + 6 * i;
    o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
  }
  if (counters && nfall) atomicAdd(counters, nfall);
}

}  // namespace ssc

using namespace ssc;

extern "C" int ssc_surrogate_leapfrog(const ssc_host_potential* host, const ssc_sources* clumps,
                                      const double* scaling, const void* packed,
                                      const uint8_t* flags, const double* w0,
                                      const int32_t* start_step, int64_t n, double t0, double dt,
                                      int32_t nsteps, double* w_out,
                                      unsigned long long* fallback_count, void* stream) {
  int rc = validate_host(host);
  if (rc) return rc;
  if (!clumps || clumps->kind != SSC_SRC_STATIC || clumps->n <= 0)
    return fail(SSC_E_ARG, "ssc_surrogate_leapfrog: needs the STATIC clump table for the fallback");
  rc = validate_sources(clumps, nsteps);
  if (rc) return rc;
  if (!scaling || !packed || !flags) return fail(SSC_E_ARG, "ssc_surrogate_leapfrog: NULL argument");
  if (n < 0 || nsteps < 0) return fail(SSC_E_ARG, "ssc_surrogate_leapfrog: negative size");
  if (n == 0) return SSC_OK;
  if (!w0 || !w_out) return fail(SSC_E_ARG, "ssc_surrogate_leapfrog: NULL state");
  SurrogateDesc sd;
  for (int k = 0; k < 3; ++k) {
    sd.center[k] = scaling[k];
    sd.inv_half[k] = 1.0 / scaling[3 + k];
    sd.grid_lo[k] = scaling[8 + k];
    sd.inv_cell[k] = 1.0 / scaling[11 + k];
    sd.grid_n[k] = (int32_t)scaling[14 + k];
    if (!(scaling[3 + k] > 0) || !(scaling[11 + k] > 0) || sd.grid_n[k] < 1)
      return fail(SSC_E_ARG, "ssc_surrogate_leapfrog: bad scaling / grid");
  }
  sd.phi_scale = scaling[6];
  sd.g_scale = scaling[7];
  sd.reserved = 0;
  const size_t lds = (size_t)(mlp::FRAG_L + mlp::FRAG_4) * 2 + (size_t)mlp::PK_F32_END * 4;
  const uint16_t* pk16 = (const uint16_t*)packed;
  const float* pk32 = (const float*)((const char*)packed + mlp::PK_F32_BASE_BYTES);
  constexpr int B = SG_BLOCK_THREADS;
  hipError_t e = hipFuncSetAttribute((const void*)surrogate_leapfrog_kernel<B>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess)
    return fail(SSC_E_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  const int64_t parts = B / 4;
  const int64_t blocks = (n + parts - 1) / parts;
  if (blocks > (int64_t)INT_MAX) return fail(SSC_E_LIMIT, "ssc_surrogate_leapfrog: too many particles");
  hipLaunchKernelGGL(surrogate_leapfrog_kernel<B>, dim3((unsigned)blocks), dim3(B), lds,
                     (hipStream_t)stream, *host, *clumps, sd, pk16, pk32, flags, w0, start_step, n,
                     t0, dt, nsteps, w_out, fallback_count);
  return check_launch("ssc_surrogate_leapfrog");
}
