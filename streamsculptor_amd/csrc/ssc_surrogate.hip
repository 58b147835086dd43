// Leapfrog with the learned clump-force surrogate (gfx950).
//
// A block of 12 waves (three per SIMD, so one wave's MFMAs overlap another's activations)
// integrates 384 particles for the whole grid.  A wave holds 32 particles as two MLP point
// tiles of 16: the FP64 state of particle (tile t, column c) lives in lanes c + 32 t and
// c + 32 t + 16, and every lane of column c takes part in both tiles' MFMAs, so each weight
// fragment read from LDS feeds two MFMAs.  Every step:
//   kick + drift (FP64) -> MLP(x) on the matrix cores: the two 256x256 layers and the output
//   layer run as BF16 v_mfma_f32_16x16x32 with fp32 accumulation, the activations chained in
//   VGPRs and each layer's weight fragments streamed through LDS in chunks of 32 KB -> clump
//   force = g_scale * MLP
//   -> fallback: particles in flagged cells (or outside the training box) are compacted into an
//   LDS list and get the exact clump sum, one wave per particle with the clumps split over
//   lanes -> host force (FP64, analytic) -> kick.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdlib>

#include "ssc_device.h"
#include "ssc_mlp.h"
#include "ssc_mlp_tile.h"
#include "ssc_surrogate.h"

namespace ssc {

// Block shape: 12 waves (three per SIMD, 168-register budget each), 32 particles per wave.
constexpr int SG_BLOCK_THREADS = 512;
// Chunks per chained 256x256 layer in the LDS weight stream.
constexpr int SG_NQ = 2;

// Dynamic LDS of the integrator: two weight-chunk buffers, the output-layer fragments, the fp32
// section of the packed network (W1, biases) and the layer-1 table.
template <int SG_BLOCK>
constexpr size_t surrogate_lds_bytes() {
  return (size_t)2 * LdsWeights<SG_BLOCK, false, SG_NQ>::CH * 16 + (size_t)mlp::FRAG_4 * 2 +
         (size_t)mlp::PK_F32_END * 4 + (size_t)mlp::H * 16;
}

// One clump of the exact-sum fallback as the pair loop consumes it.
struct FbClump {
  double x, y, z, rs2, gm;
};

// CK: also write checkpoint rows ckpt[r][n][9] = (x, v, a) at grid rows min(r E, nsteps),
// r = 0..ceil(nsteps / E) (a particle that has not started holds its initial row and the
// acceleration there) for the discrete adjoint.
template <int SG_BLOCK, bool CK>
__global__ __launch_bounds__(SG_BLOCK) void surrogate_leapfrog_kernel(
    ssc_host_potential hp, ssc_sources clumps, SurrogateDesc sd, const uint16_t* __restrict__ pk16,
    const float* __restrict__ pk32g, const uint8_t* __restrict__ flags,
    const double* __restrict__ w0, const int32_t* __restrict__ start, int64_t n, double t0,
    double dt, int nsteps, double* __restrict__ w_out, unsigned long long* __restrict__ counters,
    int E, double* __restrict__ ckpt) {
  using namespace mlp;
  using LW = LdsWeights<SG_BLOCK, false, SG_NQ>;
  constexpr int SG_PARTS = SG_BLOCK / 2;
  extern __shared__ __align__(16) char lds_raw[];
  __shared__ int kmin_sh;
  __shared__ int nfb;
  __shared__ double fb_buf[SG_PARTS][3];  // fallback positions, overwritten by their forces

  bf16x8* wbuf = reinterpret_cast<bf16x8*>(lds_raw);
  bf16x8* w4 = reinterpret_cast<bf16x8*>(lds_raw + 2 * LW::CH * 16);
  float* pk32 = reinterpret_cast<float*>(lds_raw + 2 * LW::CH * 16 + FRAG_4 * 2);
  float4* l1tab = reinterpret_cast<float4*>(pk32 + PK_F32_END);  // 16-byte aligned offset

  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tl = l >> 5;  // MLP point tile of this lane's particle
  const int pl = wv * 32 + tl * 16 + (l & 15);
  const int64_t i = (int64_t)blockIdx.x * SG_PARTS + pl;
  const bool valid = i < n;
  const bool owner = ((l & 16) == 0) && valid;  // lanes l and l + 16 hold the same particle

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
  // resident pieces: output-layer fragments, W1 and biases, the layer-1 table
  {
    const bf16x8* g4 = reinterpret_cast<const bf16x8*>(pk16 + PK_W4);
    for (int e = threadIdx.x; e < (int)(FRAG_4 / 8); e += SG_BLOCK) w4[e] = g4[e];
    for (int e = threadIdx.x; e < (int)PK_F32_END; e += SG_BLOCK) pk32[e] = pk32g[e];
    build_l1_table(pk32g, l1tab, threadIdx.x, SG_BLOCK);
  }
  __syncthreads();
  atomicMin(&kmin_sh, s);
  __syncthreads();
  const int kmin = kmin_sh;

  LW lw{wbuf, w4, nullptr, reinterpret_cast<const bf16x8*>(pk16 + PK_W2), nullptr, {}};
  lw.prefetch(0);
  unsigned long long nfall = 0;

  // Exact-sum fallback from LDS: when some particle of the block falls back, thread t loads
  // clump t (at most one clump per thread) while the particles are compacted and writes it into
  // the weight-chunk buffer that the network's last chunk left free (buffer 0: chunk 2 of 4;
  // the next call's first chunk overwrites it only after the fallback's closing barrier).
  // Larger tables are read from global memory in the pair loop.
  static_assert(LW::NCH % 2 == 0, "a network call must end on chunk buffer 1");
  static_assert(SG_BLOCK * sizeof(FbClump) <= LW::CH * 16, "fallback table exceeds a buffer");
  const bool fb_lds = clumps.n <= SG_BLOCK;
  FbClump* fb_tab = reinterpret_cast<FbClump*>(wbuf);

  // total acceleration at (x, y, z); block-uniform call (barriers inside)
  auto force = [&](double& ax, double& ay, double& az) {
    const float u0 = (float)((x - sd.center[0]) * sd.inv_half[0]);
    const float u1 = (float)((y - sd.center[1]) * sd.inv_half[1]);
    const float u2 = (float)((z - sd.center[2]) * sd.inv_half[2]);
    // the other tile's particle of this column sits 32 lanes away
    const float p0 = __shfl_xor(u0, 32, 64), p1 = __shfl_xor(u1, 32, 64),
                p2 = __shfl_xor(u2, 32, 64);
    const float xin[2][3] = {{tl ? p0 : u0, tl ? p1 : u1, tl ? p2 : u2},
                             {tl ? u0 : p0, tl ? u1 : p1, tl ? u2 : p2}};
    float o[2][4];
    mlp_tiles_forward<2>(lw, pk32, l1tab, xin, o);
    // both tiles' outputs live in lane group 0: take this lane's tile from lane l % 16
    const int src = l & 15;
    const float a1 = __shfl(o[0][1], src, 64), b1 = __shfl(o[1][1], src, 64);
    const float a2 = __shfl(o[0][2], src, 64), b2 = __shfl(o[1][2], src, 64);
    const float a3 = __shfl(o[0][3], src, 64), b3 = __shfl(o[1][3], src, 64);
    double gx = sd.g_scale * (double)(tl ? b1 : a1), gy = sd.g_scale * (double)(tl ? b2 : a2),
           gz = sd.g_scale * (double)(tl ? b3 : a3);
    // fallback test: outside the training box or a flagged cell
    const bool fb = valid && surrogate_fallback(sd, flags, x, y, z);
    if (__syncthreads_or(fb)) {
      FbClump mine{};
      if (fb_lds && (int)threadIdx.x < clumps.n) {
        const int j = threadIdx.x;
        const double rs = clumps.rs[j];
        mine = FbClump{clumps.pos[3 * j], clumps.pos[3 * j + 1], clumps.pos[3 * j + 2], rs * rs,
                       clumps.gm[j]};
      }
      int slot = -1;
      if (fb && (l & 16) == 0) {
        slot = atomicAdd(&nfb, 1);
        fb_buf[slot][0] = x;
        fb_buf[slot][1] = y;
        fb_buf[slot][2] = z;
        ++nfall;
      }
      slot = __shfl(slot, l & ~16, 64);
      if (fb_lds && (int)threadIdx.x < clumps.n) fb_tab[threadIdx.x] = mine;
      __syncthreads();
      const int cnt = nfb;
      for (int f = wv; f < cnt; f += SG_BLOCK / 64) {
        const double px = fb_buf[f][0], py = fb_buf[f][1], pz = fb_buf[f][2];
        double sx = 0.0, sy = 0.0, sz = 0.0;
        if (fb_lds) {
          for (int j = l; j < clumps.n; j += 64) {
            const FbClump c = fb_tab[j];
            plummer_grad(px - c.x, py - c.y, pz - c.z, c.gm, c.rs2, sx, sy, sz);
          }
        } else {
          for (int j = l; j < clumps.n; j += 64) {
            const double rs = clumps.rs[j];
            plummer_grad(px - clumps.pos[3 * j], py - clumps.pos[3 * j + 1],
                         pz - clumps.pos[3 * j + 2], clumps.gm[j], rs * rs, sx, sy, sz);
          }
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
    host_grad<false>(hp, x, y, z, hx, hy, hz);
    ax = -(hx + gx);
    ay = -(hy + gy);
    az = -(hz + gz);
  };

  const double h = 0.5 * dt;
  double ax = 0.0, ay = 0.0, az = 0.0;
  const int nrow = CK ? (nsteps + E - 1) / E : 0;  // last checkpoint row (grid row nsteps)
  auto save = [&](int r) {
    if (owner) {
      double* o = ckpt + ((int64_t)r * n + i) * 9;
      o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
      o[6] = ax; o[7] = ay; o[8] = az;
    }
  };
  if (kmin < nsteps) force(ax, ay, az);
  if (CK)  // rows at or before the block's first step: nothing has moved yet
    for (int r = 0; r <= nrow && (r == nrow ? nsteps : r * E) <= kmin; ++r) save(r);
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
    if (CK && ((k + 1) % E == 0 || k + 1 == nsteps)) save(k + 1 == nsteps ? nrow : (k + 1) / E);
  }
  if (owner) {
    double* o = w_out + 6 * i;
    o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
  }
  if (counters && nfall) atomicAdd(counters, nfall);
}

}  // namespace ssc

using namespace ssc;

namespace ssc {

// Launch of the surrogate leapfrog (checkpoints when ckpt != NULL); arguments validated here.
int launch_surrogate_leapfrog(const ssc_host_potential* host, const ssc_sources* clumps,
                              const double* scaling, const void* packed, const uint8_t* flags,
                              const double* w0, const int32_t* start_step, int64_t n, double t0,
                              double dt, int32_t nsteps, double* w_out,
                              unsigned long long* fallback_count, int E, double* ckpt,
                              hipStream_t stream, const char* who) {
  int rc = validate_host(host);
  if (rc) return rc;
  if (!clumps || clumps->kind != SSC_SRC_STATIC || clumps->n <= 0)
    return fail(SSC_E_ARG, std::string(who) + ": needs the STATIC clump table for the fallback");
  rc = validate_sources(clumps, nsteps);
  if (rc) return rc;
  if (!scaling || !packed || !flags) return fail(SSC_E_ARG, std::string(who) + ": NULL argument");
  if (n < 0 || nsteps < 0) return fail(SSC_E_ARG, std::string(who) + ": negative size");
  if (n == 0) return SSC_OK;
  if (!w0 || !w_out) return fail(SSC_E_ARG, std::string(who) + ": NULL state");
  SurrogateDesc sd;
  rc = parse_scaling(scaling, sd, who);
  if (rc) return rc;
  constexpr int B = SG_BLOCK_THREADS;
  const size_t lds = surrogate_lds_bytes<B>();
  const uint16_t* pk16 = (const uint16_t*)packed;
  const float* pk32 = (const float*)((const char*)packed + mlp::PK_F32_BASE_BYTES);
  const void* kern = ckpt ? (const void*)surrogate_leapfrog_kernel<B, true>
                          : (const void*)surrogate_leapfrog_kernel<B, false>;
  hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess)
    return fail(SSC_E_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  const int64_t parts = B / 2;
  const int64_t blocks = (n + parts - 1) / parts;
  if (blocks > (int64_t)INT_MAX) return fail(SSC_E_LIMIT, std::string(who) + ": too many particles");
  if (ckpt)
    hipLaunchKernelGGL((surrogate_leapfrog_kernel<B, true>), dim3((unsigned)blocks), dim3(B), lds,
                       stream, *host, *clumps, sd, pk16, pk32, flags, w0, start_step, n, t0, dt,
                       nsteps, w_out, fallback_count, E, ckpt);
  else
    hipLaunchKernelGGL((surrogate_leapfrog_kernel<B, false>), dim3((unsigned)blocks), dim3(B), lds,
                       stream, *host, *clumps, sd, pk16, pk32, flags, w0, start_step, n, t0, dt,
                       nsteps, w_out, fallback_count, 1, nullptr);
  return check_launch(who);
}

}  // namespace ssc

extern "C" int ssc_surrogate_leapfrog(const ssc_host_potential* host, const ssc_sources* clumps,
                                      const double* scaling, const void* packed,
                                      const uint8_t* flags, const double* w0,
                                      const int32_t* start_step, int64_t n, double t0, double dt,
                                      int32_t nsteps, double* w_out,
                                      unsigned long long* fallback_count, void* stream) {
  return launch_surrogate_leapfrog(host, clumps, scaling, packed, flags, w0, start_step, n, t0, dt,
                                   nsteps, w_out, fallback_count, 1, nullptr,
                                   (hipStream_t)stream, "ssc_surrogate_leapfrog");
}
