// Discrete adjoint of the surrogate leapfrog on gfx950 (checkpointed reverse mode, FP64 state
// and adjoint, BF16 MFMA network forward and input-VJP).
//
// Forward: the surrogate leapfrog kernel with checkpoint rows (x, v, a) every E grid steps.
// Backward, segment by segment from the last: a block of 8 waves (two per SIMD, 256 registers
// each) takes 128 particles, each in the four lanes p + 16 g of its wave like the forward.  Starting from the segment's end checkpoint it
// walks the leapfrog backwards -- the map is explicitly invertible:
//   vh = v_{k+1} - h a_{k+1};  x_k = x_{k+1} - dt vh;  v_k = vh - h a(x_k)
// -- so positions are reconstructed in registers instead of being stored, and at every row
// applies the same adjoint recursion as the plain adjoint kernel:
//   lh = lv + dt lx;  lv = lh + F_v(k);  lx += F_x(k);  W_k = h lh + h lv (h lh at the start row)
//   lx -= (H_host(x_k) + H_clump(x_k)) W_k;  dL/dtheta_host += (da/dtheta)^T W_k
// One network call per row yields both a(x_k) (forward outputs) and the clump Hessian-vector
// product g_scale (1/half) J^T (0, W_k) (input VJP).  Rows in flagged cells or outside the
// residual grid use the exact clump sum and its Hessian, one wave per particle with the clumps
// over lanes, as in the forward.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>

#include "ssc_device.h"
#include "ssc_mlp.h"
#include "ssc_mlp_tile.h"
#include "ssc_surrogate.h"

namespace ssc {

int launch_surrogate_leapfrog(const ssc_host_potential* host, const ssc_sources* clumps,
                              const double* scaling, const void* packed, const uint8_t* flags,
                              const double* w0, const int32_t* start_step, int64_t n, double t0,
                              double dt, int32_t nsteps, double* w_out,
                              unsigned long long* fallback_count, int E, double* ckpt,
                              hipStream_t stream, const char* who);

constexpr int SA_BLOCK = 256;            // 4 waves
constexpr int SA_PARTS = SA_BLOCK / 4;   // 64 particles per block
constexpr int SA_NQ = 4;                 // weight chunks per chained layer
constexpr int SA_MAXC = 2 * SSC_MAX_COMPONENTS;

// dynamic LDS layout (bytes)
constexpr size_t SA_LDS_W = (size_t)mlp::FRAG_L * 2 * 2 / SA_NQ;       // two chunk buffers
constexpr size_t SA_LDS_W4 = SA_LDS_W;                                  // W4 fragments
constexpr size_t SA_LDS_W4T = SA_LDS_W4 + (size_t)mlp::FRAG_4 * 2;      // fp32 W4^T
constexpr size_t SA_LDS_PK32 = SA_LDS_W4T + (size_t)mlp::H * mlp::OUT * 4;
constexpr size_t SA_LDS_FB = SA_LDS_PK32 + (size_t)mlp::PK_F32_END * 4;  // [parts][6] doubles
constexpr size_t SA_LDS_BYTES = SA_LDS_FB + (size_t)SA_PARTS * 6 * 8;

// One checkpoint segment [k0, k1] for all particles (block-uniform walk).
//   ck1:  checkpoint row at k1, [n][9] = (x, v, a)
//   lam:  in/out adjoint of each particle's state (at k1 on entry, at max(k0, start) on exit;
//         lam_T is read instead when k1 == nsteps)
//   gpart: [gridDim.x][SA_MAXC] per-block host-parameter partials (accumulated)
template <int B>
__global__ __launch_bounds__(B) void surrogate_adjoint_segment_kernel(
    ssc_host_potential hp, ssc_sources clumps, SurrogateDesc sd, const uint16_t* __restrict__ pk16,
    const float* __restrict__ pk32g, const uint16_t* __restrict__ pkv16,
    const float* __restrict__ w4tg, const uint8_t* __restrict__ flags,
    const double* __restrict__ ck1, const int32_t* __restrict__ start, int64_t n, double dt,
    int nsteps, int k0, int k1, const double* __restrict__ lam_T,
    const double* __restrict__ forcing, double* __restrict__ lam, double* __restrict__ gpart,
    unsigned long long* __restrict__ counters) {
  using namespace mlp;
  constexpr int PARTS = B / 4;
  extern __shared__ __align__(16) char lds_raw[];
  __shared__ int kmin_sh;
  __shared__ int nfb;
  __shared__ double red[B / 64][SA_MAXC];
  bf16x8* wbuf = reinterpret_cast<bf16x8*>(lds_raw);
  bf16x8* w4 = reinterpret_cast<bf16x8*>(lds_raw + SA_LDS_W4);
  float* w4t = reinterpret_cast<float*>(lds_raw + SA_LDS_W4T);
  float* pk32 = reinterpret_cast<float*>(lds_raw + SA_LDS_PK32);
  double(*fbuf)[6] = reinterpret_cast<double(*)[6]>(lds_raw + SA_LDS_FB);

  const int nc = hp.n_components, nc2 = 2 * nc;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, lg = l >> 4;
  const int pl = wv * 16 + (l & 15);
  const int64_t i = (int64_t)blockIdx.x * PARTS + pl;
  const bool valid = i < n;
  const bool lead = valid && l < 16;

  int s = nsteps;
  double x = 10.0, y = 0.0, z = 0.0, vx = 0.0, vy = 0.0, vz = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
  if (valid) {
    s = start ? start[i] : 0;
    s = s < 0 ? 0 : (s > nsteps ? nsteps : s);
    const double* w = ck1 + 9 * i;
    x = w[0]; y = w[1]; z = w[2]; vx = w[3]; vy = w[4]; vz = w[5];
    ax = w[6]; ay = w[7]; az = w[8];
  }
  const int kb = max(k0, s);            // this particle's first row in the segment
  const bool act = valid && kb < k1;    // moves inside the segment
  auto F = [&](int k, int c) -> double {
    return forcing ? forcing[((int64_t)k * n + i) * 6 + c] : 0.0;
  };
  double lx[3] = {0.0, 0.0, 0.0}, lv[3] = {0.0, 0.0, 0.0};
  if (valid) {
    const double* L = (k1 == nsteps ? lam_T : lam) + 6 * i;
#pragma unroll
    for (int c = 0; c < 3; ++c) { lx[c] = L[c]; lv[c] = L[3 + c]; }
    if (k1 == nsteps) {
#pragma unroll
      for (int c = 0; c < 3; ++c) { lx[c] += F(nsteps, c); lv[c] += F(nsteps, 3 + c); }
    }
  }
  if (threadIdx.x == 0) {
    kmin_sh = INT_MAX;
    nfb = 0;
  }
  // host-parameter partials of this particle: lane group lg keeps components lg and lg + 4
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
  __syncthreads();
  atomicMin(&kmin_sh, act ? kb : INT_MAX);
  __syncthreads();
  const int kmin = kmin_sh;
  unsigned long long nfall = 0;

  if (kmin < k1) {
    // resident pieces: output-layer fragments, W4^T, W1 and biases
    {
      const bf16x8* g4 = reinterpret_cast<const bf16x8*>(pk16 + PK_W4);
      for (int e = threadIdx.x; e < (int)(FRAG_4 / 8); e += B) w4[e] = g4[e];
      for (int e = threadIdx.x; e < H * OUT; e += B) w4t[e] = w4tg[e];
      for (int e = threadIdx.x; e < (int)PK_F32_END; e += B) pk32[e] = pk32g[e];
    }
    __syncthreads();
    LdsWeights<B, true, SA_NQ> lw{wbuf, w4, w4t, reinterpret_cast<const bf16x8*>(pk16 + PK_W2),
                                  reinterpret_cast<const bf16x8*>(pkv16 + PKV_W3T), {}};
    lw.prefetch(0);
    const double h = 0.5 * dt;

    // Adjoint contribution of row x for the particles with mv set: returns the acceleration at
    // x in (nax, nay, naz) and applies lx -= H W.  Block-uniform call (barriers inside).
    auto apply = [&](bool mv, const double* W, double& nax, double& nay,
                     double& naz) __attribute__((always_inline)) {
      const float u0 = (float)((x - sd.center[0]) * sd.inv_half[0]);
      const float u1 = (float)((y - sd.center[1]) * sd.inv_half[1]);
      const float u2 = (float)((z - sd.center[2]) * sd.inv_half[2]);
      const float dy[4] = {0.0f, (float)W[0], (float)W[1], (float)W[2]};
      float o[4], dxu[3];
      mlp_tile_input_vjp(lw, pk32, u0, u1, u2, dy, o, dxu);
      const float o1 = __shfl(o[1], l & 15, 64), o2 = __shfl(o[2], l & 15, 64),
                  o3 = __shfl(o[3], l & 15, 64);
      double gcx = sd.g_scale * (double)o1, gcy = sd.g_scale * (double)o2,
             gcz = sd.g_scale * (double)o3;
      double hcx = sd.g_scale * sd.inv_half[0] * (double)dxu[0];
      double hcy = sd.g_scale * sd.inv_half[1] * (double)dxu[1];
      double hcz = sd.g_scale * sd.inv_half[2] * (double)dxu[2];
      const bool fb = mv && surrogate_fallback(sd, flags, x, y, z);
      if (__syncthreads_or(fb)) {
        int slot = -1;
        if (fb && l < 16) {
          slot = atomicAdd(&nfb, 1);
          fbuf[slot][0] = x; fbuf[slot][1] = y; fbuf[slot][2] = z;
          fbuf[slot][3] = W[0]; fbuf[slot][4] = W[1]; fbuf[slot][5] = W[2];
          ++nfall;
        }
        slot = __shfl(slot, l & 15, 64);
        __syncthreads();
        const int cnt = nfb;
        for (int f = wv; f < cnt; f += B / 64) {
          const double px = fbuf[f][0], py = fbuf[f][1], pz = fbuf[f][2];
          const double W0 = fbuf[f][3], W1 = fbuf[f][4], W2 = fbuf[f][5];
          double g0 = 0.0, g1 = 0.0, g2 = 0.0, q0 = 0.0, q1 = 0.0, q2 = 0.0;
          for (int j = l; j < clumps.n; j += 64) {
            const double rs = clumps.rs[j];
            const double dx = px - clumps.pos[3 * j], ddy = py - clumps.pos[3 * j + 1],
                         dz = pz - clumps.pos[3 * j + 2];
            const double s2 = fma(dx, dx, fma(ddy, ddy, fma(dz, dz, rs * rs)));
            const double is = rsq(s2);
            const double fq = clumps.gm[j] * is * is * is;
            const double c5 = 3.0 * fq * is * is * (dx * W0 + ddy * W1 + dz * W2);
            g0 += fq * dx; g1 += fq * ddy; g2 += fq * dz;
            q0 += fq * W0 - c5 * dx; q1 += fq * W1 - c5 * ddy; q2 += fq * W2 - c5 * dz;
          }
          g0 = group_sum<64>(g0); g1 = group_sum<64>(g1); g2 = group_sum<64>(g2);
          q0 = group_sum<64>(q0); q1 = group_sum<64>(q1); q2 = group_sum<64>(q2);
          if (l == 0) {
            fbuf[f][0] = g0; fbuf[f][1] = g1; fbuf[f][2] = g2;
            fbuf[f][3] = q0; fbuf[f][4] = q1; fbuf[f][5] = q2;
          }
        }
        __syncthreads();
        if (fb) {
          gcx = fbuf[slot][0]; gcy = fbuf[slot][1]; gcz = fbuf[slot][2];
          hcx = fbuf[slot][3]; hcy = fbuf[slot][4]; hcz = fbuf[slot][5];
        }
        __syncthreads();
        if (threadIdx.x == 0) nfb = 0;
      }
      double phi, g[3], Hh[9];
      host_full(hp, x, y, z, phi, g, Hh);
      nax = -(g[0] + gcx);
      nay = -(g[1] + gcy);
      naz = -(g[2] + gcz);
      if (mv) {
        lx[0] -= Hh[0] * W[0] + Hh[1] * W[1] + Hh[2] * W[2] + hcx;
        lx[1] -= Hh[3] * W[0] + Hh[4] * W[1] + Hh[5] * W[2] + hcy;
        lx[2] -= Hh[6] * W[0] + Hh[7] * W[1] + Hh[8] * W[2] + hcz;
        for (int c = 0; c < nc; ++c) {
          if ((c & 3) != lg) continue;
          double gc[3] = {0.0, 0.0, 0.0}, dp[3];
          comp_grad(hp.c[c], x, y, z, gc[0], gc[1], gc[2]);
          comp_grad_dp1(hp.c[c], x, y, z, dp);
          const double dm = (gc[0] * W[0] + gc[1] * W[1] + gc[2] * W[2]) / hp.c[c].p[0];
          const double dr = dp[0] * W[0] + dp[1] * W[1] + dp[2] * W[2];
          if (c < 4) { acc0 -= dm; acc1 -= dr; } else { acc2 -= dm; acc3 -= dr; }
        }
      }
    };

    // rows k1 (terminal only: W = h lv at x_T, whose acceleration is the checkpoint's), then
    // k1 - 1 down to kmin; one network call per row
    for (int k = (k1 == nsteps ? k1 : k1 - 1); k >= kmin; --k) {
      const bool term = k == nsteps;
      const bool mv = act && (term || k >= kb);
      double W[3] = {0.0, 0.0, 0.0};
      double hx = 0.0, hy = 0.0, hz = 0.0;
      if (mv && term) {
#pragma unroll
        for (int c = 0; c < 3; ++c) W[c] = h * lv[c];
      } else if (mv) {
        // reverse drift-kick: x_k from the state at k + 1
        hx = vx - h * ax; hy = vy - h * ay; hz = vz - h * az;
        x -= dt * hx; y -= dt * hy; z -= dt * hz;
        double lh[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) lh[c] = lv[c] + dt * lx[c];
#pragma unroll
        for (int c = 0; c < 3; ++c) { lv[c] = lh[c] + F(k, 3 + c); lx[c] += F(k, c); }
#pragma unroll
        for (int c = 0; c < 3; ++c) W[c] = h * lh[c] + (k > s ? h * lv[c] : 0.0);
      }
      double nax, nay, naz;
      apply(mv, W, nax, nay, naz);
      if (mv && !term) {
        ax = nax; ay = nay; az = naz;
        vx = hx - h * ax; vy = hy - h * ay; vz = hz - h * az;
      }
    }
  }
  if (lead) {
    double* L = lam + 6 * i;
#pragma unroll
    for (int c = 0; c < 3; ++c) { L[c] = lx[c]; L[3 + c] = lv[c]; }
  }
  // block reduction: over the 16 particles of each lane group, then over the waves
  acc0 = group_sum<16>(acc0);
  acc1 = group_sum<16>(acc1);
  acc2 = group_sum<16>(acc2);
  acc3 = group_sum<16>(acc3);
  if ((l & 15) == 0) {
    red[wv][2 * lg] = acc0;
    red[wv][2 * lg + 1] = acc1;
    red[wv][2 * lg + 8] = acc2;
    red[wv][2 * lg + 9] = acc3;
  }
  __syncthreads();
  if (threadIdx.x < nc2) {
    double v = 0.0;
    for (int w = 0; w < B / 64; ++w) v += red[w][threadIdx.x];
    gpart[(int64_t)blockIdx.x * SA_MAXC + threadIdx.x] += v;
  }
  if (counters && nfall) atomicAdd(counters, nfall);
}

__global__ void sa_zero_kernel(double* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}


// ------------------------------------------------------------------------------------------
// workspace: checkpoints [(nseg + 1)][n][9], adjoint state [n][6], block partials
// ------------------------------------------------------------------------------------------
struct SaWs {
  double *ckpt, *lam, *gpart;
};

static int64_t sa_nseg(int nsteps, int E) { return nsteps == 0 ? 0 : (nsteps + E - 1) / E; }

static int64_t sa_bytes(int64_t n, int nsteps, int E, SaWs* w, char* base) {
  const int64_t nseg = sa_nseg(nsteps, E);
  const int64_t nblk = (n + SA_PARTS - 1) / SA_PARTS;
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* r = base ? base + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return r;
  };
  char* c1 = take((nseg + 1) * n * 9 * 8);
  char* c2 = take(n * 6 * 8);
  char* c3 = take(nblk * SA_MAXC * 8);
  if (w) {
    w->ckpt = (double*)c1;
    w->lam = (double*)c2;
    w->gpart = (double*)c3;
  }
  return off;
}

}  // namespace ssc

using namespace ssc;

extern "C" {

int64_t ssc_surrogate_adjoint_workspace_bytes(int64_t n, int32_t nsteps, int32_t checkpoint_every) {
  if (n <= 0 || nsteps < 0 || checkpoint_every <= 0) return 0;
  return sa_bytes(n, nsteps, checkpoint_every, nullptr, nullptr);
}

int ssc_surrogate_adjoint_forward(const ssc_host_potential* host, const ssc_sources* clumps,
                                  const double* scaling, const void* packed, const uint8_t* flags,
                                  const double* w0, const int32_t* start_step, int64_t n,
                                  double t0, double dt, int32_t nsteps, int32_t checkpoint_every,
                                  double* w_out, unsigned long long* fallback_count,
                                  void* workspace, int64_t workspace_bytes, void* stream) {
  if (n <= 0 || nsteps < 0 || checkpoint_every <= 0 || !workspace)
    return fail(SSC_E_ARG, "ssc_surrogate_adjoint_forward: bad argument");
  if (workspace_bytes < sa_bytes(n, nsteps, checkpoint_every, nullptr, nullptr))
    return fail(SSC_E_ARG, "ssc_surrogate_adjoint_forward: workspace too small");
  SaWs w;
  sa_bytes(n, nsteps, checkpoint_every, &w, (char*)workspace);
  return launch_surrogate_leapfrog(host, clumps, scaling, packed, flags, w0, start_step, n, t0, dt,
                                   nsteps, w_out, fallback_count, checkpoint_every, w.ckpt,
                                   (hipStream_t)stream, "ssc_surrogate_adjoint_forward");
}

int ssc_surrogate_adjoint_backward(const ssc_host_potential* host, const ssc_sources* clumps,
                                   const double* scaling, const void* packed,
                                   const void* packed_vjp, const uint8_t* flags,
                                   const int32_t* start_step, int64_t n, double dt,
                                   int32_t nsteps, int32_t checkpoint_every, const double* lam_T,
                                   const double* forcing, double* lam_0, double* g_host,
                                   unsigned long long* fallback_count, void* workspace,
                                   int64_t workspace_bytes, void* stream) {
  const char* who = "ssc_surrogate_adjoint_backward";
  int rc = validate_host(host);
  if (rc) return rc;
  if (!clumps || clumps->kind != SSC_SRC_STATIC || clumps->n <= 0)
    return fail(SSC_E_ARG, std::string(who) + ": needs the STATIC clump table for the fallback");
  rc = validate_sources(clumps, nsteps);
  if (rc) return rc;
  if (n <= 0 || nsteps < 1 || checkpoint_every <= 0 || !lam_T || !lam_0 || !g_host || !workspace ||
      !scaling || !packed || !packed_vjp || !flags)
    return fail(SSC_E_ARG, std::string(who) + ": bad argument (needs n >= 1, nsteps >= 1)");
  if (workspace_bytes < sa_bytes(n, nsteps, checkpoint_every, nullptr, nullptr))
    return fail(SSC_E_ARG, std::string(who) + ": workspace too small");
  SurrogateDesc sd;
  rc = parse_scaling(scaling, sd, who);
  if (rc) return rc;
  SaWs w;
  sa_bytes(n, nsteps, checkpoint_every, &w, (char*)workspace);
  hipStream_t st = (hipStream_t)stream;
  const int nc = host->n_components;
  const size_t lds = SA_LDS_BYTES;
  constexpr int B = SA_BLOCK;
  hipError_t e = hipFuncSetAttribute((const void*)surrogate_adjoint_segment_kernel<B>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess)
    return fail(SSC_E_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  const int64_t nblk = (n + SA_PARTS - 1) / SA_PARTS;
  if (nblk > (int64_t)INT_MAX) return fail(SSC_E_LIMIT, std::string(who) + ": too many particles");
  const int64_t nseg = sa_nseg(nsteps, checkpoint_every);
  const uint16_t* pk16 = (const uint16_t*)packed;
  const float* pk32 = (const float*)((const char*)packed + mlp::PK_F32_BASE_BYTES);
  const uint16_t* pkv16 = (const uint16_t*)packed_vjp;
  const float* w4t = (const float*)((const char*)packed_vjp + mlp::PKV_W4T_BYTES);
  hipLaunchKernelGGL(sa_zero_kernel, dim3((unsigned)((nblk * SA_MAXC + 255) / 256)), dim3(256), 0,
                     st, w.gpart, nblk * SA_MAXC);
  for (int64_t c = nseg - 1; c >= 0; --c) {
    const int k0 = (int)(c * checkpoint_every);
    const int k1 = (int)std::min<int64_t>((c + 1) * checkpoint_every, nsteps);
    hipLaunchKernelGGL((surrogate_adjoint_segment_kernel<B>), dim3((unsigned)nblk), dim3(B), lds,
                       st, *host, *clumps, sd, pk16, pk32, pkv16, w4t, flags,
                       w.ckpt + (c + 1) * n * 9, start_step, n, dt, (int)nsteps, k0, k1, lam_T,
                       forcing, w.lam, w.gpart, fallback_count);
  }
  hipMemcpyAsync(lam_0, w.lam, n * 6 * 8, hipMemcpyDeviceToDevice, st);
  if (nc > 0)
    hipLaunchKernelGGL(column_sum_kernel<256>, dim3(2 * nc), dim3(256), 0, st, w.gpart, nblk,
                       SA_MAXC, g_host);
  return check_launch(who);
}

}  // extern "C"
