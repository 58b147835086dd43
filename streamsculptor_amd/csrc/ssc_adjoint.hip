// Discrete adjoint of the fixed-step leapfrog on gfx950 (checkpointed reverse mode).
//
// Forward: ssc_leapfrog with checkpoint rows every E grid steps (FP64, stored in the caller's
// workspace).  Backward, segment by segment from the last: each lane re-runs its orbit from the
// segment's checkpoint storing the positions in HBM (E = 1: every row is a checkpoint, the
// forward's rows are read in place and nothing is re-run), then sweeps back through the segment:
//   lv_half = lv + dt lx;  lv = lv_half + F_v(k);  lx += F_x(k)
//   W_k = h lv_half + h lv (interior; only h lv_half at the start row, h lv at the last row)
//   lx -= H(x_k) W_k   (H: Hessian of host + sources),  dL/dtheta_host += (da/dtheta)^T W_k
// and records W_k.  A second kernel turns the recorded (x_k, W_k) of all particles into the
// source adjoints dL/dS_jk = sum_i H_j(x_ik - S_jk) W_ik and dL/d(Gm_j) = -sum (d.W)/s^3 with the
// sources on the lanes, so no cross-lane reduction sits in the sweep.
#include <hip/hip_runtime.h>

#include <climits>
#include <algorithm>
#include <cstring>

#include "ssc_device.h"
#include "ssc_sources_lds.h"

namespace ssc {

constexpr int AJ_BLOCK = 256;
constexpr int AJ_MAXC = 2 * SSC_MAX_COMPONENTS;
// Segment length of the backward sweep when every row is stored (checkpoint_every = 1).
constexpr int AJ_STORED_SEG = 100;

// Adjoint contribution at one grid point: lx -= H W, host-parameter gradients += (da/dp)^T W.
// With G lanes per orbit, lane r takes sources r, r + G, ... and the group sums the source part
// of H W; the host part is computed by every lane of the group.
// One host component's share: its Hessian accumulated into H, its own gradient and parameter
// derivative (from the same evaluation) into the partials gh[2c], gh[2c + 1].  K >= 0 fixes
// the component's kind at compile time.
template <int K, bool LEAN>
__device__ __forceinline__ void adj_comp(const ssc_component& cp, int c, double x, double y,
                                         double z, const double* W, double* H, double* gh) {
  double phic = 0.0, gc[3] = {0.0, 0.0, 0.0}, dp[3];
  comp_full_as<K, LEAN>(cp, x, y, z, phic, gc, H, dp);
  gh[2 * c] -= (gc[0] * W[0] + gc[1] * W[1] + gc[2] * W[2]) / cp.p[0];
  gh[2 * c + 1] -= dp[0] * W[0] + dp[1] * W[1] + dp[2] * W[2];
}

template <int SK, int G, int HS>
__device__ __forceinline__ void adj_apply(const ssc_host_potential& hp, const SourceLDS<SK>& lds,
                                          int k, double t, int rl, double x, double y, double z,
                                          const double* W, double* lx, double* gh) {
  // host Hessian accumulated over the components (as host_full), each component's own
  // gradient and parameter derivative taken from the same evaluation for its partials
  double H[9];
#pragma unroll
  for (int r = 0; r < 9; ++r) H[r] = 0.0;
  if constexpr (HS == HS_GENERIC) {
    for (int c = 0; c < hp.n_components; ++c)
      adj_comp<-1, lean_host(SK)>(hp.c[c], c, x, y, z, W, H, gh);
  } else {
    adj_comp<hs_kind(HS, 0), lean_host(SK)>(hp.c[0], 0, x, y, z, W, H, gh);
    adj_comp<hs_kind(HS, 1), lean_host(SK)>(hp.c[1], 1, x, y, z, W, H, gh);
    adj_comp<hs_kind(HS, 2), lean_host(SK)>(hp.c[2], 2, x, y, z, W, H, gh);
    if constexpr (hs_components(HS) > 3) adj_comp<hs_kind(HS, 3), lean_host(SK)>(hp.c[3], 3, x, y, z, W, H, gh);
  }
  H[3] = H[1];
  H[6] = H[2];
  H[7] = H[5];
  double hw[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) hw[r] = H[3 * r] * W[0] + H[3 * r + 1] * W[1] + H[3 * r + 2] * W[2];
  if (SK != SSC_SRC_NONE) {
    double sw[3] = {0.0, 0.0, 0.0};
    for (int j = rl; j < lds.n; j += G) {
      double sx, sy, sz, gm, q2;
      if (!lds.get(k, t, j, sx, sy, sz, gm, q2)) continue;
      // H_j W = G m (W / s^3 - 3 d (d.W) / s^5), accumulated by two FMAs per component
      const double dx = x - sx, dy = y - sy, dz = z - sz;
      const double s2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, q2)));
      const double is = rsq(s2);
      const double is2 = is * is;
      const double i3 = gm * (is2 * is);
      const double c = (i3 * is2) * (-3.0 * fma(dx, W[0], fma(dy, W[1], dz * W[2])));
      sw[0] = fma(i3, W[0], sw[0]); sw[0] = fma(c, dx, sw[0]);
      sw[1] = fma(i3, W[1], sw[1]); sw[1] = fma(c, dy, sw[1]);
      sw[2] = fma(i3, W[2], sw[2]); sw[2] = fma(c, dz, sw[2]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) hw[c] += group_sum<G>(sw[c]);
  }
  lx[0] -= hw[0];
  lx[1] -= hw[1];
  lx[2] -= hw[2];
}

// Total acceleration of one orbit whose G lanes split the source sum (all lanes of the group
// must be active: they always are, since the group shares one orbit).
template <int SK, int G, int HS>
__device__ __forceinline__ void accel_split(const ssc_host_potential& hp, const SourceLDS<SK>& lds,
                                            int k, double t, int rl, double x, double y, double z,
                                            double& ax, double& ay, double& az) {
  double sx = 0.0, sy = 0.0, sz = 0.0;
  lds.template add_grad<1>(k, t, &x, &y, &z, &sx, &sy, &sz, rl, G);
  double gx, gy, gz;
  host_grad_sig<HS, lean_host(SK)>(hp, x, y, z, gx, gy, gz);
  ax = -(gx + group_sum<G>(sx));
  ay = -(gy + group_sum<G>(sy));
  az = -(gz + group_sum<G>(sz));
}

// One checkpoint segment [k0, k1] for all particles (block-uniform grid walk), G lanes per
// particle (lane rl of the group takes sources rl, rl + G, ...; lane 0 writes the records).
//   ckpt: state at row k0 (or the initial row for particles starting later)
//   lam:  in/out adjoint of each particle's state (at k1 on entry, at max(k0, start) on exit)
//   X:    [k1 - k0 + 1][n][xs] positions of rows k0..k1; xs = 3: written here by re-running the
//         segment from ckpt; xs = 6: the stored forward rows (x, v), read only
//   Wb:   [k1 - k0 + 1][n][3] recorded weights (zero where inactive)
//   gpart: [gridDim.x][2*n_comp] per-block host-parameter partials (accumulated)
template <int SK, int G, int HS>
__global__ __launch_bounds__(AJ_BLOCK) void adjoint_segment_kernel(
    ssc_host_potential hp, ssc_sources src, const double* __restrict__ ckpt,
    const int32_t* __restrict__ start, int64_t n, double t0, double dt, int nsteps, int k0, int k1,
    const double* __restrict__ lam_T, const double* __restrict__ forcing, double* __restrict__ lam,
    double* __restrict__ X, int xs, double* __restrict__ Wb, double* __restrict__ gpart) {
  extern __shared__ __align__(16) char lds_raw[];
  __shared__ int kmin_sh;
  __shared__ double red[AJ_BLOCK / 64][AJ_MAXC];
  const int rl = threadIdx.x % G;
  const int64_t i = (int64_t)blockIdx.x * (AJ_BLOCK / G) + threadIdx.x / G;
  const bool valid = i < n, lead = valid && rl == 0;
  int s = nsteps;
  double x = 1.0, y = 0.0, z = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
  if (valid) {
    s = start ? start[i] : 0;
    s = s < 0 ? 0 : (s > nsteps ? nsteps : s);
    const double* w = ckpt + 6 * i;
    x = w[0]; y = w[1]; z = w[2]; vx = w[3]; vy = w[4]; vz = w[5];
  }
  const int kb = max(k0, s);                  // this particle's first row in the segment
  const bool act = valid && kb < k1;          // moves inside the segment
  if (threadIdx.x == 0) kmin_sh = INT_MAX;
  __syncthreads();
  atomicMin(&kmin_sh, act ? kb : INT_MAX);
  SourceLDS<SK> lds;
  lds.init(src, lds_raw);
  __syncthreads();
  const int kmin = kmin_sh;
  const int rows = k1 - k0 + 1;
  const double h = 0.5 * dt;
  // zero the records of this particle (rows it does not reach stay zero)
  if (lead)
    for (int r = 0; r < rows; ++r) {
      double* wb = Wb + ((int64_t)r * n + i) * 3;
      wb[0] = wb[1] = wb[2] = 0.0;
    }
  // host-parameter partials: 2 per component, as many as the host signature has components
  // (a fixed count keeps them in registers indexed at compile time)
  constexpr int NGH = HS == HS_GENERIC ? AJ_MAXC : 2 * hs_components(HS);
  double gh[NGH];
#pragma unroll
  for (int c = 0; c < NGH; ++c) gh[c] = 0.0;
  if (kmin < k1 && xs == 3) {
    // ---- recompute the segment, recording positions
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (act && lead) {
      double* o = X + ((int64_t)(kb - k0) * n + i) * 3;
      o[0] = x; o[1] = y; o[2] = z;
    }
    lds.stage(src, kmin, nsteps + 1);
    accel_split<SK, G, HS>(hp, lds, kmin, t0 + kmin * dt, rl, x, y, z, ax, ay, az);
    for (int k = kmin; k < k1; ++k) {
      lds.stage(src, k, nsteps + 1);
      const bool mv = act && k >= kb;
      if (mv) {
        vx += h * ax; vy += h * ay; vz += h * az;
        x += dt * vx; y += dt * vy; z += dt * vz;
      }
      accel_split<SK, G, HS>(hp, lds, k + 1, t0 + (k + 1) * dt, rl, x, y, z, ax, ay, az);
      if (mv) {
        vx += h * ax; vy += h * ay; vz += h * az;
        if (lead) {
          double* o = X + ((int64_t)(k + 1 - k0) * n + i) * 3;
          o[0] = x; o[1] = y; o[2] = z;
        }
      }
    }
  }
  if (kmin < k1) {
    // ---- reverse sweep
    double lx[3] = {0, 0, 0}, lv[3] = {0, 0, 0};
    if (valid) {
      const double* L = (k1 == nsteps ? lam_T : lam) + 6 * i;
#pragma unroll
      for (int c = 0; c < 3; ++c) { lx[c] = L[c]; lv[c] = L[3 + c]; }
    }
    auto F = [&](int k, int c) -> double {
      return forcing ? forcing[((int64_t)k * n + i) * 6 + c] : 0.0;
    };
    if (k1 == nsteps) {  // terminal row: forcing there and the closing half kick
      if (valid) {
#pragma unroll
        for (int c = 0; c < 3; ++c) { lx[c] += F(nsteps, c); lv[c] += F(nsteps, 3 + c); }
      }
      lds.stage_row(src, k1, nsteps + 1);
      if (act) {
        const double* xp = X + ((int64_t)(k1 - k0) * n + i) * xs;
        const double W[3] = {h * lv[0], h * lv[1], h * lv[2]};
        adj_apply<SK, G, HS>(hp, lds, k1, t0 + k1 * dt, rl, xp[0], xp[1], xp[2], W, lx, gh);
        if (lead) {
          double* wb = Wb + ((int64_t)(k1 - k0) * n + i) * 3;
          wb[0] = W[0]; wb[1] = W[1]; wb[2] = W[2];
        }
      }
    }
    for (int k = k1 - 1; k >= kmin; --k) {
      lds.stage_row(src, k, nsteps + 1);
      if (act && k >= kb) {
        double lh[3], W[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) lh[c] = lv[c] + dt * lx[c];
#pragma unroll
        for (int c = 0; c < 3; ++c) { lv[c] = lh[c] + F(k, 3 + c); lx[c] += F(k, c); }
#pragma unroll
        for (int c = 0; c < 3; ++c) W[c] = h * lh[c] + (k > s ? h * lv[c] : 0.0);
        const double* xp = X + ((int64_t)(k - k0) * n + i) * xs;
        adj_apply<SK, G, HS>(hp, lds, k, t0 + k * dt, rl, xp[0], xp[1], xp[2], W, lx, gh);
        if (lead) {
          double* wb = Wb + ((int64_t)(k - k0) * n + i) * 3;
          wb[0] = W[0]; wb[1] = W[1]; wb[2] = W[2];
        }
      }
    }
    if (lead) {
      double* L = lam + 6 * i;
#pragma unroll
      for (int c = 0; c < 3; ++c) { L[c] = lx[c]; L[3 + c] = lv[c]; }
    }
  } else if (lead && k1 == nsteps) {
    // nothing moves in the last segment: the adjoint of the (unmoved) row is lam_T + forcing
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const double f = forcing ? forcing[((int64_t)nsteps * n + i) * 6 + c] : 0.0;
      lam[6 * i + c] = lam_T[6 * i + c] + f;
    }
  }
  // block reduction of the host-parameter gradients (wave shuffles, then LDS); every lane of a
  // group holds the same values, so only lane 0 of each group contributes
  const int nc2 = 2 * hp.n_components;
#pragma unroll
  for (int c = 0; c < NGH; ++c) {
    if (c >= nc2) break;
    double v = rl == 0 ? gh[c] : 0.0;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < nc2) {
    double v = 0.0;
    for (int wv = 0; wv < AJ_BLOCK / 64; ++wv) v += red[wv][threadIdx.x];
    gpart[(int64_t)blockIdx.x * AJ_MAXC + threadIdx.x] += v;
  }
}

// Source adjoints of one segment: lanes = sources, one (row, particle chunk) per block.
//   part[chunk][r][j][4] = (dL/dS_jk (3), dL/d(Gm_j)) summed over the chunk's particles
constexpr int SR_CHUNK = 2048;

// The chunk's particle rows are staged 64 at a time through LDS with one coalesced load (lane l
// reads particle l of the tile), and every lane then reads each particle as an LDS broadcast:
// a wave-uniform load per particle straight from memory left the loop waiting on one round trip
// per particle.  The tile's particles with a non-zero weight (those that move at this row) form
// a wave-uniform bit mask, and the pair loop visits only its set bits.  Per pair: 3 differences,
// 3 FMAs for s^2, 1/s (hardware estimate + Newton step), s^-3 and s^-5, d.W, and 7 accumulating
// FMAs; G m multiplies the sums once at the end.
// One particle's term for one source: H_j W / (G m) and -(d.W) / s^3 added to (gx, gy, gz, gg).
__device__ __forceinline__ void pair_adjoint(const double* t, double sx, double sy, double sz,
                                             double q2, double& gx, double& gy, double& gz,
                                             double& gg) {
  const double W0 = t[3], W1 = t[4], W2 = t[5];
  const double dx = t[0] - sx, dy = t[1] - sy, dz = t[2] - sz;
  const double s2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, q2)));
  const double is = rsq(s2);
  const double is2 = is * is;
  const double i3 = is2 * is;
  const double dW = fma(dx, W0, fma(dy, W1, dz * W2));
  const double c = (i3 * is2) * (-3.0 * dW);
  gx = fma(i3, W0, gx); gx = fma(c, dx, gx);
  gy = fma(i3, W1, gy); gy = fma(c, dy, gy);
  gz = fma(i3, W2, gz); gz = fma(c, dz, gz);
  gg = fma(-i3, dW, gg);
}

template <int SK>
__global__ __launch_bounds__(64) void source_adjoint_kernel(
    ssc_sources src, const double* __restrict__ X, int xs, const double* __restrict__ Wb, int64_t n,
    int k0, int rows, double t0, double dt, double* __restrict__ part) {
  __shared__ double tile[64][6];  // x, y, z, W0, W1, W2 of 64 particles
  const int r = blockIdx.y, chunk = blockIdx.x;
  const int k = k0 + r;
  const double t = t0 + k * dt;
  const int64_t ib = (int64_t)chunk * SR_CHUNK, ie = min(n, ib + SR_CHUNK);
  const int nj = (src.n + 63) / 64;  // block-uniform trip count (barriers inside)
  for (int jj = 0; jj < nj; ++jj) {
    const int j = jj * 64 + threadIdx.x;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    const bool on = j < src.n && source_position(src, j, k, t, sx, sy, sz);
    const double q2 = j < src.n ? src.rs[j] * src.rs[j] : 1.0;  // finite pairs on idle lanes
    double gx = 0.0, gy = 0.0, gz = 0.0, gg = 0.0, hx = 0.0, hy = 0.0, hz = 0.0, hg = 0.0;
    // the next tile's row is loaded into registers while the current tile is processed
    double nx[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    auto fetch = [&](int64_t i0) {
      const int64_t i = i0 + threadIdx.x;
      if (i < ie) {
        const double* xp = X + ((int64_t)r * n + i) * xs;
        const double* w = Wb + ((int64_t)r * n + i) * 3;
        nx[0] = xp[0]; nx[1] = xp[1]; nx[2] = xp[2]; nx[3] = w[0]; nx[4] = w[1]; nx[5] = w[2];
      } else {
        nx[3] = nx[4] = nx[5] = 0.0;
      }
    };
    fetch(ib);
    for (int64_t i0 = ib; i0 < ie; i0 += 64) {
      __syncthreads();  // the previous tile has been read
#pragma unroll
      for (int c = 0; c < 6; ++c) tile[threadIdx.x][c] = nx[c];
      const uint64_t mask = __ballot(nx[3] != 0.0 || nx[4] != 0.0 || nx[5] != 0.0);
      __syncthreads();
      if (i0 + 64 < ie) fetch(i0 + 64);
      // two particles per iteration into separate sums: their dependent chains interleave
      uint64_t mk = mask;
      while (mk != 0) {
        const int p = __builtin_ctzll(mk);
        mk &= mk - 1;
        pair_adjoint(tile[p], sx, sy, sz, q2, gx, gy, gz, gg);
        if (mk != 0) {
          const int p2 = __builtin_ctzll(mk);
          mk &= mk - 1;
          pair_adjoint(tile[p2], sx, sy, sz, q2, hx, hy, hz, hg);
        }
      }
    }
    gx += hx; gy += hy; gz += hz; gg += hg;
    if (j < src.n) {
      const double gm = on ? src.gm[j] : 0.0;
      double* o = part + (((int64_t)chunk * rows + r) * src.n + j) * 4;
      o[0] = gm * gx; o[1] = gm * gy; o[2] = gm * gz; o[3] = on ? gg : 0.0;
    }
  }
}

// Sum the chunk partials of each (row, source) into row_sums[row][j][4] (one thread per
// (row, j)); TRACK sources own one position row per grid row, so their g_pos[k0 + row][j] is
// added here directly.
__global__ void source_adjoint_reduce(const double* __restrict__ part, int nchunk, int rows, int m,
                                      int k0, int track, double* __restrict__ g_pos,
                                      double* __restrict__ row_sums) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)rows * m) return;
  const int r = (int)(e / m), j = (int)(e - (int64_t)r * m);
  double a[4] = {0, 0, 0, 0};
  for (int c = 0; c < nchunk; ++c) {
    const double* p = part + (((int64_t)c * rows + r) * m + j) * 4;
    a[0] += p[0]; a[1] += p[1]; a[2] += p[2]; a[3] += p[3];
  }
  if (g_pos && track) {
    double* o = g_pos + ((int64_t)(k0 + r) * m + j) * 3;
    o[0] += a[0]; o[1] += a[1]; o[2] += a[2];
  }
  double* s = row_sums + ((int64_t)r * m + j) * 4;
  s[0] = a[0]; s[1] = a[1]; s[2] = a[2]; s[3] = a[3];
}

// Sum a segment's rows per source (one thread per source, the only writer of its outputs):
// g_gm[j] always, and g_pos[j] for STATIC sources, whose one position every row shares.
__global__ void rows_reduce(const double* __restrict__ row_sums, int rows, int m, int track,
                            double* __restrict__ g_pos, double* __restrict__ g_gm) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  double a[4] = {0, 0, 0, 0};
  for (int r = 0; r < rows; ++r) {
    const double* s = row_sums + ((int64_t)r * m + j) * 4;
    a[0] += s[0]; a[1] += s[1]; a[2] += s[2]; a[3] += s[3];
  }
  if (g_gm) g_gm[j] += a[3];
  if (g_pos && !track) {
    g_pos[3 * j] += a[0];
    g_pos[3 * j + 1] += a[1];
    g_pos[3 * j + 2] += a[2];
  }
}


__global__ void ckpt_index_kernel(int32_t* __restrict__ idx, int64_t nseg, int E, int nsteps) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c <= nseg) idx[c] = (int32_t)min((int64_t)nsteps, c * E);
}

// checkpoint c = rows[min(c E, nsteps)] of orbits saved at every grid row
__global__ void ckpt_from_rows_kernel(const double* __restrict__ rows, int64_t n, int64_t nseg,
                                      int E, int nsteps, double* __restrict__ ckpt) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = n * 6;
  if (e >= (nseg + 1) * per) return;
  const int64_t c = e / per, rem = e - c * per;
  const int64_t k = c * E < (int64_t)nsteps ? c * E : (int64_t)nsteps;
  ckpt[e] = rows[k * per + rem];
}

__global__ void zero_kernel(double* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ------------------------------------------------------------------------------------------
// workspace
// ------------------------------------------------------------------------------------------
struct AdjWs {
  double *ckpt, *lam, *X, *Wb, *gpart, *part, *gm_rows;
  int32_t* save_idx;
};

static int64_t nseg_of(int nsteps, int E) { return nsteps == 0 ? 0 : (nsteps + E - 1) / E; }

// rows per backward segment: E, or AJ_STORED_SEG when every row is stored (E = 1)
static int seg_len(int E) { return E == 1 ? AJ_STORED_SEG : E; }

static int64_t adj_bytes(int64_t n, int nsteps, int E, int m, AdjWs* w, char* base) {
  const int64_t nseg = nseg_of(nsteps, E);
  // segment-kernel blocks: G lanes per particle (the same choice adjoint_backward makes)
  const int64_t nblk = (n * group_lanes(n, m) + AJ_BLOCK - 1) / AJ_BLOCK;
  const int64_t nchunk = (n + SR_CHUNK - 1) / SR_CHUNK;
  const int64_t rows = (int64_t)seg_len(E) + 1;
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* r = base ? base + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return r;
  };
  char* c1 = take((nseg + 1) * n * 6 * 8);
  char* c2 = take(n * 6 * 8);
  char* c3 = take(E == 1 ? 0 : rows * n * 3 * 8);  // stored rows are read from the checkpoints
  char* c4 = take(rows * n * 3 * 8);
  char* c5 = take(nblk * AJ_MAXC * 8);
  char* c6 = take(nchunk * rows * (m > 0 ? m : 1) * 4 * 8);
  char* c7 = take(rows * (m > 0 ? m : 1) * 4 * 8);
  char* c8 = take((nseg + 1) * 4);
  if (w) {
    w->ckpt = (double*)c1; w->lam = (double*)c2; w->X = (double*)c3; w->Wb = (double*)c4;
    w->gpart = (double*)c5; w->part = (double*)c6; w->gm_rows = (double*)c7;
    w->save_idx = (int32_t*)c8;
  }
  return off;
}

// Segments covering rows [k_lo, k_hi), from the last (see ssc_adjoint_backward_range).
template <int SK, int HS>
static int adjoint_backward(const ssc_host_potential& hp, const ssc_sources& src,
                            const int32_t* start, int64_t n, double t0, double dt, int nsteps,
                            int E, int k_lo, int k_hi, const double* lam_T, const double* forcing,
                            double* lam_0, double* g_host, double* g_gm, double* g_pos, AdjWs& w,
                            hipStream_t st) {
  const int m = SK == SSC_SRC_NONE ? 0 : src.n;
  const bool stored = E == 1;
  const int L = seg_len(E);
  const int G = SK == SSC_SRC_NONE ? 1 : group_lanes(n, m);
  const unsigned nblk = (unsigned)((n * G + AJ_BLOCK - 1) / AJ_BLOCK);
  const int64_t nchunk = (n + SR_CHUNK - 1) / SR_CHUNK;
  const size_t lds = SourceLDS<SK>::bytes(src);
  auto seg_kernel = [&](void) -> const void* {
    switch (G) {
      case 2: return (const void*)adjoint_segment_kernel<SK, 2, HS>;
      case 4: return (const void*)adjoint_segment_kernel<SK, 4, HS>;
      case 8: return (const void*)adjoint_segment_kernel<SK, 8, HS>;
      case 16: return (const void*)adjoint_segment_kernel<SK, 16, HS>;
      default: return (const void*)adjoint_segment_kernel<SK, 1, HS>;
    }
  };
  if (lds > 60 * 1024) {
    hipError_t e = hipFuncSetAttribute(seg_kernel(), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return fail(SSC_E_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  }
  hipLaunchKernelGGL(zero_kernel, dim3((nblk * AJ_MAXC + 255) / 256), dim3(256), 0, st, w.gpart,
                     (int64_t)nblk * AJ_MAXC);
  if (g_gm && m) hipLaunchKernelGGL(zero_kernel, dim3((m + 255) / 256), dim3(256), 0, st, g_gm, (int64_t)m);
  if (g_pos && m) {
    const int64_t np = (SK == SSC_SRC_TRACK ? (int64_t)(nsteps + 1) : 1) * m * 3;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, g_pos, np);
  }
  // stored rows: segments of L rows ending at k_hi; checkpoints: the E-row segments in range
  for (int k1 = k_hi; k1 > k_lo;) {
    const int k0 = stored ? std::max(k_lo, k1 - L) : (k1 - 1) / E * E;
    const int rows = k1 - k0 + 1;
    const double* ck = w.ckpt + (int64_t)(stored ? k0 : k0 / E) * n * 6;
    double* X = stored ? (double*)ck : w.X;
    const int xs = stored ? 6 : 3;
#define SSC_SEG(GV)                                                                                \
  hipLaunchKernelGGL((adjoint_segment_kernel<SK, GV, HS>), dim3(nblk), dim3(AJ_BLOCK), lds, st, hp, src, \
                     ck, start, n, t0, dt, nsteps, k0, k1, lam_T, forcing, w.lam, X, xs, w.Wb,      \
                     w.gpart)
    switch (G) {
      case 2: SSC_SEG(2); break;
      case 4: SSC_SEG(4); break;
      case 8: SSC_SEG(8); break;
      case 16: SSC_SEG(16); break;
      default: SSC_SEG(1); break;
    }
#undef SSC_SEG
    if (m > 0 && (g_gm || g_pos)) {
      hipLaunchKernelGGL(source_adjoint_kernel<SK>, dim3((unsigned)nchunk, (unsigned)rows), dim3(64),
                         0, st, src, X, xs, w.Wb, n, k0, rows, t0, dt, w.part);
      hipLaunchKernelGGL(source_adjoint_reduce, dim3((unsigned)((rows * m + 255) / 256)), dim3(256),
                         0, st, w.part, (int)nchunk, rows, m, k0, SK == SSC_SRC_TRACK ? 1 : 0,
                         g_pos, w.gm_rows);
      hipLaunchKernelGGL(rows_reduce, dim3((m + 255) / 256), dim3(256), 0, st, w.gm_rows, rows, m,
                         SK == SSC_SRC_TRACK ? 1 : 0, g_pos, g_gm);
    }
    k1 = k0;
  }
  hipMemcpyAsync(lam_0, w.lam, n * 6 * 8, hipMemcpyDeviceToDevice, st);
  if (hp.n_components > 0)
    hipLaunchKernelGGL(column_sum_kernel<256>, dim3(2 * hp.n_components), dim3(256), 0, st,
                       w.gpart, (int64_t)nblk, AJ_MAXC, g_host);
  return check_launch("ssc_adjoint_backward");
}

}  // namespace ssc

using namespace ssc;

extern "C" {

int64_t ssc_adjoint_workspace_bytes(int64_t n, int32_t nsteps, int32_t checkpoint_every,
                                    int32_t n_src) {
  if (n <= 0 || nsteps < 0 || checkpoint_every <= 0) return 0;
  return adj_bytes(n, nsteps, checkpoint_every, n_src, nullptr, nullptr);
}

int ssc_adjoint_forward(const ssc_host_potential* host, const ssc_sources* src, const double* w0,
                        const int32_t* start_step, int64_t n, double t0, double dt, int32_t nsteps,
                        int32_t checkpoint_every, double* w_out, void* workspace,
                        int64_t workspace_bytes, void* stream) {
  if (n <= 0 || nsteps < 0 || checkpoint_every <= 0 || !w0 || !w_out || !workspace)
    return fail(SSC_E_ARG, "ssc_adjoint_forward: bad argument");
  const int m = (src && src->kind != SSC_SRC_NONE) ? src->n : 0;
  if (workspace_bytes < adj_bytes(n, nsteps, checkpoint_every, m, nullptr, nullptr))
    return fail(SSC_E_ARG, "ssc_adjoint_forward: workspace too small");
  AdjWs w;
  adj_bytes(n, nsteps, checkpoint_every, m, &w, (char*)workspace);
  const int64_t nseg = nseg_of(nsteps, checkpoint_every);
  hipLaunchKernelGGL(ckpt_index_kernel, dim3((unsigned)((nseg + 256) / 256)), dim3(256), 0,
                     (hipStream_t)stream, w.save_idx, nseg, (int)checkpoint_every, (int)nsteps);
  return ssc_leapfrog(host, src, w0, start_step, n, t0, dt, nsteps, w_out, w.save_idx,
                      (int32_t)(nseg + 1), w.ckpt, stream);
}

int ssc_adjoint_checkpoints_from_rows(const double* rows, int64_t n, int32_t nsteps,
                                      int32_t checkpoint_every, int32_t n_src, void* workspace,
                                      int64_t workspace_bytes, void* stream) {
  if (n <= 0 || nsteps < 0 || checkpoint_every <= 0 || n_src < 0 || !rows || !workspace)
    return fail(SSC_E_ARG, "ssc_adjoint_checkpoints_from_rows: bad argument");
  if (workspace_bytes < adj_bytes(n, nsteps, checkpoint_every, n_src, nullptr, nullptr))
    return fail(SSC_E_ARG, "ssc_adjoint_checkpoints_from_rows: workspace too small");
  AdjWs w;
  adj_bytes(n, nsteps, checkpoint_every, n_src, &w, (char*)workspace);
  const int64_t nseg = nseg_of(nsteps, checkpoint_every);
  const int64_t tot = (nseg + 1) * n * 6;
  hipLaunchKernelGGL(ckpt_from_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, rows, n, nseg, (int)checkpoint_every, (int)nsteps,
                     w.ckpt);
  return check_launch("ssc_adjoint_checkpoints_from_rows");
}

int ssc_adjoint_backward_range(const ssc_host_potential* host, const ssc_sources* src,
                               const int32_t* start_step, int64_t n, double t0, double dt,
                               int32_t nsteps, int32_t checkpoint_every, int32_t k_lo,
                               int32_t k_hi, const double* lam_T, const double* forcing,
                               double* lam_0, double* g_host, double* g_gm, double* g_pos,
                               void* workspace, int64_t workspace_bytes, void* stream) {
  int rc = validate_host(host);
  if (rc) return rc;
  rc = validate_sources(src, nsteps);
  if (rc) return rc;
  if (n <= 0 || nsteps < 1 || checkpoint_every <= 0 || !lam_T || !lam_0 || !g_host || !workspace)
    return fail(SSC_E_ARG, "ssc_adjoint_backward: bad argument (needs n >= 1, nsteps >= 1)");
  if (k_lo < 0 || k_hi > nsteps || k_lo >= k_hi ||
      (checkpoint_every > 1 && (k_lo % checkpoint_every != 0 ||
                                (k_hi != nsteps && k_hi % checkpoint_every != 0))))
    return fail(SSC_E_ARG, "ssc_adjoint_backward_range: needs 0 <= k_lo < k_hi <= nsteps, both "
                           "checkpoint rows (multiples of checkpoint_every, or k_hi = nsteps)");
  ssc_sources s;
  if (src) s = *src; else memset(&s, 0, sizeof(s));
  if (s.n == 0) s.kind = SSC_SRC_NONE;
  const int m = s.kind == SSC_SRC_NONE ? 0 : s.n;
  if (workspace_bytes < adj_bytes(n, nsteps, checkpoint_every, m, nullptr, nullptr))
    return fail(SSC_E_ARG, "ssc_adjoint_backward: workspace too small");
  AdjWs w;
  adj_bytes(n, nsteps, checkpoint_every, m, &w, (char*)workspace);
  hipStream_t st = (hipStream_t)stream;
  const int hs = host_signature(*host);
#define SSC_AB3(K, H)                                                                             \
  adjoint_backward<K, H>(*host, s, start_step, n, t0, dt, nsteps, checkpoint_every, k_lo, k_hi,   \
                         lam_T, forcing, lam_0, g_host, g_gm, g_pos, w, st)
#define SSC_AB(K)                                                                            \
  (hs == HS_MN_HQ_NFW ? SSC_AB3(K, HS_MN_HQ_NFW)                                             \
                      : (hs == HS_MN_HQ_HQ_NFW ? SSC_AB3(K, HS_MN_HQ_HQ_NFW) : SSC_AB3(K, HS_GENERIC)))
  switch (s.kind) {
    case SSC_SRC_NONE: return SSC_AB(SSC_SRC_NONE);
    case SSC_SRC_STATIC: return SSC_AB(SSC_SRC_STATIC);
    case SSC_SRC_TRACK: return SSC_AB(SSC_SRC_TRACK);
#undef SSC_AB
#undef SSC_AB3
    default:
      return fail(SSC_E_ARG, "ssc_adjoint_backward: LINE and Hernquist sources are not "
                             "differentiated");
  }
}

int ssc_adjoint_backward(const ssc_host_potential* host, const ssc_sources* src,
                         const int32_t* start_step, int64_t n, double t0, double dt,
                         int32_t nsteps, int32_t checkpoint_every, const double* lam_T,
                         const double* forcing, double* lam_0, double* g_host, double* g_gm,
                         double* g_pos, void* workspace, int64_t workspace_bytes, void* stream) {
  return ssc_adjoint_backward_range(host, src, start_step, n, t0, dt, nsteps, checkpoint_every, 0,
                                    nsteps, lam_T, forcing, lam_0, g_host, g_gm, g_pos, workspace,
                                    workspace_bytes, stream);
}

}  // extern "C"
