// First-order perturbation responses (the fields-interface OTF perturbation ODE) on gfx950.
//
// Per particle the state is the unperturbed row w = (x, v) in the host potential plus, for
// every perturbing source j (Plummer spheres, or Hernquist spheres for straight-line Hernquist
// subhalos), the response dw_j = d w / d eps_j at eps = 0 (mass block)
// and optionally d^2 w / d a_j d eps_j (radius block).  Discretisation: the exact derivative of
// the kick-drift-kick map of Phi_host + sum_j eps_j Phi_j, i.e. the responses follow
//   d(dx)/dt = dv,  d(dv)/dt = -grad Phi_j(x) - H(x) dx          (mass)
//   d(rx)/dt = rv,  d(rv)/dt = -grad(dPhi_j/da_j)(x) - H(x) rx    (radius)
// as in reference MassRadiusPerturbation_OTF.term (streamsculptor/fields.py:175-206), with
// dPhi_j/da_j the structural potential of SubhaloLinePotential_dRadius (potential.py:852-904)
// or SubhaloLinePotentialCustom_dRadius_fromFunc (potential.py:1215-1268).
//
// Layout: L = ceil(m / G) lanes per particle; lane g of a particle owns the G responses g,
// g + L, g + 2L, ... (strided: at every response slot the L lanes read consecutive source
// records, which keeps their LDS reads free of bank conflicts).  With a workspace the
// grid is cut into segments of rows: perturb_base_kernel (one particle per lane) advances the
// base orbits through a segment and records, at every row a particle is kicked at, its position
// and the host Hessian; perturb_rows_kernel then advances the responses through the segment
// from those records.  Without one, perturb_kernel has each of the L lanes recompute the base
// orbit and the host Hessian (2.2x the arithmetic of the split form, measured by counters).
// Velocities are carried at half steps so that forces are consumed as soon as they are formed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <string>

#include "ssc_device.h"
#include "ssc_sources_lds.h"

namespace ssc {

constexpr int PT_BLOCK = 256;

// host gradient and Hessian (upper triangle filled into all 9) at one point
template <int HS>
__device__ __forceinline__ void host_grad_hess(const ssc_host_potential& hp, double x, double y,
                                               double z, double* g, double* H) {
  double phi;
  host_full_sig<HS, true>(hp, x, y, z, phi, g, H);
}

template <int SK, int G, int RAD, int HS>
__global__ __launch_bounds__(PT_BLOCK) void perturb_kernel(
    ssc_host_potential hp, ssc_sources pert, const double* __restrict__ w0,
    const double* __restrict__ dw0, const int32_t* __restrict__ start, int64_t n, int m, int L,
    double t0, double dt, int nsteps, double* __restrict__ w_out, double* __restrict__ dw_out) {
  extern __shared__ __align__(16) char lds_raw[];
  __shared__ int kmin_sh;
  constexpr int DD = RAD ? 12 : 6;
  constexpr bool HQ = SK == SSC_SRC_LINE_HERNQUIST;
  const int ppb = PT_BLOCK / L;
  const int lp = threadIdx.x / L, grp = threadIdx.x - lp * L;
  const int64_t i = (int64_t)blockIdx.x * ppb + lp;
  const bool valid = lp < ppb && i < n;

  double x = 1.0, y = 0.0, z = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
  double dx[G][3], dv[G][3], rx[RAD ? G : 1][3], rv[RAD ? G : 1][3];
  int s = nsteps;
  if (valid) {
    const double* w = w0 + 6 * i;
    x = w[0]; y = w[1]; z = w[2]; vx = w[3]; vy = w[4]; vz = w[5];
    s = start ? start[i] : 0;
    s = s < 0 ? 0 : (s > nsteps ? nsteps : s);
  }
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int j = grp + L * q;
    const bool ok = valid && j < m && dw0;
    const double* d = ok ? dw0 + ((int64_t)i * m + j) * DD : nullptr;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      dx[q][c] = ok ? d[c] : 0.0;
      dv[q][c] = ok ? d[3 + c] : 0.0;
      if (RAD) {
        rx[q][c] = ok ? d[6 + c] : 0.0;
        rv[q][c] = ok ? d[9 + c] : 0.0;
      }
    }
  }
  if (threadIdx.x == 0) kmin_sh = INT_MAX;
  __syncthreads();
  atomicMin(&kmin_sh, s);
  SourceLDS<SK> lds;
  lds.init(pert, lds_raw);
  __syncthreads();
  const int kmin = kmin_sh;
  const double h = 0.5 * dt;

  // Forces at grid row k: base acceleration a, and for every owned response the force
  // F = -grad Phi_j - H dx (and the radius-block force), applied as vel += c * F.
  auto apply = [&](int k, double c, bool move) {
    const double t = t0 + k * dt;
    double g[3], H[9];
    host_grad_hess<HS>(hp, x, y, z, g, H);
    if (move) {
      vx -= c * g[0]; vy -= c * g[1]; vz -= c * g[2];
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int j = grp + L * q;
      double sx, sy, sz, gm, q2;
      const bool act = (j < m) && lds.get(k, t, j, sx, sy, sz, gm, q2);
      double fx = 0.0, fy = 0.0, fz = 0.0, ux = 0.0, uy = 0.0, uz = 0.0;
      if (act) {
        const double ex = x - sx, ey = y - sy, ez = z - sz;
        if (HQ) {  // Hernquist (q2 holds a): f = G m / (r (r + a)^2)
          const double r2 = fma(ex, ex, fma(ey, ey, ez * ez));
          const double ir = rsq(r2);
          const double ia = 1.0 / fma(r2, ir, q2);
          const double f = gm * ir * ia * ia;
          fx = f * ex; fy = f * ey; fz = f * ez;
          if (RAD) {  // grad of dPhi_j/da_j = -2 G m e / (r (r + a)^3)
            const double r5 = -2.0 * f * ia;
            ux = r5 * ex; uy = r5 * ey; uz = r5 * ez;
          }
        } else {  // Plummer (q2 holds a^2)
          const double s2 = fma(ex, ex, fma(ey, ey, fma(ez, ez, q2)));
          const double is = rsq(s2);
          const double f = gm * is * is * is;
          fx = f * ex; fy = f * ey; fz = f * ez;
          if (RAD) {  // grad of dPhi_j/da_j = -3 G m a e / s^5
            const double r5 = -3.0 * f * sqrt(q2) * is * is;
            ux = r5 * ex; uy = r5 * ey; uz = r5 * ez;
          }
        }
      }
      if (move) {
        const double hx = H[0] * dx[q][0] + H[1] * dx[q][1] + H[2] * dx[q][2];
        const double hy = H[3] * dx[q][0] + H[4] * dx[q][1] + H[5] * dx[q][2];
        const double hz = H[6] * dx[q][0] + H[7] * dx[q][1] + H[8] * dx[q][2];
        dv[q][0] -= c * (fx + hx);
        dv[q][1] -= c * (fy + hy);
        dv[q][2] -= c * (fz + hz);
        if (RAD) {
          const double px = H[0] * rx[q][0] + H[1] * rx[q][1] + H[2] * rx[q][2];
          const double py = H[3] * rx[q][0] + H[4] * rx[q][1] + H[5] * rx[q][2];
          const double pz = H[6] * rx[q][0] + H[7] * rx[q][1] + H[8] * rx[q][2];
          rv[q][0] -= c * (ux + px);
          rv[q][1] -= c * (uy + py);
          rv[q][2] -= c * (uz + pz);
        }
      }
    }
  };

  if (kmin < nsteps) {
    lds.stage(pert, kmin, nsteps + 1);
    apply(kmin, h, valid && s == kmin);  // first half kick of particles starting at kmin
  }
  for (int k = kmin; k < nsteps; ++k) {
    lds.stage(pert, k, nsteps + 1);
    const bool act = valid && k >= s;
    if (act) {
      x += dt * vx; y += dt * vy; z += dt * vz;
#pragma unroll
      for (int q = 0; q < G; ++q) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          dx[q][c] += dt * dv[q][c];
          if (RAD) rx[q][c] += dt * rv[q][c];
        }
      }
    }
    // active: full kick to the next half step (or the closing half kick at the last row);
    // starting at k + 1: its opening half kick
    const bool last = (k + 1 == nsteps);
    const bool starts = valid && s == k + 1 && s < nsteps;
    apply(k + 1, (act && !last) ? dt : h, act || starts);
  }
  if (valid) {
    if (grp == 0) {
      double* o = w_out + 6 * i;
      o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int j = grp + L * q;
      if (j < m) {
        double* o = dw_out + ((int64_t)i * m + j) * DD;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          o[c] = dx[q][c];
          o[3 + c] = dv[q][c];
          if (RAD) {
            o[6 + c] = rx[q][c];
            o[9 + c] = rv[q][c];
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------------------------------------
// Split form.  The kick rows of a particle starting at s < nsteps are s (opening half kick),
// s + 1 .. nsteps - 1 (a drift, then a full kick) and nsteps (a drift, then the closing half
// kick) -- the steps perturb_kernel takes.  A segment covers rows [ka, kb].  Record of row k:
// base position and the upper triangle of the host Hessian, [kb - ka + 1][n][9].
// ------------------------------------------------------------------------------------------
constexpr int PR_REC = 9;

__global__ void pt_zero_kernel(double* __restrict__ p, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = 0.0;
}

// Base orbits in place in w (initialised to w0): one particle per lane.
template <int HS>
__global__ __launch_bounds__(256) void perturb_base_kernel(
    ssc_host_potential hp, double* __restrict__ w, const int32_t* __restrict__ start, int64_t n,
    double dt, int nsteps, int ka, int kb, double* __restrict__ rows) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int s = start ? start[i] : 0;
  s = s < 0 ? 0 : (s > nsteps ? nsteps : s);
  if (s >= nsteps || s > kb) return;
  double* wi = w + 6 * i;
  double x = wi[0], y = wi[1], z = wi[2], vx = wi[3], vy = wi[4], vz = wi[5];
  const double h = 0.5 * dt;
  for (int k = ka > s ? ka : s; k <= kb; ++k) {
    if (k > s) {
      x += dt * vx; y += dt * vy; z += dt * vz;
    }
    double g[3], H[9];
    host_grad_hess<HS>(hp, x, y, z, g, H);
    double* r = rows + ((int64_t)(k - ka) * n + i) * PR_REC;
    r[0] = x; r[1] = y; r[2] = z;
    r[3] = H[0]; r[4] = H[1]; r[5] = H[2]; r[6] = H[4]; r[7] = H[5]; r[8] = H[8];
    const double c = (k == s || k == nsteps) ? h : dt;
    vx -= c * g[0]; vy -= c * g[1]; vz -= c * g[2];
  }
  wi[0] = x; wi[1] = y; wi[2] = z; wi[3] = vx; wi[4] = vy; wi[5] = vz;
}

// Responses in place in dw ([n][m][DD], initialised to dw0 or zeros) through rows [ka, kb].
template <int SK, int G, int RAD>
__global__ __launch_bounds__(PT_BLOCK) void perturb_rows_kernel(
    ssc_sources pert, const double* __restrict__ rows, const int32_t* __restrict__ start,
    int64_t n, int m, int L, double t0, double dt, int nsteps, int ka, int kb,
    double* __restrict__ dw) {
  extern __shared__ __align__(16) char lds_raw[];
  __shared__ int kmin_sh;
  constexpr int DD = RAD ? 12 : 6;
  constexpr bool HQ = SK == SSC_SRC_LINE_HERNQUIST;
  constexpr bool FOLD = (SK == SSC_SRC_STATIC || SK == SSC_SRC_TRACK) && !RAD;
  const int ppb = PT_BLOCK / L;
  const int lp = threadIdx.x / L, grp = threadIdx.x - lp * L;
  const int64_t i = (int64_t)blockIdx.x * ppb + lp;
  const bool valid = lp < ppb && i < n;
  int s = nsteps;
  if (valid) {
    s = start ? start[i] : 0;
    s = s < 0 ? 0 : (s > nsteps ? nsteps : s);
  }
  const bool moves = valid && s < nsteps && s <= kb;
  double dx[G][3], dv[G][3], rx[RAD ? G : 1][3], rv[RAD ? G : 1][3];
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int j = grp + L * q;
    const bool ok = moves && j < m;
    const double* d = ok ? dw + ((int64_t)i * m + j) * DD : nullptr;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      dx[q][c] = ok ? d[c] : 0.0;
      dv[q][c] = ok ? d[3 + c] : 0.0;
      if (RAD) {
        rx[q][c] = ok ? d[6 + c] : 0.0;
        rv[q][c] = ok ? d[9 + c] : 0.0;
      }
    }
  }
  if (threadIdx.x == 0) kmin_sh = INT_MAX;
  __syncthreads();
  if (moves) atomicMin(&kmin_sh, s);
  SourceLDS<SK> lds;
  lds.init(pert, lds_raw);
  __syncthreads();
  const int kfirst = kmin_sh < ka ? ka : kmin_sh;  // block-uniform
  const double h = 0.5 * dt;
  for (int k = kfirst; k <= kb; ++k) {
    lds.stage(pert, k, nsteps + 1);
    if (!(moves && k >= s)) continue;
    if (k > s) {
#pragma unroll
      for (int q = 0; q < G; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          dx[q][c] += dt * dv[q][c];
          if (RAD) rx[q][c] += dt * rv[q][c];
        }
    }
    const double* r = rows + ((int64_t)(k - ka) * n + i) * PR_REC;
    const double x = r[0], y = r[1], z = r[2];
    const double H[9] = {r[3], r[4], r[5], r[4], r[6], r[7], r[5], r[7], r[8]};
    const double t = t0 + k * dt;
    const double c = (k == s || k == nsteps) ? h : dt;
    if constexpr (FOLD) {
      // Plummer sources that are always active: no branch per response, the folded
      // reciprocal-square-root correction of plummer_grad_k, and the source force added into
      // the Hessian product by one FMA per component
#pragma unroll
      for (int q = 0; q < G; ++q) {
        const int j = grp + L * q;
        double sx, sy, sz, q2, c1, c2;
        lds.get_k(k, j < m ? j : m - 1, sx, sy, sz, q2, c1, c2);
        const double ex = x - sx, ey = y - sy, ez = z - sz;
        const double s2 = fma(ex, ex, fma(ey, ey, fma(ez, ez, q2)));
        const double yr = __builtin_amdgcn_rsq(s2);
        const double y2 = yr * yr;
        double f = yr * y2 * fma(s2 * y2, c1, c2);
        f = j < m ? f : 0.0;
        const double hx = fma(H[2], dx[q][2], fma(H[1], dx[q][1], H[0] * dx[q][0]));
        const double hy = fma(H[5], dx[q][2], fma(H[4], dx[q][1], H[3] * dx[q][0]));
        const double hz = fma(H[8], dx[q][2], fma(H[7], dx[q][1], H[6] * dx[q][0]));
        dv[q][0] = fma(-c, fma(f, ex, hx), dv[q][0]);
        dv[q][1] = fma(-c, fma(f, ey, hy), dv[q][1]);
        dv[q][2] = fma(-c, fma(f, ez, hz), dv[q][2]);
      }
      continue;
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int j = grp + L * q;
      double sx, sy, sz, gm, q2;
      const bool act = (j < m) && lds.get(k, t, j, sx, sy, sz, gm, q2);
      double fx = 0.0, fy = 0.0, fz = 0.0, ux = 0.0, uy = 0.0, uz = 0.0;
      if (act) {
        const double ex = x - sx, ey = y - sy, ez = z - sz;
        if (HQ) {  // Hernquist (q2 holds a): f = G m / (r (r + a)^2)
          const double r2 = fma(ex, ex, fma(ey, ey, ez * ez));
          const double ir = rsq(r2);
          const double ia = 1.0 / fma(r2, ir, q2);
          const double f = gm * ir * ia * ia;
          fx = f * ex; fy = f * ey; fz = f * ez;
          if (RAD) {  // grad of dPhi_j/da_j = -2 G m e / (r (r + a)^3)
            const double r5 = -2.0 * f * ia;
            ux = r5 * ex; uy = r5 * ey; uz = r5 * ez;
          }
        } else {  // Plummer (q2 holds a^2)
          const double s2 = fma(ex, ex, fma(ey, ey, fma(ez, ez, q2)));
          const double is = rsq(s2);
          const double f = gm * is * is * is;
          fx = f * ex; fy = f * ey; fz = f * ez;
          if (RAD) {  // grad of dPhi_j/da_j = -3 G m a e / s^5
            const double r5 = -3.0 * f * sqrt(q2) * is * is;
            ux = r5 * ex; uy = r5 * ey; uz = r5 * ez;
          }
        }
      }
      const double hx = H[0] * dx[q][0] + H[1] * dx[q][1] + H[2] * dx[q][2];
      const double hy = H[3] * dx[q][0] + H[4] * dx[q][1] + H[5] * dx[q][2];
      const double hz = H[6] * dx[q][0] + H[7] * dx[q][1] + H[8] * dx[q][2];
      dv[q][0] -= c * (fx + hx);
      dv[q][1] -= c * (fy + hy);
      dv[q][2] -= c * (fz + hz);
      if (RAD) {
        const double px = H[0] * rx[q][0] + H[1] * rx[q][1] + H[2] * rx[q][2];
        const double py = H[3] * rx[q][0] + H[4] * rx[q][1] + H[5] * rx[q][2];
        const double pz = H[6] * rx[q][0] + H[7] * rx[q][1] + H[8] * rx[q][2];
        rv[q][0] -= c * (ux + px);
        rv[q][1] -= c * (uy + py);
        rv[q][2] -= c * (uz + pz);
      }
    }
  }
  if (moves) {
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int j = grp + L * q;
      if (j < m) {
        double* o = dw + ((int64_t)i * m + j) * DD;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          o[c] = dx[q][c];
          o[3 + c] = dv[q][c];
          if (RAD) {
            o[6 + c] = rx[q][c];
            o[9 + c] = rv[q][c];
          }
        }
      }
    }
  }
}

// Rows per segment that fit in workspace_bytes (0 when not even one does).
static int64_t perturb_segment_rows(int64_t n, int nsteps, int64_t workspace_bytes) {
  const int64_t per_row = n * PR_REC * 8;
  const int64_t e = per_row > 0 ? workspace_bytes / per_row : 0;
  return e < (int64_t)nsteps + 1 ? e : (int64_t)nsteps + 1;
}

template <int SK, int G, int RAD, int HS>
static int launch_perturb_split(const ssc_host_potential& hp, const ssc_sources& pert,
                                const double* w0, const double* dw0, const int32_t* start,
                                int64_t n, double t0, double dt, int nsteps, double* w_out,
                                double* dw_out, double* rows, int64_t E, hipStream_t st) {
  const int m = pert.n;
  constexpr int DD = RAD ? 12 : 6;
  int L = (m + G - 1) / G;
  int Lp = 1;
  while (Lp < L) Lp <<= 1;
  if (Lp > PT_BLOCK) return fail(SSC_E_LIMIT, "ssc_perturb_leapfrog: too many sources");
  const size_t lds = SourceLDS<SK>::bytes(pert);
  if (lds > 60 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)perturb_rows_kernel<SK, G, RAD>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return fail(SSC_E_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  }
  hipMemcpyAsync(w_out, w0, (size_t)n * 6 * 8, hipMemcpyDeviceToDevice, st);
  if (dw0)
    hipMemcpyAsync(dw_out, dw0, (size_t)n * m * DD * 8, hipMemcpyDeviceToDevice, st);
  else
    hipLaunchKernelGGL(pt_zero_kernel, dim3((unsigned)((n * m * DD + 255) / 256)), dim3(256), 0,
                       st, dw_out, n * m * DD);
  const unsigned bblocks = (unsigned)((n + 255) / 256);
  const unsigned rblocks = (unsigned)((n + PT_BLOCK / Lp - 1) / (PT_BLOCK / Lp));
  for (int64_t ka = 0; ka <= nsteps; ka += E) {
    const int kb = (int)std::min<int64_t>(ka + E - 1, nsteps);
    hipLaunchKernelGGL((perturb_base_kernel<HS>), dim3(bblocks), dim3(256), 0, st, hp, w_out,
                       start, n, dt, nsteps, (int)ka, kb, rows);
    hipLaunchKernelGGL((perturb_rows_kernel<SK, G, RAD>), dim3(rblocks), dim3(PT_BLOCK), lds, st,
                       pert, rows, start, n, m, Lp, t0, dt, nsteps, (int)ka, kb, dw_out);
  }
  return check_launch("ssc_perturb_leapfrog");
}

template <int SK, int G, int RAD, int HS>
static int launch_perturb(const ssc_host_potential& hp, const ssc_sources& pert, const double* w0,
                          const double* dw0, const int32_t* start, int64_t n, double t0, double dt,
                          int nsteps, double* w_out, double* dw_out, hipStream_t st) {
  const int m = pert.n;
  int L = (m + G - 1) / G;
  int Lp = 1;
  while (Lp < L) Lp <<= 1;  // lanes per particle: a power of two dividing the block
  if (Lp > PT_BLOCK) return fail(SSC_E_LIMIT, "ssc_perturb_leapfrog: too many sources");
  const int ppb = PT_BLOCK / Lp;
  const size_t lds = SourceLDS<SK>::bytes(pert);
  if (lds > 60 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)perturb_kernel<SK, G, RAD, HS>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return fail(SSC_E_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  }
  const int64_t blocks = (n + ppb - 1) / ppb;
  hipLaunchKernelGGL((perturb_kernel<SK, G, RAD, HS>), dim3((unsigned)blocks), dim3(PT_BLOCK), lds, st,
                     hp, pert, w0, dw0, start, n, m, Lp, t0, dt, nsteps, w_out, dw_out);
  return check_launch("ssc_perturb_leapfrog");
}

}  // namespace ssc

using namespace ssc;

extern "C" int64_t ssc_perturb_workspace_bytes(int64_t n, int32_t nsteps) {
  if (n <= 0 || nsteps < 0) return 0;
  const int64_t per_row = n * PR_REC * 8;
  int64_t e = ((int64_t)1 << 31) / per_row;  // rows of at most 2 GiB per segment
  if (e < 1) e = 1;
  if (e > (int64_t)nsteps + 1) e = (int64_t)nsteps + 1;
  return e * per_row;
}

extern "C" int ssc_perturb_leapfrog(const ssc_host_potential* host, const ssc_sources* pert,
                                    const double* w0, const double* dw0, const int32_t* start_step,
                                    int64_t n, double t0, double dt, int32_t nsteps,
                                    int32_t with_radius, double* w_out, double* dw_out,
                                    void* workspace, int64_t workspace_bytes, void* stream) {
  int rc = validate_host(host);
  if (rc) return rc;
  if (!pert || pert->kind == SSC_SRC_NONE || pert->n <= 0)
    return fail(SSC_E_ARG, "ssc_perturb_leapfrog: needs at least one perturbing source");
  rc = validate_sources(pert, nsteps);
  if (rc) return rc;
  if (n < 0 || nsteps < 0) return fail(SSC_E_ARG, "ssc_perturb_leapfrog: negative size");
  if (n == 0) return SSC_OK;
  if (!w0 || !w_out || !dw_out) return fail(SSC_E_ARG, "ssc_perturb_leapfrog: NULL array");
  hipStream_t st = (hipStream_t)stream;
  const ssc_sources s = *pert;
  const int hs = host_signature(*host);
  int64_t E = 0;
  if (workspace) {
    E = perturb_segment_rows(n, nsteps, workspace_bytes);
    if (E < 1)
      return fail(SSC_E_ARG, "ssc_perturb_leapfrog: workspace smaller than one row of records (" +
                                 std::to_string(n * PR_REC * 8) + " bytes)");
    if (w_out == w0 || (dw0 && dw_out == dw0))
      return fail(SSC_E_ARG, "ssc_perturb_leapfrog: outputs must not alias the inputs");
  }
#define SSC_PT3(K, R, H)                                                                        \
  (E > 0 ? launch_perturb_split<K, 8, R, H>(*host, s, w0, dw0, start_step, n, t0, dt, nsteps,    \
                                            w_out, dw_out, (double*)workspace, E, st)           \
         : launch_perturb<K, 8, R, H>(*host, s, w0, dw0, start_step, n, t0, dt, nsteps, w_out,  \
                                      dw_out, st))
#define SSC_PT(K, R)                                                                 \
  (hs == HS_MN_HQ_NFW ? SSC_PT3(K, R, HS_MN_HQ_NFW)                                  \
                      : (hs == HS_MN_HQ_HQ_NFW ? SSC_PT3(K, R, HS_MN_HQ_HQ_NFW)      \
                                               : SSC_PT3(K, R, HS_GENERIC)))
  if (with_radius) {
    switch (s.kind) {
      case SSC_SRC_STATIC: return SSC_PT(SSC_SRC_STATIC, 1);
      case SSC_SRC_TRACK: return SSC_PT(SSC_SRC_TRACK, 1);
      case SSC_SRC_LINE: return SSC_PT(SSC_SRC_LINE, 1);
      case SSC_SRC_LINE_HERNQUIST: return SSC_PT(SSC_SRC_LINE_HERNQUIST, 1);
    }
  } else {
    switch (s.kind) {
      case SSC_SRC_STATIC: return SSC_PT(SSC_SRC_STATIC, 0);
      case SSC_SRC_TRACK: return SSC_PT(SSC_SRC_TRACK, 0);
      case SSC_SRC_LINE: return SSC_PT(SSC_SRC_LINE, 0);
      case SSC_SRC_LINE_HERNQUIST: return SSC_PT(SSC_SRC_LINE_HERNQUIST, 0);
    }
  }
#undef SSC_PT
#undef SSC_PT3
  return fail(SSC_E_ARG, "ssc_perturb_leapfrog: bad source kind");
}
