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
// Layout: L = ceil(m / G) lanes per particle, each owning G responses; the base orbit and the
// host Hessian are recomputed by each of the L lanes (cheaper than exchanging them).  Velocities
// are carried at half steps so that forces are consumed as soon as they are formed.
#include <hip/hip_runtime.h>

#include <climits>

#include "ssc_device.h"
#include "ssc_sources_lds.h"

namespace ssc {

constexpr int PT_BLOCK = 256;

// host gradient and Hessian (upper triangle filled into all 9) at one point
template <int HS>
__device__ __forceinline__ void host_grad_hess(const ssc_host_potential& hp, double x, double y,
                                               double z, double* g, double* H) {
  double phi;
  host_full_sig<HS>(hp, x, y, z, phi, g, H);
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
  const int j0 = grp * G;

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
    const int j = j0 + q;
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
      const int j = j0 + q;
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
      const int j = j0 + q;
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

extern "C" int ssc_perturb_leapfrog(const ssc_host_potential* host, const ssc_sources* pert,
                                    const double* w0, const double* dw0, const int32_t* start_step,
                                    int64_t n, double t0, double dt, int32_t nsteps,
                                    int32_t with_radius, double* w_out, double* dw_out,
                                    void* stream) {
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
#define SSC_PT3(K, R, H) launch_perturb<K, 8, R, H>(*host, s, w0, dw0, start_step, n, t0, dt, nsteps, w_out, dw_out, st)
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
