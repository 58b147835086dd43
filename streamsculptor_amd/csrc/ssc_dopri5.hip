// Adaptive Dormand-Prince 5(4) orbit kernel: one orbit per lane, FP64 state and the seven
// stage derivatives in VGPRs, Plummer source tables (STATIC or LINE) in LDS.  Every lane runs
// its own step-size controller, so lanes of one wave may sit at different times; the wave
// retires when its last lane reaches t1 or runs out of steps.
//
// Replaces integrate_orbit / gen_stream_vmapped with solver=Dopri5 (reference
// streamsculptor/main.py:125-163,343-368): diffrax's diffeqsolve with
// PIDController(rtol, atol, dtmin, dtmax, force_dtmin=True, jump_ts) and SaveAt(ts).  Step
// control: integral controller on the RMS norm of err / (atol + rtol max(|y0|, |y1|)), factor
// 0.9 (1/e)^(1/5) clipped to [1, 10] on accepted and [0.2, 10] on rejected steps; a step whose
// size was clamped up to dtmin is accepted whatever its error; the starting step follows
// Hairer, Norsett & Wanner (Solving ODEs I, sec. II.4); steps ending within 1e-10 of t1 are
// stretched to t1 (or, when rejected, halved towards it).  Saved values between step ends come
// from the quartic dense output through y0, y1, h f0, h f1 and Shampine's midpoint value.
#include <hip/hip_runtime.h>

#include <climits>
#include <string>

#include "ssc_device.h"
#include "ssc_dp5.h"
#include "ssc_sources_lds.h"

namespace ssc {
namespace {

constexpr int DP_BLOCK = 256;
// Dense-output record of one accepted step, in tau = dir t: tau at the step start, the step h,
// y0[6], f0 = dy/dtau at y0 [6], Shampine's midpoint value ym[6].  The record after the last
// step holds (tau1, 0, y1, f1, 0): the quartic of step k needs y1 and f1 from record k + 1.
constexpr int DP_REC = 20;

using namespace dp5;

// One orbit's group of sources read straight from global memory: in a batch of independent
// simulations (grouped sources) every orbit has its own few subhalos, so there is no shared
// table to stage in LDS; the group's records stay in L1/L2 across the stages.
template <int SK>
struct GroupSources {
  const double* __restrict__ gm;
  const double* __restrict__ rs;
  const double* __restrict__ pos;
  const double* __restrict__ vel;
  const double* __restrict__ t0;
  const double* __restrict__ tw;
  int n;
  // False (and no source is ever read) when group[i] is outside [0, n / group_size).
  __device__ bool init(const ssc_sources& s, int64_t i) {
    n = s.group_size;
    const int g = s.group[i];
    if (g < 0 || g >= s.n / n) {
      n = 0;
      return false;
    }
    const int64_t j0 = (int64_t)g * n;
    gm = s.gm + j0;
    rs = s.rs + j0;
    pos = s.pos + 3 * j0;
    vel = LINE ? s.vel + 3 * j0 : nullptr;
    t0 = LINE ? s.t0 + j0 : nullptr;
    tw = LINE ? s.t_window + j0 : nullptr;
    return true;
  }
  static constexpr bool LINE = SK == SSC_SRC_LINE || SK == SSC_SRC_LINE_HERNQUIST;
  __device__ __forceinline__ void add(double t, double x, double y, double z, double& gx,
                                      double& gy, double& gz) const {
    for (int j = 0; j < n; ++j) {
      double sx = pos[3 * j], sy = pos[3 * j + 1], sz = pos[3 * j + 2];
      if (LINE) {
        const double dt = t - t0[j];
        if (!(fabs(dt) < tw[j])) continue;
        sx = fma(vel[3 * j], dt, sx);
        sy = fma(vel[3 * j + 1], dt, sy);
        sz = fma(vel[3 * j + 2], dt, sz);
      }
      const double a = rs[j];
      if (SK == SSC_SRC_LINE_HERNQUIST)
        hernquist_grad(x - sx, y - sy, z - sz, gm[j], a, gx, gy, gz);
      else
        plummer_grad(x - sx, y - sy, z - sz, gm[j], a * a, gx, gy, gz);
    }
  }
};

// dy/dtau at tau (t = dir * tau) for the orbit ODE [v, -grad Phi].
template <int SK, bool GR, int HS>
__device__ __forceinline__ void rhs(const ssc_host_potential& hp, const SourceLDS<SK>& lds,
                                    const GroupSources<SK>& grp, double dir, double tau,
                                    const double* y, double* k) {
  double gx, gy, gz;
  host_grad_sig<HS, false>(hp, y[0], y[1], y[2], gx, gy, gz);
  if (GR)
    grp.add(dir * tau, y[0], y[1], y[2], gx, gy, gz);
  else
    lds.template add_grad<1>(0, dir * tau, &y[0], &y[1], &y[2], &gx, &gy, &gz, 0, 1);
  k[0] = dir * y[3];
  k[1] = dir * y[4];
  k[2] = dir * y[5];
  k[3] = -dir * gx;
  k[4] = -dir * gy;
  k[5] = -dir * gz;
}

__device__ __forceinline__ double rms6(const double* u) {
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < 6; ++c) s = fma(u[c], u[c], s);
  return sqrt(s * (1.0 / 6.0));
}

// End a step at the first stop strictly after tau0 when the step would pass it.
__device__ __forceinline__ double clip_stops(const double* __restrict__ stops, int n_stops,
                                             double dir, double tau0, double tau1) {
  double best = tau1;
  for (int j = 0; j < n_stops; ++j) {
    const double s = dir * stops[j];
    if (s > tau0 && s < best) best = s;
  }
  return best;
}

template <int SK, bool GR, int HS>
__global__ __launch_bounds__(DP_BLOCK) void dopri5_kernel(
    ssc_host_potential hp, ssc_sources src, const double* __restrict__ w0,
    const double* __restrict__ t_start, const double* __restrict__ t_end, int64_t n, double t0,
    double t1, const double* __restrict__ ts, int nts, int64_t ts_stride,
    const double* __restrict__ stops, Dp5Ctl ctl, double* __restrict__ w_out,
    double* __restrict__ ys, int32_t* __restrict__ stats, double* __restrict__ dense,
    int dense_cap) {
  extern __shared__ __align__(16) char lds_raw[];
  SourceLDS<SK> lds;
  if (!GR) lds.init(src, lds_raw);  // the only barrier: every lane reaches it
  const int64_t i = (int64_t)blockIdx.x * DP_BLOCK + threadIdx.x;
  if (i >= n) return;
  GroupSources<SK> grp;
  if (GR && !grp.init(src, i)) {  // bad group index: status 2, NaN rows, nothing integrated
    if (stats) {
      stats[3 * i] = 0;
      stats[3 * i + 1] = 0;
      stats[3 * i + 2] = 2;
    }
    double* yr = ys ? ys + (int64_t)6 * nts * i : nullptr;
    if (yr)
      for (int j = 0; j < 6 * nts; ++j) yr[j] = __builtin_nan("");
    if (w_out)
      for (int c = 0; c < 6; ++c) w_out[6 * i + c] = __builtin_nan("");
    return;
  }

  // orbit i runs from t_start[i] to t_end[i] (or the shared t0 / t1) and saves at its own row
  // of ts when ts_stride = nts (the shared times when ts_stride = 0)
  const double ta = t_start ? t_start[i] : t0;
  const double tb = t_end ? t_end[i] : t1;
  const double dir = tb >= ta ? 1.0 : -1.0;
  const double tau0 = dir * ta, tau1 = dir * tb;
  const double* __restrict__ tsi = ts ? ts + ts_stride * i : nullptr;
  double y[6], f0[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    y[c] = w0[6 * i + c];
    f0[c] = 0.0;
  }
  double* __restrict__ yrow = ys ? ys + (int64_t)6 * nts * i : nullptr;
  double* __restrict__ drow = dense ? dense + (int64_t)dense_cap * DP_REC * i : nullptr;
  int js = 0;
  for (; js < nts && dir * tsi[js] < tau0; ++js)
#pragma unroll
    for (int c = 0; c < 6; ++c) yrow[6 * js + c] = __builtin_nan("");
  for (; js < nts && dir * tsi[js] == tau0; ++js)
#pragma unroll
    for (int c = 0; c < 6; ++c) yrow[6 * js + c] = y[c];

  int nacc = 0, nrej = 0, nsteps = 0;
  double tprev = tau0;
  if (tau0 < tau1) {
    rhs<SK, GR, HS>(hp, lds, grp, dir, tau0, y, f0);
    double tnext;
    bool at_dtmin = false;
    {
      double sc[6], u[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) sc[c] = ctl.atol + fabs(y[c]) * ctl.rtol;
#pragma unroll
      for (int c = 0; c < 6; ++c) u[c] = y[c] / sc[c];
      const double d0 = rms6(u);
#pragma unroll
      for (int c = 0; c < 6; ++c) u[c] = f0[c] / sc[c];
      const double d1 = rms6(u);
      double dt0 = ctl.dt0;
      if (!(dt0 > 0.0)) {
        const double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
        double y1[6], f1[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) y1[c] = y[c] + h0 * f0[c];
        rhs<SK, GR, HS>(hp, lds, grp, dir, tau0 + h0, y1, f1);
#pragma unroll
        for (int c = 0; c < 6; ++c) u[c] = (f1[c] - f0[c]) / sc[c];
        const double d2 = rms6(u) / h0;
        const double md = fmax(d1, d2);
        const double h1 = md <= 1e-15 ? fmax(1e-6, h0 * 1e-3) : pow(0.01 / md, 0.2);
        dt0 = fmin(100.0 * h0, h1);
      }
      // the controller's init clamps the first step to [dtmin, dtmax] like every later one:
      // a first step raised to dtmin is accepted whatever its error
      if (ctl.dtmax > 0.0) dt0 = fmin(dt0, ctl.dtmax);
      if (ctl.dtmin > 0.0) {
        at_dtmin = dt0 <= ctl.dtmin;
        dt0 = fmax(dt0, ctl.dtmin);
      }
      tnext = clip_stops(stops, ctl.n_stops, dir, tau0, tau0 + dt0);
      if (tnext > tau1) tnext = tau1;
    }
    while (tprev < tau1 && nsteps < ctl.max_steps) {
      const double h = tnext - tprev;
      double k2[6], k3[6], k4[6], k5[6], k6[6], k7[6], yt[6], yn[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) yt[c] = y[c] + h * (A21 * f0[c]);
      rhs<SK, GR, HS>(hp, lds, grp, dir, tprev + C2 * h, yt, k2);
#pragma unroll
      for (int c = 0; c < 6; ++c) yt[c] = y[c] + h * (A31 * f0[c] + A32 * k2[c]);
      rhs<SK, GR, HS>(hp, lds, grp, dir, tprev + C3 * h, yt, k3);
#pragma unroll
      for (int c = 0; c < 6; ++c) yt[c] = y[c] + h * (A41 * f0[c] + A42 * k2[c] + A43 * k3[c]);
      rhs<SK, GR, HS>(hp, lds, grp, dir, tprev + C4 * h, yt, k4);
#pragma unroll
      for (int c = 0; c < 6; ++c)
        yt[c] = y[c] + h * (A51 * f0[c] + A52 * k2[c] + A53 * k3[c] + A54 * k4[c]);
      rhs<SK, GR, HS>(hp, lds, grp, dir, tprev + C5 * h, yt, k5);
#pragma unroll
      for (int c = 0; c < 6; ++c)
        yt[c] = y[c] + h * (A61 * f0[c] + A62 * k2[c] + A63 * k3[c] + A64 * k4[c] + A65 * k5[c]);
      rhs<SK, GR, HS>(hp, lds, grp, dir, tnext, yt, k6);
#pragma unroll
      for (int c = 0; c < 6; ++c)
        yn[c] = y[c] + h * (B1 * f0[c] + B3 * k3[c] + B4 * k4[c] + B5 * k5[c] + B6 * k6[c]);
      rhs<SK, GR, HS>(hp, lds, grp, dir, tnext, yn, k7);
      double u[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const double err = h * (E1 * f0[c] + E3 * k3[c] + E4 * k4[c] + E5 * k5[c] + E6 * k6[c] +
                                E7 * k7[c]);
        u[c] = err / (ctl.atol + ctl.rtol * fmax(fabs(y[c]), fabs(yn[c])));
      }
      const double e = rms6(u);
      const bool keep = (e < 1.0) || at_dtmin;
      double factor = 0.9 * pow(1.0 / e, 0.2);
      factor = fmin(fmax(factor, keep ? 1.0 : 0.2), 10.0);
      double dtn = h * factor;
      if (ctl.dtmax > 0.0 && dtn > ctl.dtmax) dtn = ctl.dtmax;
      if (ctl.dtmin > 0.0) {
        at_dtmin = dtn <= ctl.dtmin;
        if (dtn < ctl.dtmin) dtn = ctl.dtmin;
      }
      double t_new = tprev;
      if (keep && drow) {
        double* r = drow + (int64_t)nacc * DP_REC;
        r[0] = tprev;
        r[1] = h;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          r[2 + c] = y[c];
          r[8 + c] = f0[c];
          r[14 + c] = y[c] + h * (M1 * f0[c] + M3 * k3[c] + M4 * k4[c] + M5 * k5[c] +
                                  M6 * k6[c] + M7 * k7[c]);
        }
      }
      if (keep) {
        if (yrow && js < nts && dir * tsi[js] <= tnext) {
          // dense output on (tprev, tnext]
          double F0[6], a4[6], b3[6], c2[6];
#pragma unroll
          for (int c = 0; c < 6; ++c) {
            const double ym = y[c] + h * (M1 * f0[c] + M3 * k3[c] + M4 * k4[c] + M5 * k5[c] +
                                          M6 * k6[c] + M7 * k7[c]);
            F0[c] = h * f0[c];
            const double D1 = yn[c] - y[c] - F0[c], D2 = h * k7[c] - F0[c],
                         D3 = ym - y[c] - 0.5 * F0[c];
            a4[c] = 2.0 * D2 - 8.0 * D1 + 16.0 * D3;
            b3[c] = D2 - 2.0 * D1 - 2.0 * a4[c];
            c2[c] = D1 - a4[c] - b3[c];
          }
          for (; js < nts && dir * tsi[js] <= tnext; ++js) {
            const double tj = dir * tsi[js];
            if (tj == tnext) {
#pragma unroll
              for (int c = 0; c < 6; ++c) yrow[6 * js + c] = yn[c];
            } else {
              const double s = (tj - tprev) / h;
#pragma unroll
              for (int c = 0; c < 6; ++c)
                yrow[6 * js + c] = y[c] + s * (F0[c] + s * (c2[c] + s * (b3[c] + s * a4[c])));
            }
          }
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          y[c] = yn[c];
          f0[c] = k7[c];
        }
        t_new = tnext;
        ++nacc;
      } else {
        ++nrej;
      }
      double tn1 = clip_stops(stops, ctl.n_stops, dir, t_new, t_new + dtn);
      tprev = t_new < tau1 ? t_new : tau1;
      if (tn1 > tau1 - 1e-10) tn1 = keep ? tau1 : tprev + 0.5 * (tau1 - tprev);
      tnext = tn1;
      ++nsteps;
    }
  }
  const bool failed = tprev < tau1;
  if (stats) {
    stats[3 * i] = nacc;
    stats[3 * i + 1] = nrej;
    stats[3 * i + 2] = failed ? 1 : 0;
  }
  if (failed) {
    const double inf = __builtin_inf();
    if (yrow)
      for (int j = 0; j < nts; ++j)
#pragma unroll
        for (int c = 0; c < 6; ++c) yrow[6 * j + c] = inf;
    if (w_out)
#pragma unroll
      for (int c = 0; c < 6; ++c) w_out[6 * i + c] = inf;
    return;
  }
  if (yrow)
    for (; js < nts; ++js)
#pragma unroll
      for (int c = 0; c < 6; ++c) yrow[6 * js + c] = __builtin_nan("");
  if (w_out)
#pragma unroll
    for (int c = 0; c < 6; ++c) w_out[6 * i + c] = y[c];
  if (drow) {  // closing record: the end point of the last step
    double* r = drow + (int64_t)nacc * DP_REC;
    r[0] = tprev;
    r[1] = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      r[2 + c] = y[c];
      r[8 + c] = f0[c];
      r[14 + c] = 0.0;
    }
  }
}

int set_lds_dp(const void* fn, size_t lds) {
  if (lds > 160 * 1024) return fail(SSC_E_LIMIT, "ssc_dopri5: source table exceeds LDS");
  if (lds > 60 * 1024) {
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess)
      return fail(SSC_E_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  }
  return SSC_OK;
}

template <int SK, bool GR, int HS>
int launch_dopri5(const ssc_host_potential& hp, const ssc_sources& src, const double* w0,
                  const double* t_start, const double* t_end, int64_t n, double t0, double t1,
                  const double* ts, int nts, int64_t ts_stride, const double* stops,
                  const Dp5Ctl& ctl, double* w_out, double* ys, int32_t* stats, double* dense,
                  int dense_cap, hipStream_t st) {
  const size_t lds = GR ? 0 : SourceLDS<SK>::bytes(src);
  int rc = set_lds_dp((const void*)dopri5_kernel<SK, GR, HS>, lds);
  if (rc) return rc;
  const int64_t blocks = (n + DP_BLOCK - 1) / DP_BLOCK;
  hipLaunchKernelGGL((dopri5_kernel<SK, GR, HS>), dim3((unsigned)blocks), dim3(DP_BLOCK), lds, st, hp,
                     src, w0, t_start, t_end, n, t0, t1, ts, nts, ts_stride, stops, ctl, w_out,
                     ys, stats, dense, dense_cap);
  return check_launch("ssc_dopri5");
}

}  // namespace
}  // namespace ssc

using namespace ssc;

static int dopri5_entry(const ssc_host_potential* host, const ssc_sources* src, const double* w0,
                        const double* t_start, const double* t_end, int64_t n, double t0,
                        double t1, const double* ts, int32_t nts, bool per_orbit_ts,
                        const ssc_dopri5_options* opt, double* w_out, double* ys,
                        int32_t* stats, double* dense, int32_t dense_cap, void* stream) {
  int rc = validate_host(host);
  if (rc) return rc;
  if (!opt) return fail(SSC_E_ARG, "ssc_dopri5: options are NULL");
  if (n < 0 || nts < 0 || opt->n_stops < 0) return fail(SSC_E_ARG, "ssc_dopri5: negative size");
  if (n == 0) return SSC_OK;
  if (!w0) return fail(SSC_E_ARG, "ssc_dopri5: w0 is NULL");
  if (!w_out && !(ys && nts > 0) && !dense)
    return fail(SSC_E_ARG, "ssc_dopri5: no output (w_out, ys, dense)");
  if (nts > 0 && (!ts || !ys)) return fail(SSC_E_ARG, "ssc_dopri5: nts > 0 needs ts and ys");
  if (opt->n_stops > 0 && !opt->stops) return fail(SSC_E_ARG, "ssc_dopri5: n_stops > 0 needs stops");
  if (!(opt->rtol >= 0.0) || !(opt->atol >= 0.0) || !(opt->rtol + opt->atol > 0.0))
    return fail(SSC_E_ARG, "ssc_dopri5: rtol/atol must be >= 0 and not both 0");
  if (opt->max_steps < 1) return fail(SSC_E_ARG, "ssc_dopri5: max_steps < 1");
  if (dense && dense_cap < opt->max_steps + 1)
    return fail(SSC_E_ARG, "ssc_dopri5_dense: dense_cap must be at least max_steps + 1");
  if ((n + DP_BLOCK - 1) / DP_BLOCK > (int64_t)INT_MAX)
    return fail(SSC_E_LIMIT, "ssc_dopri5: too many orbits");
  rc = validate_sources(src, 0, true);
  if (rc) return rc;
  ssc_sources s;
  if (src) {
    s = *src;
  } else {
    s = ssc_sources{};
  }
  if (s.n == 0) s.kind = SSC_SRC_NONE;
  if (s.kind == SSC_SRC_TRACK || s.kind == SSC_SRC_TRACK_HERNQUIST)
    return fail(SSC_E_ARG, "ssc_dopri5: TRACK sources are sampled on a fixed grid; adaptive steps "
                           "need STATIC or LINE sources");
  if (nts == 0) ys = nullptr;
  Dp5Ctl ctl{opt->rtol, opt->atol, opt->dtmin, opt->dtmax, opt->dt0, opt->max_steps, opt->n_stops};
  hipStream_t st = (hipStream_t)stream;
  const bool grouped = s.kind != SSC_SRC_NONE && s.group_size > 0;
  const int hs = host_signature(*host);
#define SSC_DP3(K, G, H)                                                                          \
  launch_dopri5<K, G, H>(*host, s, w0, t_start, t_end, n, t0, t1, ts, nts,                       \
                         per_orbit_ts ? (int64_t)nts : 0, opt->stops, ctl, w_out, ys, stats,     \
                         dense, dense_cap, st)
#define SSC_DP(K, G)                                                                              \
  (hs == HS_MN_HQ_NFW ? SSC_DP3(K, G, HS_MN_HQ_NFW)                                              \
                      : (hs == HS_MN_HQ_HQ_NFW ? SSC_DP3(K, G, HS_MN_HQ_HQ_NFW)                  \
                                               : SSC_DP3(K, G, HS_GENERIC)))
  switch (s.kind) {
    case SSC_SRC_NONE: return SSC_DP(SSC_SRC_NONE, false);
    case SSC_SRC_STATIC: return grouped ? SSC_DP(SSC_SRC_STATIC, true) : SSC_DP(SSC_SRC_STATIC, false);
    case SSC_SRC_LINE: return grouped ? SSC_DP(SSC_SRC_LINE, true) : SSC_DP(SSC_SRC_LINE, false);
    case SSC_SRC_LINE_HERNQUIST:
      return grouped ? SSC_DP(SSC_SRC_LINE_HERNQUIST, true) : SSC_DP(SSC_SRC_LINE_HERNQUIST, false);
    default: return fail(SSC_E_ARG, "ssc_dopri5: bad source kind");
  }
#undef SSC_DP
#undef SSC_DP3
}

extern "C" int ssc_dopri5(const ssc_host_potential* host, const ssc_sources* src,
                          const double* w0, const double* t_start, int64_t n, double t0,
                          double t1, const double* ts, int32_t nts,
                          const ssc_dopri5_options* opt, double* w_out, double* ys,
                          int32_t* stats, void* stream) {
  return dopri5_entry(host, src, w0, t_start, nullptr, n, t0, t1, ts, nts, false, opt, w_out, ys,
                      stats, nullptr, 0, stream);
}

extern "C" int ssc_dopri5_orbits(const ssc_host_potential* host, const ssc_sources* src,
                                 const double* w0, const double* t_start, const double* t_end,
                                 int64_t n, const double* ts, int32_t nts,
                                 const ssc_dopri5_options* opt, double* w_out, double* ys,
                                 int32_t* stats, void* stream) {
  if (n > 0 && (!t_start || !t_end))
    return fail(SSC_E_ARG, "ssc_dopri5_orbits: t_start and t_end are required");
  return dopri5_entry(host, src, w0, t_start, t_end, n, 0.0, 0.0, ts, nts, true, opt, w_out, ys,
                      stats, nullptr, 0, stream);
}

extern "C" int ssc_dopri5_dense(const ssc_host_potential* host, const ssc_sources* src,
                                const double* w0, const double* t_start, int64_t n, double t0,
                                double t1, const ssc_dopri5_options* opt, double* w_out,
                                double* dense, int32_t dense_cap, int32_t* stats, void* stream) {
  if (!dense || !stats) return fail(SSC_E_ARG, "ssc_dopri5_dense: dense and stats are required");
  return dopri5_entry(host, src, w0, t_start, nullptr, n, t0, t1, nullptr, 0, false, opt, w_out,
                      nullptr, stats, dense, dense_cap, stream);
}
