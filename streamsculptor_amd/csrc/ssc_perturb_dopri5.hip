// Adaptive Dormand-Prince 5(4) integration of base orbits together with their first-order
// responses to perturbing subhalos (the fields-interface OTF perturbation ODE with a
// PID-controlled Runge-Kutta solver, as reference GenerateMassRadiusPerturbation runs it:
// streamsculptor/perturbative.py:52-60 for the progenitor, :95-135 for the particles, through
// fields.integrate_field, fields.py:34-99, and MassRadiusPerturbation_OTF.term, :175-206).
//
// The state of one particle is the reference's pytree [w (6), dw (m, 6 | 12)]: the base row and,
// per subhalo j, the mass-block response d w / d eps_j and (with the radius block)
// d^2 w / (d a_j d eps_j).  The step controller is shared by the whole state: its error norm is
// the RMS over all 6 + 6 nvec components (nvec = m or 2 m response vectors), as diffrax's
// PIDController takes the RMS norm over every leaf of the pytree.  Step control otherwise
// follows the adaptive orbit kernel (integral controller, dtmin forcing, Hairer's starting step,
// end-of-interval clipping, quartic dense output at save times).
//
// Layout: a particle is spread over a group of L lanes (a power of two, 8 <= L <= 256); lane l
// owns the response vectors l, l + L, ... (G = 1, 2 or 4 of them) and, for l < 6, base
// component l.  At each
// stage the six base owners publish the stage row through LDS, lane c of the group forms host
// component c's gradient and Hessian there and publishes them, and every lane sums the
// components, so the base orbit and the host terms are formed once per group.  The stage values
// and the error norm are block-synchronous: every lane of the block passes the same barriers,
// and a group whose particle has finished idles until the block's last particle is done.
#include <hip/hip_runtime.h>

#include <climits>
#include <string>
#include <type_traits>

#include "ssc_device.h"
#include "ssc_dp5.h"
#include "ssc_sources_lds.h"

namespace ssc {
namespace {

using namespace dp5;

constexpr int PD_BLOCK = 256;
constexpr int PD_MIN_L = 8;
constexpr int PD_GROUPS = PD_BLOCK / PD_MIN_L;
constexpr int PD_WAVES = PD_BLOCK / 64;

// Sum of v over the L lanes of a particle group, bit-identical in every lane of the group: the
// butterfly forms each pairwise sum from the same two operands in both partner lanes; groups
// wider than a wave add their wave partials in wave order through LDS (one barrier, taken when
// L > 64, which is uniform over the launch).
__device__ __forceinline__ double group_sum(double v, int L, double* part) {
  const int w = L < 64 ? L : 64;
  for (int o = w >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, w);
  if (L <= 64) return v;
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  const int nw = L >> 6, first = (threadIdx.x / L) * nw;
  double s = 0.0;
  for (int k = 0; k < nw; ++k) s += part[first + k];
  return s;
}

// Stage inputs: S = 0 the state itself, 1..5 the Dormand-Prince stages 2..6, 6 the 5th-order
// solution (input of the FSAL stage), 7 the starting-step probe y + h f0.
template <int S>
__device__ __forceinline__ double stage_in(double y, const double (&k)[7], double h) {
  if constexpr (S == 0) return y;
  if constexpr (S == 1) return y + h * (A21 * k[0]);
  if constexpr (S == 2) return y + h * (A31 * k[0] + A32 * k[1]);
  if constexpr (S == 3) return y + h * (A41 * k[0] + A42 * k[1] + A43 * k[2]);
  if constexpr (S == 4) return y + h * (A51 * k[0] + A52 * k[1] + A53 * k[2] + A54 * k[3]);
  if constexpr (S == 5)
    return y + h * (A61 * k[0] + A62 * k[1] + A63 * k[2] + A64 * k[3] + A65 * k[4]);
  if constexpr (S == 6)
    return y + h * (B1 * k[0] + B3 * k[2] + B4 * k[3] + B5 * k[4] + B6 * k[5]);
  if constexpr (S == 7) return y + h * k[0];
  return 0.0;
}
// Output slot of stage S (the probe reuses slot 1, which the first step overwrites).
template <int S>
constexpr int stage_out() { return S == 7 ? 1 : S; }
// Abscissa of stage S as a fraction of h.
template <int S>
__device__ __forceinline__ double stage_c() {
  if constexpr (S == 1) return C2;
  if constexpr (S == 2) return C3;
  if constexpr (S == 3) return C4;
  if constexpr (S == 4) return C5;
  if constexpr (S >= 5) return 1.0;
  return 0.0;
}

__device__ __forceinline__ double err_of(double y, const double (&k)[7], double yn, double h,
                                         const Dp5Ctl& ctl) {
  const double err =
      h * (E1 * k[0] + E3 * k[2] + E4 * k[3] + E5 * k[4] + E6 * k[5] + E7 * k[6]);
  const double u = err / (ctl.atol + ctl.rtol * fmax(fabs(y), fabs(yn)));
  return u * u;
}

// Quartic through y (s = 0), yn (s = 1), slopes F0 = h f0 and F7 = h f7, and the midpoint
// value ym, at s in (0, 1).
__device__ __forceinline__ double quartic(double y, double yn, double F0, double F7, double ym,
                                          double s) {
  const double D1 = yn - y - F0, D2 = F7 - F0, D3 = ym - y - 0.5 * F0;
  const double a4 = 2.0 * D2 - 8.0 * D1 + 16.0 * D3;
  const double b3 = D2 - 2.0 * D1 - 2.0 * a4;
  const double c2 = D1 - a4 - b3;
  return y + s * (F0 + s * (c2 + s * (b3 + s * a4)));
}

// Dense output of a component with stored slopes k[0..6].
__device__ __forceinline__ double dense_at(double y, const double (&k)[7], double yn, double h,
                                           double s) {
  const double ym =
      y + h * (M1 * k[0] + M3 * k[2] + M4 * k[3] + M5 * k[4] + M6 * k[5] + M7 * k[6]);
  return quartic(y, yn, h * k[0], h * k[6], ym, s);
}

// Position components of a response vector keep no slopes of their own: the slope of position
// component c at stage j is dir V_j, V_j the stage value of velocity component c,
// V_j = v0 + h sum_l a_jl kv_l.  A weighted sum h sum_j w_j kx_j is then
// dir h (sum_j w_j v0 + h sum_l (sum_j w_j a_jl) kv_l), with the products of the tableau
// formed at compile time.  Weight rows T: 0..6 the stage inputs (6 = b, the FSAL input),
// 7 the starting-step probe (w = e_0), 8 the error weights b - bhat, 9 the midpoint weights.
constexpr double butcher(int j, int l) {
  constexpr double a[7][6] = {{0, 0, 0, 0, 0, 0},
                              {A21, 0, 0, 0, 0, 0},
                              {A31, A32, 0, 0, 0, 0},
                              {A41, A42, A43, 0, 0, 0},
                              {A51, A52, A53, A54, 0, 0},
                              {A61, A62, A63, A64, A65, 0},
                              {B1, 0, B3, B4, B5, B6}};
  return a[j][l];
}
constexpr double wrow(int T, int j) {
  constexpr double e[7] = {E1, 0, E3, E4, E5, E6, E7};
  constexpr double mw[7] = {M1, 0, M3, M4, M5, M6, M7};
  if (T == 7) return j == 0 ? 1.0 : 0.0;
  if (T == 8) return e[j];
  if (T == 9) return mw[j];
  return j < 6 ? butcher(T, j) : 0.0;
}
constexpr double qcoef(int T, int l) {
  double s = 0.0;
  for (int j = 0; j < 7; ++j) s += wrow(T, j) * butcher(j, l);
  return s;
}
constexpr double ccoef(int T) {
  if (T == 8) return 0.0;  // b and bhat both sum to one
  double s = 0.0;
  for (int j = 0; j < 7; ++j) s += wrow(T, j);
  return s;
}
// sum_{l <= L} qcoef(T, l) kv[l], zero coefficients skipped.
template <int T, int L>
__device__ __forceinline__ double qsum(const double (&kv)[7]) {
  if constexpr (L < 0) {
    return 0.0;
  } else {
    constexpr double q = qcoef(T, L);
    if constexpr (q != 0.0)
      return fma(q, kv[L], qsum<T, L - 1>(kv));
    else
      return qsum<T, L - 1>(kv);
  }
}
// x0 + h sum_j w_j kx_j for weight row T, from the velocity component's v0 and slopes.
template <int T>
__device__ __forceinline__ double pos_sum(double x0, double v0, const double (&kv)[7], double h,
                                          double dir) {
  if constexpr (T == 0) {
    return x0;
  } else {
    constexpr double c = ccoef(T);
    return x0 + dir * h * (c * v0 + h * qsum<T, 5>(kv));
  }
}

// One response vector per lane fits two waves per SIMD with a few spilled registers (1.5x
// faster than one wave per SIMD without spills); two vectors per lane need one wave per SIMD,
// four (513 to 1024 vectors) spill about a hundred registers at one wave per SIMD.
template <int SK, int RAD, int G>
__global__ __launch_bounds__(PD_BLOCK) __attribute__((amdgpu_waves_per_eu(G == 1 ? 2 : 1,
                                                                          G == 1 ? 2 : 1)))
void perturb_dopri5_kernel(
    ssc_host_potential hp, ssc_sources pert, const double* __restrict__ w0,
    const double* __restrict__ dw0, const double* __restrict__ t_start, int64_t n, int m, int L,
    double t0, double t1, const double* __restrict__ ts, int nts, Dp5Ctl ctl,
    double* __restrict__ w_out, double* __restrict__ dw_out, double* __restrict__ ys,
    double* __restrict__ dys, int32_t* __restrict__ stats) {
  extern __shared__ __align__(16) char lds_raw[];
  __shared__ double xv_sh[PD_GROUPS][2][6];
  __shared__ double part_sh[2][PD_WAVES];
  __shared__ double hg_sh[PD_GROUPS][SSC_MAX_COMPONENTS][9];  // host component partials
  constexpr int DD = RAD ? 12 : 6;
  SourceLDS<SK> lds;
  lds.init(pert, lds_raw);  // barrier inside

  const int nvec = RAD ? 2 * m : m;
  const double ntot = 6.0 + 6.0 * nvec;
  const int ppb = PD_BLOCK / L;
  const int gi = threadIdx.x / L, lane = threadIdx.x - gi * L;
  const int64_t i = (int64_t)blockIdx.x * ppb + gi;
  const bool valid = i < n;
  const int bc = lane < 6 ? lane : -1;  // owned base component

  bool vh[G];  // vector present
  int vj[G], voff[G];
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int v = lane + q * L;
    vh[q] = v < nvec;
    const bool rad = RAD && v >= m;
    vj[q] = vh[q] ? (rad ? v - m : v) : 0;
    voff[q] = rad ? 6 : 0;
  }

  const double ta = valid ? (t_start ? t_start[i] : t0) : t1;
  const double dir = t1 >= ta ? 1.0 : -1.0;
  const double tau0 = dir * ta, tau1 = dir * t1;

  // state: base component (owners) and response vectors, with the seven stage slopes
  double by = 0.0, byn = 0.0, bk[7];
  double y[G][6], yn[G][6], kv[G][3][7];  // kv: slopes of the velocity components
#pragma unroll
  for (int s = 0; s < 7; ++s) bk[s] = 0.0;
  if (valid && bc >= 0) by = w0[6 * i + bc];
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const bool ok = valid && vh[q] && dw0;
    const double* d = ok ? dw0 + ((int64_t)i * m + vj[q]) * DD + voff[q] : nullptr;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      y[q][c] = ok ? d[c] : 0.0;
      yn[q][c] = 0.0;
#pragma unroll
      for (int s = 0; s < 7; ++s)
        if (c < 3) kv[q][c][s] = 0.0;
    }
  }

  // saves at or before the start
  auto save_row = [&](int js, bool interp, double s, double h) {
    const int64_t row = (int64_t)i * nts + js;
    if (bc >= 0)
      ys[row * 6 + bc] = interp ? dense_at(by, bk, byn, h, s) : by;
#pragma unroll
    for (int q = 0; q < G; ++q) {
      if (!vh[q]) continue;
      double* o = dys + (row * m + vj[q]) * DD + voff[q];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!interp) {
          o[c] = y[q][c];
          o[3 + c] = y[q][3 + c];
          continue;
        }
        const double x0 = y[q][c], v0 = y[q][3 + c];
        o[c] = quartic(x0, yn[q][c], dir * h * v0, dir * h * yn[q][3 + c],
                       pos_sum<9>(x0, v0, kv[q][c], h, dir), s);
        o[3 + c] = dense_at(v0, kv[q][c], yn[q][3 + c], h, s);
      }
    }
  };
  auto fill_row = [&](int js, double val) {
    const int64_t row = (int64_t)i * nts + js;
    if (bc >= 0) ys[row * 6 + bc] = val;
#pragma unroll
    for (int q = 0; q < G; ++q) {
      if (!vh[q]) continue;
      double* o = dys + (row * m + vj[q]) * DD + voff[q];
#pragma unroll
      for (int c = 0; c < 6; ++c) o[c] = val;
    }
  };
  int js = 0;
  if (valid && nts > 0) {
    for (; js < nts && dir * ts[js] < tau0; ++js) fill_row(js, __builtin_nan(""));
    for (; js < nts && dir * ts[js] == tau0; ++js) save_row(js, false, 0.0, 0.0);
  }

  bool live = valid && tau0 < tau1;
  int xb = 0, pb = 0;

  // Stage S at tau_s = tau + c_S h: publish the base stage value; lane c < n_components of the
  // group forms host component c's gradient and Hessian at the stage position and publishes
  // them; every lane sums the components in order, then forms its slopes.  Two barriers.
  auto stage = [&](auto S_, double tau, double h) {
    constexpr int S = decltype(S_)::value;
    constexpr int O = stage_out<S>();
    const double bval = stage_in<S>(by, bk, h);
    if (live && bc >= 0) xv_sh[gi][xb][bc] = bval;
    if (S == 6) byn = bval;
    __syncthreads();
    const double* xv = xv_sh[gi][xb];
    if (live && lane < hp.n_components) {
      double phi = 0.0, gc[3] = {0.0, 0.0, 0.0}, Hc[9];
#pragma unroll
      for (int r = 0; r < 9; ++r) Hc[r] = 0.0;
      comp_full(hp.c[lane], xv[0], xv[1], xv[2], phi, gc, Hc);
      double* o = hg_sh[gi][lane];
      o[0] = gc[0]; o[1] = gc[1]; o[2] = gc[2];
      o[3] = Hc[0]; o[4] = Hc[1]; o[5] = Hc[2]; o[6] = Hc[4]; o[7] = Hc[5]; o[8] = Hc[8];
    }
    __syncthreads();
    if (live) {
      const double px = xv[0], py = xv[1], pz = xv[2];
      double g[3] = {0.0, 0.0, 0.0}, H[9];
#pragma unroll
      for (int r = 0; r < 9; ++r) H[r] = 0.0;
      for (int c = 0; c < hp.n_components; ++c) {
        const double* o = hg_sh[gi][c];
        g[0] += o[0]; g[1] += o[1]; g[2] += o[2];
        H[0] += o[3]; H[1] += o[4]; H[2] += o[5]; H[4] += o[6]; H[5] += o[7]; H[8] += o[8];
      }
      H[3] = H[1]; H[6] = H[2]; H[7] = H[5];
      if (bc >= 0) bk[O] = bc < 3 ? dir * xv[3 + bc] : -dir * g[bc - 3];
      const double t = dir * (tau + stage_c<S>() * h);
#pragma unroll
      for (int q = 0; q < G; ++q) {
        double yt[6];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          yt[c] = pos_sum<S>(y[q][c], y[q][3 + c], kv[q][c], h, dir);
          yt[3 + c] = stage_in<S>(y[q][3 + c], kv[q][c], h);
        }
        if (S == 6)
#pragma unroll
          for (int c = 0; c < 6; ++c) yn[q][c] = yt[c];
        double fx = 0.0, fy = 0.0, fz = 0.0;
        double sx, sy, sz, gm, qs;
        if (vh[q] && lds.get(0, t, vj[q], sx, sy, sz, gm, qs)) {
          const double ex = px - sx, ey = py - sy, ez = pz - sz;
          double f, fr;  // grad Phi_j = f e; grad(dPhi_j/da_j) = fr e
          if (SK == SSC_SRC_LINE_HERNQUIST) {  // qs = a
            const double r2 = fma(ex, ex, fma(ey, ey, ez * ez));
            const double ir = rsq(r2);
            const double ia = 1.0 / fma(r2, ir, qs);
            f = gm * ir * ia * ia;
            fr = -2.0 * f * ia;
          } else {  // qs = a^2
            const double is = rsq(fma(ex, ex, fma(ey, ey, fma(ez, ez, qs))));
            f = gm * is * is * is;
            fr = -3.0 * f * sqrt(qs) * is * is;
          }
          const double ff = voff[q] ? fr : f;
          fx = ff * ex; fy = ff * ey; fz = ff * ez;
        }
        const double hx = H[0] * yt[0] + H[1] * yt[1] + H[2] * yt[2];
        const double hy = H[3] * yt[0] + H[4] * yt[1] + H[5] * yt[2];
        const double hz = H[6] * yt[0] + H[7] * yt[1] + H[8] * yt[2];
        kv[q][0][O] = -dir * (fx + hx);
        kv[q][1][O] = -dir * (fy + hy);
        kv[q][2][O] = -dir * (fz + hz);
      }
    }
    xb ^= 1;
  };
  auto reduce = [&](double v) {
    const double s = group_sum(v, L, part_sh[pb]);
    pb ^= 1;
    return s;
  };

  int nacc = 0, nrej = 0, nst = 0;
  double tprev = tau0, tnext = tau0;
  bool at_dtmin = false;

  stage(std::integral_constant<int, 0>{}, tau0, 0.0);  // f0 at the start
  {
    // starting step (Hairer, Norsett & Wanner, Solving ODEs I, sec. II.4) on the whole state
    double s0 = 0.0, s1 = 0.0;
    if (live) {
      if (bc >= 0) {
        const double sc = ctl.atol + fabs(by) * ctl.rtol;
        s0 = (by / sc) * (by / sc);
        s1 = (bk[0] / sc) * (bk[0] / sc);
      }
#pragma unroll
      for (int q = 0; q < G; ++q)
        if (vh[q])
#pragma unroll
          for (int c = 0; c < 6; ++c) {
            const double sc = ctl.atol + fabs(y[q][c]) * ctl.rtol;
            const double f0 = c < 3 ? dir * y[q][3 + c] : kv[q][c - 3][0];
            const double a = y[q][c] / sc, b = f0 / sc;
            s0 = fma(a, a, s0);
            s1 = fma(b, b, s1);
          }
    }
    double dt0 = ctl.dt0;
    if (!(dt0 > 0.0)) {  // uniform over the launch
      const double d0 = sqrt(reduce(s0) / ntot);
      const double d1 = sqrt(reduce(s1) / ntot);
      const double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
      stage(std::integral_constant<int, 7>{}, tau0, h0);  // f(tau0 + h0, y + h0 f0) -> slot 1
      double s2 = 0.0;
      if (live) {
        if (bc >= 0) {
          const double sc = ctl.atol + fabs(by) * ctl.rtol;
          const double u = (bk[1] - bk[0]) / sc;
          s2 = u * u;
        }
#pragma unroll
        for (int q = 0; q < G; ++q)
          if (vh[q])
#pragma unroll
            for (int c = 0; c < 6; ++c) {
              // position: f1 - f0 = dir (v0 + h0 kv0) - dir v0
              const double df =
                  c < 3 ? dir * stage_in<7>(y[q][3 + c], kv[q][c], h0) - dir * y[q][3 + c]
                        : kv[q][c - 3][1] - kv[q][c - 3][0];
              const double u = df / (ctl.atol + fabs(y[q][c]) * ctl.rtol);
              s2 = fma(u, u, s2);
            }
      }
      const double d2 = sqrt(reduce(s2) / ntot) / h0;
      const double md = fmax(d1, d2);
      const double h1 = md <= 1e-15 ? fmax(1e-6, h0 * 1e-3) : pow(0.01 / md, 0.2);
      dt0 = fmin(100.0 * h0, h1);
    }
    // the first step is clamped to [dtmin, dtmax] like every later one (a first step raised
    // to dtmin is accepted whatever its error)
    if (ctl.dtmax > 0.0) dt0 = fmin(dt0, ctl.dtmax);
    if (ctl.dtmin > 0.0) {
      at_dtmin = dt0 <= ctl.dtmin;
      dt0 = fmax(dt0, ctl.dtmin);
    }
    tnext = tau0 + dt0;
    if (tnext > tau1) tnext = tau1;
  }

  while (__syncthreads_or(live ? 1 : 0)) {
    const double h = tnext - tprev;
    stage(std::integral_constant<int, 1>{}, tprev, h);
    stage(std::integral_constant<int, 2>{}, tprev, h);
    stage(std::integral_constant<int, 3>{}, tprev, h);
    stage(std::integral_constant<int, 4>{}, tprev, h);
    stage(std::integral_constant<int, 5>{}, tprev, h);
    stage(std::integral_constant<int, 6>{}, tprev, h);
    double ss = 0.0;
    if (live) {
      if (bc >= 0) ss = err_of(by, bk, byn, h, ctl);
#pragma unroll
      for (int q = 0; q < G; ++q)
        if (vh[q])
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            ss += err_of(y[q][3 + c], kv[q][c], yn[q][3 + c], h, ctl);
            const double ex = dir * h * (h * qsum<8, 5>(kv[q][c]));
            const double u = ex / (ctl.atol + ctl.rtol * fmax(fabs(y[q][c]), fabs(yn[q][c])));
            ss = fma(u, u, ss);
          }
    }
    const double e = sqrt(reduce(ss) / ntot);
    if (live) {
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
      if (keep) {
        for (; js < nts && dir * ts[js] <= tnext; ++js) {
          const double tj = dir * ts[js];
          if (tj == tnext) {
            const int64_t row = (int64_t)i * nts + js;
            if (bc >= 0) ys[row * 6 + bc] = byn;
#pragma unroll
            for (int q = 0; q < G; ++q) {
              if (!vh[q]) continue;
              double* o = dys + (row * m + vj[q]) * DD + voff[q];
#pragma unroll
              for (int c = 0; c < 6; ++c) o[c] = yn[q][c];
            }
          } else {
            save_row(js, true, (tj - tprev) / h, h);
          }
        }
        by = byn;
        bk[0] = bk[6];
#pragma unroll
        for (int q = 0; q < G; ++q)
#pragma unroll
          for (int c = 0; c < 6; ++c) {
            y[q][c] = yn[q][c];
            if (c < 3) kv[q][c][0] = kv[q][c][6];
          }
        t_new = tnext;
        ++nacc;
      } else {
        ++nrej;
      }
      double tn1 = t_new + dtn;
      tprev = t_new < tau1 ? t_new : tau1;
      if (tn1 > tau1 - 1e-10) tn1 = keep ? tau1 : tprev + 0.5 * (tau1 - tprev);
      tnext = tn1;
      ++nst;
      live = tprev < tau1 && nst < ctl.max_steps;
    }
  }

  if (!valid) return;
  const bool failed = tprev < tau1;
  if (stats && lane == 0) {
    stats[3 * i] = nacc;
    stats[3 * i + 1] = nrej;
    stats[3 * i + 2] = failed ? 1 : 0;
  }
  if (failed) {
    const double inf = __builtin_inf();
    for (int j = 0; j < nts; ++j) fill_row(j, inf);
    if (bc >= 0 && w_out) w_out[6 * i + bc] = inf;
#pragma unroll
    for (int q = 0; q < G; ++q) {
      if (!vh[q] || !dw_out) continue;
      double* o = dw_out + ((int64_t)i * m + vj[q]) * DD + voff[q];
#pragma unroll
      for (int c = 0; c < 6; ++c) o[c] = inf;
    }
    return;
  }
  for (; js < nts; ++js) fill_row(js, __builtin_nan(""));
  if (bc >= 0 && w_out) w_out[6 * i + bc] = by;
#pragma unroll
  for (int q = 0; q < G; ++q) {
    if (!vh[q] || !dw_out) continue;
    double* o = dw_out + ((int64_t)i * m + vj[q]) * DD + voff[q];
#pragma unroll
    for (int c = 0; c < 6; ++c) o[c] = y[q][c];
  }
}

template <int SK, int RAD, int G>
int launch_pd(const ssc_host_potential& hp, const ssc_sources& pert, const double* w0,
              const double* dw0, const double* t_start, int64_t n, int L, double t0, double t1,
              const double* ts, int nts, const Dp5Ctl& ctl, double* w_out, double* dw_out,
              double* ys, double* dys, int32_t* stats, hipStream_t st) {
  const size_t lds = SourceLDS<SK>::bytes(pert);
  if (lds > 120 * 1024) return fail(SSC_E_LIMIT, "ssc_perturb_dopri5: source table exceeds LDS");
  if (lds > 60 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)perturb_dopri5_kernel<SK, RAD, G>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess)
      return fail(SSC_E_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  }
  const int ppb = PD_BLOCK / L;
  const int64_t blocks = (n + ppb - 1) / ppb;
  if (blocks > (int64_t)INT_MAX) return fail(SSC_E_LIMIT, "ssc_perturb_dopri5: too many orbits");
  hipLaunchKernelGGL((perturb_dopri5_kernel<SK, RAD, G>), dim3((unsigned)blocks), dim3(PD_BLOCK),
                     lds, st, hp, pert, w0, dw0, t_start, n, pert.n, L, t0, t1, ts, nts, ctl,
                     w_out, dw_out, ys, dys, stats);
  return check_launch("ssc_perturb_dopri5");
}

template <int SK, int RAD>
int dispatch_g(const ssc_host_potential& hp, const ssc_sources& s, const double* w0,
               const double* dw0, const double* t_start, int64_t n, double t0, double t1,
               const double* ts, int nts, const Dp5Ctl& ctl, double* w_out, double* dw_out,
               double* ys, double* dys, int32_t* stats, hipStream_t st) {
  const int nvec = RAD ? 2 * s.n : s.n;
  const int G = nvec <= PD_BLOCK ? 1 : (nvec <= 2 * PD_BLOCK ? 2 : 4);
  const int per = (nvec + G - 1) / G;
  int L = PD_MIN_L;
  while (L < per) L <<= 1;
  if (G == 1)
    return launch_pd<SK, RAD, 1>(hp, s, w0, dw0, t_start, n, L, t0, t1, ts, nts, ctl, w_out,
                                 dw_out, ys, dys, stats, st);
  if (G == 2)
    return launch_pd<SK, RAD, 2>(hp, s, w0, dw0, t_start, n, L, t0, t1, ts, nts, ctl, w_out,
                                 dw_out, ys, dys, stats, st);
  return launch_pd<SK, RAD, 4>(hp, s, w0, dw0, t_start, n, L, t0, t1, ts, nts, ctl, w_out,
                               dw_out, ys, dys, stats, st);
}

}  // namespace
}  // namespace ssc

using namespace ssc;

extern "C" int ssc_perturb_dopri5(const ssc_host_potential* host, const ssc_sources* pert,
                                  const double* w0, const double* dw0, const double* t_start,
                                  int64_t n, double t0, double t1, const double* ts, int32_t nts,
                                  int32_t with_radius, const ssc_dopri5_options* opt,
                                  double* w_out, double* dw_out, double* ys, double* dys,
                                  int32_t* stats, void* stream) {
  int rc = validate_host(host);
  if (rc) return rc;
  if (!opt) return fail(SSC_E_ARG, "ssc_perturb_dopri5: options are NULL");
  if (!pert || pert->kind == SSC_SRC_NONE || pert->n <= 0)
    return fail(SSC_E_ARG, "ssc_perturb_dopri5: needs at least one perturbing source");
  if (n < 0 || nts < 0) return fail(SSC_E_ARG, "ssc_perturb_dopri5: negative size");
  if (opt->n_stops != 0)
    return fail(SSC_E_ARG, "ssc_perturb_dopri5: jump times are not supported");
  if (!(opt->rtol >= 0.0) || !(opt->atol >= 0.0) || !(opt->rtol + opt->atol > 0.0))
    return fail(SSC_E_ARG, "ssc_perturb_dopri5: rtol/atol must be >= 0 and not both 0");
  if (opt->max_steps < 1) return fail(SSC_E_ARG, "ssc_perturb_dopri5: max_steps < 1");
  rc = validate_sources(pert, 0, true);
  if (rc) return rc;
  const ssc_sources s = *pert;
  if (s.group_size > 0)
    return fail(SSC_E_ARG, "ssc_perturb_dopri5: grouped sources are not supported");
  const int nvec = with_radius ? 2 * s.n : s.n;
  if (nvec > 4 * PD_BLOCK)
    return fail(SSC_E_LIMIT, "ssc_perturb_dopri5: at most 1024 response vectors per particle "
                             "(1024 subhalos, or 512 with the radius block)");
  if (n == 0) return SSC_OK;
  if (!w0) return fail(SSC_E_ARG, "ssc_perturb_dopri5: w0 is NULL");
  if (nts > 0 && (!ts || !ys || !dys))
    return fail(SSC_E_ARG, "ssc_perturb_dopri5: nts > 0 needs ts, ys and dys");
  if (nts == 0 && (!w_out || !dw_out))
    return fail(SSC_E_ARG, "ssc_perturb_dopri5: no output (w_out and dw_out, or ys and dys)");
  if (nts == 0) ys = dys = nullptr;
  Dp5Ctl ctl{opt->rtol, opt->atol, opt->dtmin, opt->dtmax, opt->dt0, opt->max_steps, 0};
  hipStream_t st = (hipStream_t)stream;
#define SSC_PD(K, R)                                                                          \
  dispatch_g<K, R>(*host, s, w0, dw0, t_start, n, t0, t1, ts, nts, ctl, w_out, dw_out, ys, dys, \
                   stats, st)
  switch (s.kind) {
    case SSC_SRC_STATIC: return with_radius ? SSC_PD(SSC_SRC_STATIC, 1) : SSC_PD(SSC_SRC_STATIC, 0);
    case SSC_SRC_LINE: return with_radius ? SSC_PD(SSC_SRC_LINE, 1) : SSC_PD(SSC_SRC_LINE, 0);
    case SSC_SRC_LINE_HERNQUIST:
      return with_radius ? SSC_PD(SSC_SRC_LINE_HERNQUIST, 1) : SSC_PD(SSC_SRC_LINE_HERNQUIST, 0);
    case SSC_SRC_TRACK:
    case SSC_SRC_TRACK_HERNQUIST:
      return fail(SSC_E_ARG, "ssc_perturb_dopri5: TRACK sources are sampled on a fixed grid; "
                             "adaptive steps need STATIC or LINE sources");
    default: return fail(SSC_E_ARG, "ssc_perturb_dopri5: bad source kind");
  }
#undef SSC_PD
}
