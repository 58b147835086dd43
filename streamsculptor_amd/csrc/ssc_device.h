// Device-side building blocks for the gfx950 stream-particle kernels: analytic host
// components (gradient and Hessian), Plummer source terms, and the launch/error plumbing
// shared by every translation unit of libssc_mi355x.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/ssc_mi355x.h"

namespace ssc {

// ------------------------------------------------------------------------------------------
// error plumbing (host)
// ------------------------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_launch(const char* what);
int validate_host(const ssc_host_potential* host);
int validate_sources(const ssc_sources* src, int32_t nsteps_needed, bool allow_groups = false);

// ------------------------------------------------------------------------------------------
// small vector helpers
// ------------------------------------------------------------------------------------------
struct d3 {
  double x, y, z;
};

// 1/sqrt(v) for positive finite v: the hardware estimate refined by one Newton step
// y (1 + e/2), e = 1 - v y^2, which doubles the estimate's correct bits.
// Unlike the library rsqrt it skips the zero/inf/denormal class handling, which no squared
// radius on this path needs.
__device__ __forceinline__ double rsq(double v) {
  const double y = __builtin_amdgcn_rsq(v);
  const double e = fma(-v * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

// 1/b for positive finite b: the hardware estimate refined by two Newton steps (about one unit
// in the last place; five instructions against the eleven of an IEEE division).
__device__ __forceinline__ double recip(double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = fma(r, fma(-b, r, 1.0), r);
  return fma(r, fma(-b, r, 1.0), r);
}

// log(1 + u) for u >= 0 (the NFW profile's log(1 + r / r_s)), within about two units in the
// last place.  The library's log1p evaluates in double-double arithmetic, about 130
// instructions a call on gfx950 -- half of a host-force evaluation -- and on a wave64 every
// FP64 instruction holds the SIMD for four cycles however few lanes are active, so the
// one-orbit chains (progenitors, subhalo orbits) pay for every one of them.  Method: t = 1 + u
// rounded and its rounding error c = u - (t - 1) (exact for t <= 2^53); t = m 2^e with
// m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m - 1) / (m + 1), |s| <= 0.1716, by its
// Taylor series 2 s (1 + z/3 + z^2/5 + ...), z = s^2, up to z^10 (the first omitted term is
// below 2e-18 relative); log(1 + u) = e ln 2 + log m + c / t, with ln 2 split so that e ln2_hi
// is exact.  About 37 instructions.
__device__ __forceinline__ double log1p_nonneg(double u) {
  const double t = 1.0 + u;
  const double c = u - (t - 1.0);
  int e = __builtin_amdgcn_frexp_exp(t);  // t = mant 2^e, mant in [1/2, 1)
  double m = __builtin_amdgcn_frexp_mant(t);
  if (m < 0.70710678118654752) {
    m *= 2.0;
    e -= 1;
  }
  const double num = m - 1.0;  // exact
  const double den = m + 1.0;
  double r = __builtin_amdgcn_rcp(den);
  r = fma(r, fma(-den, r, 1.0), r);
  double s = num * r;
  s = fma(r, fma(-den, s, num), s);
  const double z = s * s;
  double p = 0.09523809523809523;  // 2/21
  p = fma(p, z, 0.10526315789473684);
  p = fma(p, z, 0.11764705882352941);
  p = fma(p, z, 0.13333333333333333);
  p = fma(p, z, 0.15384615384615385);
  p = fma(p, z, 0.18181818181818182);
  p = fma(p, z, 0.2222222222222222);
  p = fma(p, z, 0.2857142857142857);
  p = fma(p, z, 0.4);
  p = fma(p, z, 0.6666666666666666);  // 2/3
  const double lm = fma(s * z, p, 2.0 * s);
  const double fe = (double)e;
  const double ln2_hi = 0.6931467056274414;       // 21 significant bits
  const double ln2_lo = 4.7493250390316726e-07;   // ln 2 - ln2_hi
  return fma(fe, ln2_hi, fma(fe, ln2_lo, lm) + c * __builtin_amdgcn_rcp(t));
}

// ------------------------------------------------------------------------------------------
// Sum of a double over aligned groups of G consecutive lanes (G = 1, 2, 4, ..., 64), the result
// in every lane of the group.  Within a 16-lane row the exchanges are DPP moves (quad
// permutations, then the half-row and row mirrors), which need no trip through the LDS unit;
// the 32- and 64-lane steps use lane shuffles.  Every lane of the wave must be active.
// ------------------------------------------------------------------------------------------
// Lanes per orbit for the grouped kernels: enough to give the launch about three waves per SIMD
// (3 x 1024 SIMDs x 64 lanes), at most 16, and at least four sources per lane.  (1e5 orbits in
// 64 live subhalos: two lanes per orbit 18.4 ms, four 19.0 ms, eight 25.9 ms per stream.)
inline int group_lanes(int64_t n, int nsrc) {
  if (nsrc < 16 || n <= 0) return 1;
  const double want = 196608.0 / (double)n;
  int g = 1;
  while (g < want && g < 16) g *= 2;
  while (g > 1 && nsrc / g < 4) g /= 2;
  return g;
}

template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// v[lane] + v[lane ^ 16] (W = 16) or v[lane] + v[lane ^ 32] (W = 32) with the gfx950 permlane
// swaps, VALU instructions in place of two ds_bpermute round trips through the LDS pipe.  With
// both operands v, a swap exchanges the matching lanes of paired 16-lane rows (W = 16) or of the
// two 32-lane halves (W = 32), so every lane ends with {v[lane], v[lane ^ W]} in some order;
// IEEE addition commutes, so the sum equals v + __shfl_xor(v, W) bit for bit.  Every lane of
// the wave must be active.
template <int W>
__device__ __forceinline__ double swap_sum(double v) {
  const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
  if constexpr (W == 16) {
    const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
  } else {
    static_assert(W == 32, "swap_sum exchanges lane ^ 16 or lane ^ 32");
    const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
  }
}

// Sum over aligned groups of G lanes; G = 32 and 64 need every lane of the wave active.
template <int G>
__device__ __forceinline__ double group_sum(double v) {
  if (G >= 2) v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
  if (G >= 4) v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
  if (G >= 8) v += dpp_move<0x141>(v);  // row_half_mirror: the other quad of the 8-lane half
  if (G >= 16) v += dpp_move<0x140>(v); // row_mirror: the other half of the 16-lane row
  if (G >= 32) v = swap_sum<16>(v);
  if (G >= 64) v = swap_sum<32>(v);
  return v;
}

// ------------------------------------------------------------------------------------------
// Host components, gradient only (the per-step force).  Formulas: reference
// streamsculptor/potential.py:66-84 (MN, NFW), :124-138 (Plummer, Hernquist), differentiated
// in closed form.
// ------------------------------------------------------------------------------------------
// Gradient of one host component at (x, y, z) in parts: grad = (f x, f y, zm z), or
// (f x, f y, zm) when zadd (the Miyamoto-Nagai z component, formed whole).  comp_grad_add
// accumulates the parts with the
// same operations everywhere (an FMA per coordinate, an add for the Miyamoto-Nagai z term), so
// a sum formed on one lane and a sum of parts gathered from several waves are bit-identical.
// LEAN (the fixed-step kernels): the refined hardware reciprocal square root and reciprocal
// (rsq, recip) where the library would expand an IEEE square root or division, and
// log1p_nonneg -- about one unit in the last place each, and half the instructions of a host
// force (241 -> 115 per step of the one-lane leapfrog with MN + Hernquist + NFW).  The
// adaptive kernels keep the IEEE forms (LEAN = false): their step sequences follow the last
// bits of every force, and parity with the reference's step sequence is stated there.
template <bool LEAN = true>
__device__ __forceinline__ void parts_mn(const ssc_component& c, double x, double y, double z,
                                         double& f, double& zm) {
  const double gm = c.p[0], a = c.p[1], b = c.p[2];
  if constexpr (LEAN) {
    const double v = z * z + b * b;
    const double iz = rsq(v);  // 1 / zeta
    const double zeta = v * iz;
    const double S = a + zeta;
    const double D2 = x * x + y * y + S * S;
    const double iD = rsq(D2);
    f = gm * iD * iD * iD;
    zm = f * z * S * iz;
  } else {
    const double zeta = sqrt(z * z + b * b);
    const double S = a + zeta;
    const double D2 = x * x + y * y + S * S;
    const double iD = rsq(D2);
    f = gm * iD * iD * iD;
    zm = f * z * S / zeta;
  }
}
template <bool LEAN = true>
__device__ __forceinline__ double parts_nfw(const ssc_component& c, double x, double y,
                                            double z) {
  const double gm = c.p[0], rs = c.p[1];
  const double r2 = x * x + y * y + z * z;
  const double ir = rsq(r2);
  const double r = r2 * ir;
  if constexpr (LEAN) {
    const double L = log1p_nonneg(r * recip(rs));
    return gm * ir * ir * (L * ir - recip(r + rs));
  } else {
    const double L = log1p(r / rs);
    return gm * ir * ir * (L * ir - 1.0 / (r + rs));
  }
}
template <bool LEAN = true>
__device__ __forceinline__ double parts_hernquist(const ssc_component& c, double x, double y,
                                                  double z) {
  const double gm = c.p[0], cc = c.p[1], soft = c.p[2];
  const double r2 = x * x + y * y + z * z + soft;
  const double ir = rsq(r2);
  const double r = r2 * ir;
  const double rc = r + cc;
  if constexpr (LEAN) return gm * ir * recip(rc * rc);
  return gm * ir / (rc * rc);
}
__device__ __forceinline__ double parts_plummer(const ssc_component& c, double x, double y,
                                                double z) {
  const double gm = c.p[0], rs = c.p[1];
  const double s2 = x * x + y * y + z * z + rs * rs;
  const double is = rsq(s2);
  return gm * is * is * is;
}

template <bool LEAN = true>
__device__ __forceinline__ void comp_grad_parts(const ssc_component& c, double x, double y,
                                                double z, double& f, double& zm, bool& zadd) {
  f = 0.0;
  zm = 0.0;
  zadd = false;
  switch (c.kind) {
    case SSC_COMP_MIYAMOTO_NAGAI:
      parts_mn<LEAN>(c, x, y, z, f, zm);
      zadd = true;
      break;
    case SSC_COMP_NFW:
      f = zm = parts_nfw<LEAN>(c, x, y, z);
      break;
    case SSC_COMP_HERNQUIST:
      f = zm = parts_hernquist<LEAN>(c, x, y, z);
      break;
    case SSC_COMP_PLUMMER:
      f = zm = parts_plummer(c, x, y, z);
      break;
    default:
      break;
  }
}

__device__ __forceinline__ void comp_grad_add(double f, double zm, bool zadd, double x, double y,
                                              double z, double& gx, double& gy, double& gz) {
  gx = fma(f, x, gx);
  gy = fma(f, y, gy);
  gz = zadd ? gz + zm : fma(zm, z, gz);
}

template <bool LEAN = true>
__device__ __forceinline__ void comp_grad(const ssc_component& c, double x, double y, double z,
                                          double& gx, double& gy, double& gz) {
  double f, zm;
  bool zadd;
  comp_grad_parts<LEAN>(c, x, y, z, f, zm, zadd);
  comp_grad_add(f, zm, zadd, x, y, z, gx, gy, gz);
}

template <bool LEAN = true>
__device__ __forceinline__ void host_grad(const ssc_host_potential& hp, double x, double y,
                                          double z, double& gx, double& gy, double& gz) {
  gx = 0.0;
  gy = 0.0;
  gz = 0.0;
  for (int c = 0; c < hp.n_components; ++c) comp_grad<LEAN>(hp.c[c], x, y, z, gx, gy, gz);
}

// Host signatures with straight-line gradient code.  The generic gradient walks the components
// with a branch on each kind, so a wave finishes one component's dependent chain before it
// issues the next; for the two hosts used throughout (MN + Hernquist + NFW, and the Milky Way
// model MN + Hernquist + Hernquist + NFW) the kinds are known at compile time and the compiler
// interleaves the independent chains.  The parts are accumulated as comp_grad_add does.
constexpr int HS_GENERIC = 0, HS_MN_HQ_NFW = 1, HS_MN_HQ_HQ_NFW = 2;

inline int host_signature(const ssc_host_potential& hp) {
  const int n = hp.n_components;
  if (n == 3 && hp.c[0].kind == SSC_COMP_MIYAMOTO_NAGAI && hp.c[1].kind == SSC_COMP_HERNQUIST &&
      hp.c[2].kind == SSC_COMP_NFW)
    return HS_MN_HQ_NFW;
  if (n == 4 && hp.c[0].kind == SSC_COMP_MIYAMOTO_NAGAI && hp.c[1].kind == SSC_COMP_HERNQUIST &&
      hp.c[2].kind == SSC_COMP_HERNQUIST && hp.c[3].kind == SSC_COMP_NFW)
    return HS_MN_HQ_HQ_NFW;
  return HS_GENERIC;
}

// The fixed-step kernels take the LEAN host forms except with STATIC sources (the clumpy
// halo): there the 512-term source sum is nearly all of a step, so the host forms' cost does
// not matter, and the IEEE forms keep the plain orbits -- and with them the surrogate's
// training positions and its measured error -- as they were measured (that error moves with
// the last bits of the training data: one training seed went from 2.8 to 19.4 pc when the
// plain orbits changed only there).
constexpr bool lean_host(int SK) { return SK != SSC_SRC_STATIC; }

template <int HS, bool LEAN = true>
__device__ __forceinline__ void host_grad_sig(const ssc_host_potential& hp, double x, double y,
                                              double z, double& gx, double& gy, double& gz) {
  if constexpr (HS == HS_GENERIC) {
    host_grad<LEAN>(hp, x, y, z, gx, gy, gz);
  } else {
    double fm, zm;
    parts_mn<LEAN>(hp.c[0], x, y, z, fm, zm);
    const double f1 = parts_hernquist<LEAN>(hp.c[1], x, y, z);
    const double f2 = HS == HS_MN_HQ_HQ_NFW ? parts_hernquist<LEAN>(hp.c[2], x, y, z) : 0.0;
    const double fn = parts_nfw<LEAN>(hp.c[HS == HS_MN_HQ_HQ_NFW ? 3 : 2], x, y, z);
    gx = gy = gz = 0.0;
    comp_grad_add(fm, zm, true, x, y, z, gx, gy, gz);
    comp_grad_add(f1, f1, false, x, y, z, gx, gy, gz);
    if (HS == HS_MN_HQ_HQ_NFW) comp_grad_add(f2, f2, false, x, y, z, gx, gy, gz);
    comp_grad_add(fn, fn, false, x, y, z, gx, gy, gz);
  }
}

// Spherical Hessian accumulation: H += (phi'/r) I + (phi'' - phi'/r) x x^T / r^2.
__device__ __forceinline__ void sph_hess(double x, double y, double z, double d1r, double c,
                                         double* H) {
  H[0] += d1r + c * x * x;
  H[1] += c * x * y;
  H[2] += c * x * z;
  H[4] += d1r + c * y * y;
  H[5] += c * y * z;
  H[8] += d1r + c * z * z;
}

// Potential, gradient and Hessian of one component (H upper triangle in H[0,1,2,4,5,8]).
// Potential, gradient and Hessian (upper triangle) of one component, accumulated; with dp, also
// d(grad Phi_c)/d(p1) (as comp_grad_dp1) from the same intermediate terms.
__device__ __forceinline__ void comp_full(const ssc_component& c, double x, double y, double z,
                                          double& phi, double* g, double* H,
                                          double* dp = nullptr) {
  const double gm = c.p[0];
  switch (c.kind) {
    case SSC_COMP_MIYAMOTO_NAGAI: {
      const double a = c.p[1], b = c.p[2];
      const double zeta = sqrt(z * z + b * b);
      const double S = a + zeta;
      const double D2 = x * x + y * y + S * S;
      const double D = sqrt(D2);
      const double iD3 = 1.0 / (D2 * D);
      const double iD5 = iD3 / D2;
      const double dz = z * S / zeta;
      phi -= gm / D;
      g[0] += gm * x * iD3;
      g[1] += gm * y * iD3;
      g[2] += gm * dz * iD3;
      const double k = -3.0 * gm * iD5;
      const double ddz = S * b * b / (zeta * zeta * zeta) + z * z / (zeta * zeta);
      H[0] += k * x * x + gm * iD3;
      H[1] += k * x * y;
      H[2] += k * x * dz;
      H[4] += k * y * y + gm * iD3;
      H[5] += k * y * dz;
      H[8] += k * dz * dz + gm * ddz * iD3;
      if (dp) {
        dp[0] = -3.0 * gm * x * S * iD5;
        dp[1] = -3.0 * gm * y * S * iD5;
        dp[2] = gm * (z / zeta) * iD3 - 3.0 * gm * (z * S / zeta) * S * iD5;
      }
      break;
    }
    case SSC_COMP_NFW: {
      const double rs = c.p[1];
      const double r = sqrt(x * x + y * y + z * z);
      const double L = log1p(r / rs);
      const double rp = r + rs;
      if (dp) {
        const double f = gm * (-1.0 / (r * rs * rp) + 1.0 / (r * rp * rp)) / r;
        dp[0] = f * x; dp[1] = f * y; dp[2] = f * z;
      }
      phi -= gm * L / r;
      const double d1 = gm * (L / (r * r) - 1.0 / (r * rp));
      const double d2 = gm * (1.0 / (r * r * rp) - 2.0 * L / (r * r * r) +
                              (2.0 * r + rs) / (r * r * rp * rp));
      const double d1r = d1 / r;
      g[0] += d1r * x;
      g[1] += d1r * y;
      g[2] += d1r * z;
      sph_hess(x, y, z, d1r, (d2 - d1r) / (r * r), H);
      break;
    }
    case SSC_COMP_HERNQUIST: {
      const double cc = c.p[1], soft = c.p[2];
      const double r = sqrt(x * x + y * y + z * z + soft);
      const double rc = r + cc;
      if (dp) {
        const double f = -2.0 * gm / (r * rc * rc * rc);
        dp[0] = f * x; dp[1] = f * y; dp[2] = f * z;
      }
      phi -= gm / rc;
      const double d1 = gm / (rc * rc);
      const double d2 = -2.0 * gm / (rc * rc * rc);
      const double d1r = d1 / r;
      g[0] += d1r * x;
      g[1] += d1r * y;
      g[2] += d1r * z;
      sph_hess(x, y, z, d1r, (d2 - d1r) / (r * r), H);
      break;
    }
    case SSC_COMP_PLUMMER: {
      const double rs = c.p[1];
      const double s2 = x * x + y * y + z * z + rs * rs;
      const double s = sqrt(s2);
      const double i3 = 1.0 / (s2 * s);
      if (dp) {
        const double f = -3.0 * gm * rs / (s2 * s2 * s);
        dp[0] = f * x; dp[1] = f * y; dp[2] = f * z;
      }
      phi -= gm / s;
      const double f = gm * i3;
      g[0] += f * x;
      g[1] += f * y;
      g[2] += f * z;
      const double k = -3.0 * f / s2;
      H[0] += f + k * x * x;
      H[1] += k * x * y;
      H[2] += k * x * z;
      H[4] += f + k * y * y;
      H[5] += k * y * z;
      H[8] += f + k * z * z;
      break;
    }
    default:
      if (dp) dp[0] = dp[1] = dp[2] = 0.0;
      break;
  }
}

__device__ __forceinline__ void host_full(const ssc_host_potential& hp, double x, double y,
                                          double z, double& phi, double* g, double* H) {
  phi = 0.0;
  g[0] = g[1] = g[2] = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) H[i] = 0.0;
  for (int c = 0; c < hp.n_components; ++c) comp_full(hp.c[c], x, y, z, phi, g, H);
  H[3] = H[1];
  H[6] = H[2];
  H[7] = H[5];
}

// comp_full in the lean forms (see parts_mn): every division and square root of comp_full
// becomes a product of the refined reciprocal square root of the squared radius and of one
// refined reciprocal (rsq, recip), and log1p becomes log1p_nonneg.  Same outputs.
__device__ __forceinline__ void comp_full_lean(const ssc_component& c, double x, double y,
                                               double z, double& phi, double* g, double* H,
                                               double* dp = nullptr) {
  const double gm = c.p[0];
  switch (c.kind) {
    case SSC_COMP_MIYAMOTO_NAGAI: {
      const double a = c.p[1], b = c.p[2];
      const double v = z * z + b * b;
      const double iz = rsq(v);  // 1 / zeta
      const double zeta = v * iz;
      const double S = a + zeta;
      const double D2 = x * x + y * y + S * S;
      const double iD = rsq(D2);
      const double iD2 = iD * iD;
      const double iD3 = iD2 * iD;
      const double iD5 = iD3 * iD2;
      const double dz = z * S * iz;
      phi -= gm * iD;
      g[0] += gm * x * iD3;
      g[1] += gm * y * iD3;
      g[2] += gm * dz * iD3;
      const double k = -3.0 * gm * iD5;
      const double iz2 = iz * iz;
      const double ddz = S * b * b * (iz2 * iz) + z * z * iz2;
      H[0] += k * x * x + gm * iD3;
      H[1] += k * x * y;
      H[2] += k * x * dz;
      H[4] += k * y * y + gm * iD3;
      H[5] += k * y * dz;
      H[8] += k * dz * dz + gm * ddz * iD3;
      if (dp) {
        dp[0] = -3.0 * gm * x * S * iD5;
        dp[1] = -3.0 * gm * y * S * iD5;
        dp[2] = gm * (z * iz) * iD3 - 3.0 * gm * dz * S * iD5;
      }
      break;
    }
    case SSC_COMP_NFW: {
      const double rs = c.p[1];
      const double r2 = x * x + y * y + z * z;
      const double ir = rsq(r2);
      const double r = r2 * ir;
      const double ir2 = ir * ir;
      const double irs = recip(rs);
      const double rp = r + rs;
      const double irp = recip(rp);
      const double L = log1p_nonneg(r * irs);
      if (dp) {
        const double f = gm * (ir * irp) * (irp - irs) * ir;
        dp[0] = f * x; dp[1] = f * y; dp[2] = f * z;
      }
      phi -= gm * L * ir;
      const double d1 = gm * (L * ir2 - ir * irp);
      const double d2 = gm * (ir2 * irp - 2.0 * L * ir2 * ir + (2.0 * r + rs) * ir2 * irp * irp);
      const double d1r = d1 * ir;
      g[0] += d1r * x;
      g[1] += d1r * y;
      g[2] += d1r * z;
      sph_hess(x, y, z, d1r, (d2 - d1r) * ir2, H);
      break;
    }
    case SSC_COMP_HERNQUIST: {
      const double cc = c.p[1], soft = c.p[2];
      const double r2 = x * x + y * y + z * z + soft;
      const double ir = rsq(r2);
      const double r = r2 * ir;
      const double irc = recip(r + cc);
      const double irc2 = irc * irc;
      if (dp) {
        const double f = -2.0 * gm * ir * irc2 * irc;
        dp[0] = f * x; dp[1] = f * y; dp[2] = f * z;
      }
      phi -= gm * irc;
      const double d1 = gm * irc2;
      const double d2 = -2.0 * gm * irc2 * irc;
      const double d1r = d1 * ir;
      g[0] += d1r * x;
      g[1] += d1r * y;
      g[2] += d1r * z;
      sph_hess(x, y, z, d1r, (d2 - d1r) * (ir * ir), H);
      break;
    }
    case SSC_COMP_PLUMMER: {
      const double rs = c.p[1];
      const double s2 = x * x + y * y + z * z + rs * rs;
      const double is = rsq(s2);
      const double is2 = is * is;
      const double i3 = is2 * is;
      if (dp) {
        const double f = -3.0 * gm * rs * (i3 * is2);
        dp[0] = f * x; dp[1] = f * y; dp[2] = f * z;
      }
      phi -= gm * is;
      const double f = gm * i3;
      g[0] += f * x;
      g[1] += f * y;
      g[2] += f * z;
      const double k = -3.0 * f * is2;
      H[0] += f + k * x * x;
      H[1] += k * x * y;
      H[2] += k * x * z;
      H[4] += f + k * y * y;
      H[5] += k * y * z;
      H[8] += f + k * z * z;
      break;
    }
    default:
      if (dp) dp[0] = dp[1] = dp[2] = 0.0;
      break;
  }
}

// comp_full with the component's kind fixed at compile time (K >= 0: the kind switch folds
// away; K < 0: the kind is read from c as in comp_full); LEAN selects comp_full_lean.
template <int K, bool LEAN = false>
__device__ __forceinline__ void comp_full_as(const ssc_component& c, double x, double y, double z,
                                             double& phi, double* g, double* H,
                                             double* dp = nullptr) {
  ssc_component k = c;
  if constexpr (K >= 0) k.kind = K;
  if constexpr (LEAN)
    comp_full_lean(k, x, y, z, phi, g, H, dp);
  else
    comp_full(k, x, y, z, phi, g, H, dp);
}

// Number of components and kind of component C of a known host signature.
constexpr int hs_components(int HS) { return HS == HS_MN_HQ_HQ_NFW ? 4 : 3; }
constexpr int hs_kind(int HS, int C) {
  return C == 0 ? SSC_COMP_MIYAMOTO_NAGAI
                : (C == hs_components(HS) - 1 ? SSC_COMP_NFW : SSC_COMP_HERNQUIST);
}

// host_full as straight-line code for a known signature (same accumulation order).
template <int HS, bool LEAN = false>
__device__ __forceinline__ void host_full_sig(const ssc_host_potential& hp, double x, double y,
                                              double z, double& phi, double* g, double* H) {
  if constexpr (HS == HS_GENERIC && !LEAN) {
    host_full(hp, x, y, z, phi, g, H);
  } else if constexpr (HS == HS_GENERIC) {
    phi = 0.0;
    g[0] = g[1] = g[2] = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) H[i] = 0.0;
    for (int c = 0; c < hp.n_components; ++c) comp_full_lean(hp.c[c], x, y, z, phi, g, H);
    H[3] = H[1];
    H[6] = H[2];
    H[7] = H[5];
  } else {
    phi = 0.0;
    g[0] = g[1] = g[2] = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) H[i] = 0.0;
    comp_full_as<hs_kind(HS, 0), LEAN>(hp.c[0], x, y, z, phi, g, H);
    comp_full_as<hs_kind(HS, 1), LEAN>(hp.c[1], x, y, z, phi, g, H);
    comp_full_as<hs_kind(HS, 2), LEAN>(hp.c[2], x, y, z, phi, g, H);
    if constexpr (hs_components(HS) > 3)
      comp_full_as<hs_kind(HS, 3), LEAN>(hp.c[3], x, y, z, phi, g, H);
    H[3] = H[1];
    H[6] = H[2];
    H[7] = H[5];
  }
}

// d(grad Phi_c)/d(p1) for one component (MN a, NFW r_s, Hernquist c, Plummer r_s).
__device__ __forceinline__ void comp_grad_dp1(const ssc_component& c, double x, double y, double z,
                                              double* out) {
  const double gm = c.p[0];
  out[0] = out[1] = out[2] = 0.0;
  switch (c.kind) {
    case SSC_COMP_MIYAMOTO_NAGAI: {
      const double a = c.p[1], b = c.p[2];
      const double zeta = sqrt(z * z + b * b), S = a + zeta;
      const double D2 = x * x + y * y + S * S, D = sqrt(D2);
      const double iD3 = 1.0 / (D2 * D), iD5 = iD3 / D2;
      out[0] = -3.0 * gm * x * S * iD5;
      out[1] = -3.0 * gm * y * S * iD5;
      out[2] = gm * (z / zeta) * iD3 - 3.0 * gm * (z * S / zeta) * S * iD5;
      break;
    }
    case SSC_COMP_NFW: {
      const double rs = c.p[1];
      const double r = sqrt(x * x + y * y + z * z), rp = r + rs;
      const double f = gm * (-1.0 / (r * rs * rp) + 1.0 / (r * rp * rp)) / r;
      out[0] = f * x; out[1] = f * y; out[2] = f * z;
      break;
    }
    case SSC_COMP_HERNQUIST: {
      const double cc = c.p[1], soft = c.p[2];
      const double r = sqrt(x * x + y * y + z * z + soft), rc = r + cc;
      const double f = -2.0 * gm / (r * rc * rc * rc);
      out[0] = f * x; out[1] = f * y; out[2] = f * z;
      break;
    }
    case SSC_COMP_PLUMMER: {
      const double rs = c.p[1];
      const double s2 = x * x + y * y + z * z + rs * rs, s = sqrt(s2);
      const double f = -3.0 * gm * rs / (s2 * s2 * s);
      out[0] = f * x; out[1] = f * y; out[2] = f * z;
      break;
    }
    default:
      break;
  }
}

// dst[c] = sum_i src[i * stride + c] for c = blockIdx.x (launch one block of BLOCK threads per
// column): strided partial sums, then a tree in LDS in a fixed order.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void column_sum_kernel(const double* __restrict__ src,
                                                           int64_t rows, int stride,
                                                           double* __restrict__ dst) {
  __shared__ double part[BLOCK];
  const int c = blockIdx.x;
  double a = 0.0;
  for (int64_t i = threadIdx.x; i < rows; i += BLOCK) a += src[i * stride + c];
  part[threadIdx.x] = a;
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) dst[c] = part[0];
}

// Plummer source (x - s): gradient contribution, the N x M hot loop body.
// G m / s^3 from the hardware estimate y of 1/sqrt(s2) with the correction folded into the
// cube: (1 - e)^(-3/2) = 1 + 3e/2 + O(e^2), e = 1 - s2 y^2, one instruction fewer than a
// refined 1/sqrt followed by the cube (relative error 15 e^2 / 8, against 9 e^2 / 8 there).
__device__ __forceinline__ double gm_inv_s3(double gm, double s2) {
  const double y = __builtin_amdgcn_rsq(s2);
  const double y2 = y * y;
  const double t = fma(s2 * y2, -1.5, 2.5);  // 1 + 3e/2
  return gm * y * y2 * t;
}

// plummer_grad with G m given as (c1, c2) = (-1.5 G m, 2.5 G m), so that the product by G m
// folds into the correction of gm_inv_s3: G m / s^3 = y^3 fma(s^2 y^2, c1, c2).
__device__ __forceinline__ void plummer_grad_k(double dx, double dy, double dz, double rs2,
                                               double c1, double c2, double& gx, double& gy,
                                               double& gz) {
  const double s2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, rs2)));
  const double y = __builtin_amdgcn_rsq(s2);
  const double y2 = y * y;
  const double f = y * y2 * fma(s2 * y2, c1, c2);
  gx += f * dx;
  gy += f * dy;
  gz += f * dz;
}

__device__ __forceinline__ void plummer_grad(double dx, double dy, double dz, double gm,
                                             double rs2, double& gx, double& gy, double& gz) {
  const double s2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, rs2)));
  const double f = gm_inv_s3(gm, s2);
  gx += f * dx;
  gy += f * dy;
  gz += f * dz;
}

// Plummer source: potential, gradient and full Hessian contribution.
__device__ __forceinline__ void plummer_full(double dx, double dy, double dz, double gm,
                                             double rs2, double& phi, double* g, double* H) {
  const double s2 = dx * dx + dy * dy + dz * dz + rs2;
  const double is = rsq(s2);
  const double f = gm * is * is * is;
  phi -= gm * is;
  g[0] += f * dx;
  g[1] += f * dy;
  g[2] += f * dz;
  const double k = -3.0 * f * is * is;
  H[0] += f + k * dx * dx;
  H[1] += k * dx * dy;
  H[2] += k * dx * dz;
  H[3] += k * dx * dy;
  H[4] += f + k * dy * dy;
  H[5] += k * dy * dz;
  H[6] += k * dx * dz;
  H[7] += k * dy * dz;
  H[8] += f + k * dz * dz;
}

// Hernquist source (x - s), Phi = -G m / (r + a): gradient contribution.
__device__ __forceinline__ void hernquist_grad(double dx, double dy, double dz, double gm,
                                               double a, double& gx, double& gy, double& gz) {
  const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
  const double ir = rsq(r2);
  const double rc = fma(r2, ir, a);
  const double f = gm * ir / (rc * rc);
  gx += f * dx;
  gy += f * dy;
  gz += f * dz;
}

// Hernquist source: potential, gradient and full Hessian contribution
// (H = f I + k d d^T, f = G m / (r (r + a)^2), k = -f (3 r + a) / (r^2 (r + a))).
__device__ __forceinline__ void hernquist_full(double dx, double dy, double dz, double gm,
                                               double a, double& phi, double* g, double* H) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double r = sqrt(r2);
  const double rc = r + a;
  const double f = gm / (r * rc * rc);
  phi -= gm / rc;
  g[0] += f * dx;
  g[1] += f * dy;
  g[2] += f * dz;
  const double k = -f * (3.0 * r + a) / (r2 * rc);
  H[0] += f + k * dx * dx;
  H[1] += k * dx * dy;
  H[2] += k * dx * dz;
  H[3] += k * dx * dy;
  H[4] += f + k * dy * dy;
  H[5] += k * dy * dz;
  H[6] += k * dx * dz;
  H[7] += k * dy * dz;
  H[8] += f + k * dz * dz;
}

// Position of source j at grid row `step` / time t; false when inactive (LINE window,
// reference streamsculptor/potential.py:826).
__device__ __forceinline__ bool source_position(const ssc_sources& s, int j, int step, double t,
                                                double& sx, double& sy, double& sz) {
  if (s.kind == SSC_SRC_STATIC) {
    sx = s.pos[3 * j];
    sy = s.pos[3 * j + 1];
    sz = s.pos[3 * j + 2];
    return true;
  }
  if (s.kind == SSC_SRC_TRACK || s.kind == SSC_SRC_TRACK_HERNQUIST) {
    const double* row = s.pos + ((int64_t)step * s.n + j) * 3;
    sx = row[0];
    sy = row[1];
    sz = row[2];
    return true;
  }
  if (s.kind == SSC_SRC_LINE || s.kind == SSC_SRC_LINE_HERNQUIST) {
    const double dt = t - s.t0[j];
    if (!(fabs(dt) < s.t_window[j])) return false;
    sx = s.pos[3 * j] + s.vel[3 * j] * dt;
    sy = s.pos[3 * j + 1] + s.vel[3 * j + 1] * dt;
    sz = s.pos[3 * j + 2] + s.vel[3 * j + 2] * dt;
    return true;
  }
  return false;
}

}  // namespace ssc
