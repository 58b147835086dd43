// ORACLE -- test infrastructure only.  This is a CPU FP64 restatement of the reference
// algorithms on the stream-particle integration path, used by tests/, by smoke() and by the
// cpu_baseline leg of bench.py, as the reference results are compared against.  No product
// code may link, import or call it.
//
// Each function cites the reference file:line it restates (paths relative to the reference
// repository root).  Pins (what checks each part independently of this file) are listed in the
// Python wrapper oracle/__init__.py and exercised by tests/test_oracle_*.py.
//
// Build: oracle/Makefile (g++ -O2 -fopenmp, no -ffast-math, no FMA contraction).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

#include "../include/ssc_mi355x.h"

namespace {

// ---------------------------------------------------------------------------------------
// Forward-mode dual numbers (restates jax.jacfwd, used by the reference for release-model
// Jacobians, streamsculptor/perturbative.py:282-296, and for the tidal tensor, fields.py:193).
// ---------------------------------------------------------------------------------------
template <int N>
struct Dual {
  double v;
  double d[N];
  Dual() : v(0) { for (int i = 0; i < N; ++i) d[i] = 0; }
  Dual(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0; }
};
template <int N> Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) { Dual<N> r(a.v + b.v); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r; }
template <int N> Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) { Dual<N> r(a.v - b.v); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r; }
template <int N> Dual<N> operator-(const Dual<N>& a) { Dual<N> r(-a.v); for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r; }
template <int N> Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) { Dual<N> r(a.v * b.v); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
template <int N> Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) { double q = a.v / b.v; Dual<N> r(q); for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - q * b.d[i]) / b.v; return r; }
template <int N> Dual<N> operator+(const Dual<N>& a, double b) { Dual<N> r = a; r.v += b; return r; }
template <int N> Dual<N> operator+(double b, const Dual<N>& a) { return a + b; }
template <int N> Dual<N> operator-(const Dual<N>& a, double b) { Dual<N> r = a; r.v -= b; return r; }
template <int N> Dual<N> operator-(double b, const Dual<N>& a) { return (-a) + b; }
template <int N> Dual<N> operator*(const Dual<N>& a, double b) { Dual<N> r(a.v * b); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b; return r; }
template <int N> Dual<N> operator*(double b, const Dual<N>& a) { return a * b; }
template <int N> Dual<N> operator/(const Dual<N>& a, double b) { Dual<N> r(a.v / b); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / b; return r; }
template <int N> Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }
template <int N> Dual<N> sqrt(const Dual<N>& a) { double s = std::sqrt(a.v); Dual<N> r(s); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / (2.0 * s); return r; }
template <int N> Dual<N> log1p(const Dual<N>& a) { Dual<N> r(std::log1p(a.v)); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / (1.0 + a.v); return r; }
template <int N> Dual<N> cbrt(const Dual<N>& a) { double c = std::cbrt(a.v); Dual<N> r(c); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / (3.0 * c * c); return r; }

inline double sqrt(double x) { return std::sqrt(x); }
inline double log1p(double x) { return std::log1p(x); }
inline double cbrt(double x) { return std::cbrt(x); }
inline double val(double x) { return x; }
template <int N> inline double val(const Dual<N>& x) { return x.v; }

// ---------------------------------------------------------------------------------------
// Host components: phi, gradient and Hessian, closed forms.
// ---------------------------------------------------------------------------------------

// Spherical helper: given r (with r^2 = |x|^2 + soft), dphi/dr and d2phi/dr2, accumulate
// grad = phi' x / r and H = (phi'/r) I + (phi'' - phi'/r) x x^T / r^2.
template <typename T>
void spherical_accum(const T x[3], T r, T d1, T d2, T g[3], T H[9]) {
  T d1r = d1 / r;
  for (int i = 0; i < 3; ++i) g[i] = g[i] + d1r * x[i];
  if (H) {
    T c = (d2 - d1r) / (r * r);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        T h = c * x[i] * x[j];
        if (i == j) h = h + d1r;
        H[3 * i + j] = H[3 * i + j] + h;
      }
  }
}

// Miyamoto-Nagai disk, reference streamsculptor/potential.py:66-72:
//   phi = -G m / sqrt(R^2 + (a + sqrt(z^2 + b^2))^2)
template <typename T>
void mn_eval(const T p[4], const T x[3], T* phi, T g[3], T H[9]) {
  T gm = p[0], a = p[1], b = p[2];
  T zeta = sqrt(x[2] * x[2] + b * b);
  T S = a + zeta;
  T D2 = x[0] * x[0] + x[1] * x[1] + S * S;
  T D = sqrt(D2);
  T iD3 = 1.0 / (D2 * D);
  if (phi) *phi = *phi - gm / D;
  T dz = x[2] * S / zeta;  // d(D^2/2)/dz
  g[0] = g[0] + gm * x[0] * iD3;
  g[1] = g[1] + gm * x[1] * iD3;
  g[2] = g[2] + gm * dz * iD3;
  if (H) {
    T iD5 = iD3 / D2;
    T q[3] = {x[0], x[1], dz};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) H[3 * i + j] = H[3 * i + j] - 3.0 * gm * q[i] * q[j] * iD5;
    H[0] = H[0] + gm * iD3;
    H[4] = H[4] + gm * iD3;
    // d/dz (z S / zeta) = S b^2 / zeta^3 + z^2 / zeta^2
    T z2 = zeta * zeta;
    T ddz = S * b * b / (z2 * zeta) + x[2] * x[2] / z2;
    H[8] = H[8] + gm * ddz * iD3;
  }
}

// NFW, reference streamsculptor/potential.py:74-84: phi = -G m ln(1 + r/r_s) / r
template <typename T>
void nfw_eval(const T p[4], const T x[3], T* phi, T g[3], T H[9]) {
  T gm = p[0], rs = p[1];
  T r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  T L = log1p(r / rs);
  if (phi) *phi = *phi - gm * L / r;
  T rpr = r + rs;
  T d1 = gm * (L / (r * r) - 1.0 / (r * rpr));
  T d2 = gm * (1.0 / (r * r * rpr) - 2.0 * L / (r * r * r) + (2.0 * r + rs) / (r * r * rpr * rpr));
  spherical_accum(x, r, d1, d2, g, H);
}

// Hernquist, reference streamsculptor/potential.py:132-138: phi = -G m / (r + c),
// r = sqrt(|x|^2 + soft)
template <typename T>
void hernquist_eval(const T p[4], const T x[3], T* phi, T g[3], T H[9]) {
  T gm = p[0], c = p[1], soft = p[2];
  T r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + soft);
  T rc = r + c;
  if (phi) *phi = *phi - gm / rc;
  T d1 = gm / (rc * rc);
  T d2 = -2.0 * gm / (rc * rc * rc);
  spherical_accum(x, r, d1, d2, g, H);
}

// Plummer, reference streamsculptor/potential.py:124-130: phi = -G m / sqrt(r^2 + r_s^2)
template <typename T>
void plummer_eval(T gm, T rs, const T x[3], T* phi, T g[3], T H[9]) {
  T s2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + rs * rs;
  T s = sqrt(s2);
  T i1 = 1.0 / s;
  T i3 = i1 / s2;
  if (phi) *phi = *phi - gm * i1;
  for (int i = 0; i < 3; ++i) g[i] = g[i] + gm * i3 * x[i];
  if (H) {
    T i5 = i3 / s2;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        T h = -3.0 * gm * i5 * x[i] * x[j];
        if (i == j) h = h + gm * i3;
        H[3 * i + j] = H[3 * i + j] + h;
      }
  }
}

template <typename T>
void comp_eval(int kind, const T p[4], const T x[3], T* phi, T g[3], T H[9]) {
  switch (kind) {
    case SSC_COMP_MIYAMOTO_NAGAI: mn_eval(p, x, phi, g, H); break;
    case SSC_COMP_NFW: nfw_eval(p, x, phi, g, H); break;
    case SSC_COMP_HERNQUIST: hernquist_eval(p, x, phi, g, H); break;
    case SSC_COMP_PLUMMER: plummer_eval(p[0], p[1], x, phi, g, H); break;
    default: break;
  }
}

// Sum over components (reference Potential_Combine, streamsculptor/potential.py:1279-1296).
template <typename T>
void host_eval_T(const ssc_host_potential* hp, const T x[3], T* phi, T g[3], T H[9]) {
  if (phi) *phi = T(0.0);
  for (int i = 0; i < 3; ++i) g[i] = T(0.0);
  if (H) for (int i = 0; i < 9; ++i) H[i] = T(0.0);
  for (int c = 0; c < hp->n_components; ++c) {
    T p[4];
    for (int k = 0; k < 4; ++k) p[k] = T(hp->c[c].p[k]);
    comp_eval(hp->c[c].kind, p, x, phi, g, H);
  }
}

// d(grad Phi_c)/d(p1_c) for one component (p1 = MN a, NFW r_s, Hernquist c, Plummer r_s);
// grad Phi_c is linear in G m_c, so d/d(G m_c) = grad Phi_c / (G m_c).
void comp_grad_dp1(int kind, const double p[4], const double x[3], double out[3]) {
  const double gm = p[0];
  out[0] = out[1] = out[2] = 0.0;
  if (kind == SSC_COMP_MIYAMOTO_NAGAI) {
    double a = p[1], b = p[2];
    double zeta = std::sqrt(x[2] * x[2] + b * b), S = a + zeta;
    double D2 = x[0] * x[0] + x[1] * x[1] + S * S, D = std::sqrt(D2);
    double iD3 = 1.0 / (D2 * D), iD5 = iD3 / D2;
    out[0] = -3.0 * gm * x[0] * S * iD5;
    out[1] = -3.0 * gm * x[1] * S * iD5;
    out[2] = gm * (x[2] / zeta) * iD3 - 3.0 * gm * (x[2] * S / zeta) * S * iD5;
  } else if (kind == SSC_COMP_NFW) {
    double rs = p[1];
    double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    double rp = r + rs;
    double dd1 = gm * (-1.0 / (r * rs * rp) + 1.0 / (r * rp * rp));
    for (int i = 0; i < 3; ++i) out[i] = dd1 / r * x[i];
  } else if (kind == SSC_COMP_HERNQUIST) {
    double c = p[1], soft = p[2];
    double r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + soft);
    double rc = r + c;
    for (int i = 0; i < 3; ++i) out[i] = -2.0 * gm * x[i] / (r * rc * rc * rc);
  } else if (kind == SSC_COMP_PLUMMER) {
    double rs = p[1];
    double s2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + rs * rs, s = std::sqrt(s2);
    for (int i = 0; i < 3; ++i) out[i] = -3.0 * gm * rs * x[i] / (s2 * s2 * s);
  }
}

// Position of source j at grid row `step` / time t; returns false when the source is inactive
// (LINE window, reference streamsculptor/potential.py:826).
bool source_pos(const ssc_sources* s, int j, int step, double t, double out[3]) {
  if (s->kind == SSC_SRC_STATIC) {
    for (int i = 0; i < 3; ++i) out[i] = s->pos[3 * j + i];
    return true;
  }
  if (s->kind == SSC_SRC_TRACK || s->kind == SSC_SRC_TRACK_HERNQUIST) {
    const double* row = s->pos + ((int64_t)step * s->n + j) * 3;
    for (int i = 0; i < 3; ++i) out[i] = row[i];
    return true;
  }
  if (s->kind == SSC_SRC_LINE || s->kind == SSC_SRC_LINE_HERNQUIST) {
    double dtt = t - s->t0[j];
    if (!(std::fabs(dtt) < s->t_window[j])) return false;
    for (int i = 0; i < 3; ++i) out[i] = s->pos[3 * j + i] + s->vel[3 * j + i] * dtt;
    return true;
  }
  return false;
}

// Direct N x M sum over Plummer spheres (reference SubhaloLinePotential.potential,
// streamsculptor/potential.py:813-830) or, for LINE_HERNQUIST, Hernquist spheres (reference
// SubhaloLinePotentialCustom_fromFunc with func=HernquistPotential, potential.py:1161-1190);
// TRACK_HERNQUIST: Hernquist spheres centred on a sampled track (reference ProgenitorPotential /
// TimeDepTranslatingPotential with a Hernquist profile, potential.py:140-153,448-462).
void sources_eval(const ssc_sources* s, const double x[3], int step, double t,
                  double* phi, double g[3], double H[9]) {
  if (!s || s->kind == SSC_SRC_NONE) return;
  for (int j = 0; j < s->n; ++j) {
    double sp[3];
    if (!source_pos(s, j, step, t, sp)) continue;
    double d[3] = {x[0] - sp[0], x[1] - sp[1], x[2] - sp[2]};
    if (s->kind == SSC_SRC_LINE_HERNQUIST || s->kind == SSC_SRC_TRACK_HERNQUIST) {
      const double p[4] = {s->gm[j], s->rs[j], 0.0, 0.0};
      hernquist_eval<double>(p, d, phi, g, H);
    } else {
      plummer_eval<double>(s->gm[j], s->rs[j], d, phi, g, H);
    }
  }
}

void total_eval(const ssc_host_potential* hp, const ssc_sources* s, const double x[3], int step,
                double t, double* phi, double g[3], double H[9]) {
  host_eval_T<double>(hp, x, phi, g, H);
  sources_eval(s, x, step, t, phi, g, H);
}

void accel(const ssc_host_potential* hp, const ssc_sources* s, const double x[3], int step,
           double t, double a[3]) {
  double g[3];
  total_eval(hp, s, x, step, t, nullptr, g, nullptr);
  for (int i = 0; i < 3; ++i) a[i] = -g[i];
}

// ---------------------------------------------------------------------------------------
// Threefry-2x32 with 20 rounds (Salmon et al., SC'11; the cipher behind jax.random's default
// PRNG used by release_model, reference streamsculptor/main.py:220-228,263-266).
// ---------------------------------------------------------------------------------------
inline uint32_t rotl32(uint32_t v, int r) { return (v << r) | (v >> (32 - r)); }

void threefry2x32(uint32_t k0, uint32_t k1, uint32_t x0, uint32_t x1, uint32_t out[2]) {
  static const int R[8] = {13, 15, 26, 6, 17, 29, 16, 24};
  uint32_t ks[3] = {k0, k1, k0 ^ k1 ^ 0x1BD11BDAu};
  uint32_t v0 = x0 + ks[0], v1 = x1 + ks[1];
  for (int blk = 0; blk < 5; ++blk) {
    for (int r = 0; r < 4; ++r) {
      v0 += v1;
      v1 = rotl32(v1, R[(blk % 2) * 4 + r]);
      v1 ^= v0;
    }
    v0 += ks[(blk + 1) % 3];
    v1 += ks[(blk + 2) % 3] + (uint32_t)(blk + 1);
  }
  out[0] = v0;
  out[1] = v1;
}

// PRNGKey(seed) for an int64 seed: key = (seed >> 32, seed & 0xffffffff).
void prng_key(int64_t seed, uint32_t key[2]) {
  key[0] = (uint32_t)((uint64_t)seed >> 32);
  key[1] = (uint32_t)((uint64_t)seed & 0xffffffffu);
}

// 64-bit random words for a flat shape of n elements: counters 0..2n-1 split into halves
// (j, n + j); word j = (y0 << 32) | y1.
uint64_t random_bits64(const uint32_t key[2], int n, int j) {
  uint32_t y[2];
  threefry2x32(key[0], key[1], (uint32_t)j, (uint32_t)(n + j), y);
  return ((uint64_t)y[0] << 32) | (uint64_t)y[1];
}

// split(key) -> (k1, k2): counters 0..3 split into halves (0,2), (1,3).
void prng_split(const uint32_t key[2], uint32_t k1[2], uint32_t k2[2]) {
  uint32_t a[2], b[2];
  threefry2x32(key[0], key[1], 0u, 2u, a);
  threefry2x32(key[0], key[1], 1u, 3u, b);
  k1[0] = a[0]; k1[1] = b[0];
  k2[0] = a[1]; k2[1] = b[1];
}

// randint(key, (5,), 0, 1000) with int64 output (reference streamsculptor/main.py:221).
void randint5_1000(int64_t seed, int64_t out[5]) {
  uint32_t key[2], k1[2], k2[2];
  prng_key(seed, key);
  prng_split(key, k1, k2);
  const uint64_t span = 1000;
  uint64_t mult = ((uint64_t)1 << 32) % span;
  mult = (mult * mult) % span;
  for (int j = 0; j < 5; ++j) {
    uint64_t hb = random_bits64(k1, 5, j), lb = random_bits64(k2, 5, j);
    uint64_t off = ((hb % span) * mult + (lb % span)) % span;
    out[j] = (int64_t)off;
  }
}

// Inverse error function: Winitzki's closed-form start (rel. err ~2e-3) refined by three
// Halley steps on erf (|u| < 0.5) or erfc (tails, where 1 - |u| is exact).
double erfinv(double u) {
  if (u == 0.0) return 0.0;
  if (u >= 1.0) return INFINITY;
  if (u <= -1.0) return -INFINITY;
  double s = u < 0 ? -1.0 : 1.0;
  double a = std::fabs(u);
  const double A = 0.147;
  double ln = std::log((1.0 - a) * (1.0 + a));
  double t1 = 2.0 / (M_PI * A) + 0.5 * ln;
  double x = std::sqrt(std::sqrt(t1 * t1 - ln / A) - t1);
  const double c = 2.0 / std::sqrt(M_PI);
  for (int it = 0; it < 3; ++it) {
    double f = (a < 0.5) ? (std::erf(x) - a) : ((1.0 - a) - std::erfc(x));
    double fp = c * std::exp(-x * x);
    x = x - f / (fp + x * f);
  }
  return s * x;
}

// normal(key, (1,)) in float64: uniform on [nextafter(-1, 0), 1) from the top 52 bits of one
// 64-bit word, then sqrt(2) * erfinv.
double prng_normal(int64_t key_seed) {
  uint32_t key[2];
  prng_key(key_seed, key);
  uint64_t bits = random_bits64(key, 1, 0);
  uint64_t fb = (bits >> 12) | 0x3FF0000000000000ull;
  double f;
  std::memcpy(&f, &fb, 8);
  f -= 1.0;
  const double lo = std::nextafter(-1.0, 0.0);
  double uval = f * 2.0 + lo;  // (hi - lo) rounds to 2.0 in float64
  if (uval < lo) uval = lo;
  return std::sqrt(2.0) * erfinv(uval);
}

// ---------------------------------------------------------------------------------------
// Release model (reference Potential.release_model, streamsculptor/main.py:209-280, with
// lagrange_pts/tidalr/omega/d2phidr2 from :77-112).  Templated so the Jacobian can be
// taken with duals (reference release_func_jacobian, perturbative.py:282-296).
// in[6] = progenitor (x, v); hostp = host component parameters as scalars.
// ---------------------------------------------------------------------------------------
// H_extra (nullable): Hessian of the potential's point sources at the progenitor, added to the
// host Hessian (the reference takes d2phidr2 of the whole potential, main.py:77-85).
template <typename T>
void release_T(const ssc_host_potential* hp, const T* hostp /*[n_comp*4]*/, const T in[6],
               T g_msat, const double z[4], const double kv[8], T out[12],
               const double* H_extra = nullptr) {
  T x[3] = {in[0], in[1], in[2]}, v[3] = {in[3], in[4], in[5]};
  // Hessian of the host at x.
  T g[3], H[9];
  for (int i = 0; i < 3; ++i) g[i] = T(0.0);
  for (int i = 0; i < 9; ++i) H[i] = T(0.0);
  for (int c = 0; c < hp->n_components; ++c) comp_eval(hp->c[c].kind, hostp + 4 * c, x, (T*)nullptr, g, H);
  if (H_extra)
    for (int i = 0; i < 9; ++i) H[i] = H[i] + T(H_extra[i]);
  T rad = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  T rhat[3] = {x[0] / rad, x[1] / rad, x[2] / rad};
  // d2phidr2 = rhat^T H rhat (rhat held fixed), main.py:77-85
  T d2 = T(0.0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) d2 = d2 + rhat[i] * H[3 * i + j] * rhat[j];
  // omega = |x cross v| / r^2, main.py:89-96
  T L[3] = {x[1] * v[2] - x[2] * v[1], x[2] * v[0] - x[0] * v[2], x[0] * v[1] - x[1] * v[0]};
  T r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  T ov[3] = {L[0] / r2, L[1] / r2, L[2] / r2};
  T om = sqrt(ov[0] * ov[0] + ov[1] * ov[1] + ov[2] * ov[2]);
  // tidal radius, main.py:99-104
  T rt = cbrt(g_msat / (om * om - d2));
  T relv = om * rt;
  T Ln = sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
  T zh[3] = {L[0] / Ln, L[1] / Ln, L[2] / Ln};
  T vr = v[0] * rhat[0] + v[1] * rhat[1] + v[2] * rhat[2];
  T pv[3] = {v[0] - vr * rhat[0], v[1] - vr * rhat[1], v[2] - vr * rhat[2]};
  T pn = sqrt(pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2]);
  T ph[3] = {pv[0] / pn, pv[1] / pn, pv[2] / pn};
  // samples, main.py:263-266
  double kr = kv[0] + z[0] * kv[4];
  double kvphi = kr * (kv[1] + z[1] * kv[5]);
  double kz = kv[2] + z[2] * kv[6];
  double kvz = kv[3] + z[3] * kv[7];
  // main.py:269-278; out = [pos_lead(3), pos_trail(3), v_lead(3), v_trail(3)]
  for (int i = 0; i < 3; ++i) {
    T ptr = x[i] + kr * rhat[i] * rt;
    ptr = ptr + zh[i] * kz * rt;
    T vtr = v[i] + (kvphi * relv) * ph[i];
    vtr = vtr + (kvz * relv) * zh[i];
    T pld = x[i] + kr * rhat[i] * (-rt);
    pld = pld + zh[i] * kz * (-rt);
    T vld = v[i] + (kvphi * relv * -1.0) * ph[i];
    vld = vld + (kvz * relv * -1.0) * zh[i];
    out[i] = pld;
    out[3 + i] = ptr;
    out[6 + i] = vld;
    out[9 + i] = vtr;
  }
}

const double kDefaultK[8] = {2.0, 0.3, 0.0, 0.0, 0.4, 0.4, 0.5, 0.5};

void release_normals(int64_t seed, int64_t i, double z[4]) {
  int64_t ri[5];
  randint5_1000(seed, ri);
  for (int k = 0; k < 4; ++k) z[k] = prng_normal(i * ri[k]);
}

// ---------------------------------------------------------------------------------------
// Adaptive Dormand-Prince 5(4) as the reference runs it through diffrax: diffeqsolve with
// Dopri5 and PIDController(rtol, atol, dtmin, dtmax, force_dtmin=True, jump_ts), dt0=None,
// SaveAt(ts=ts) (reference streamsculptor/main.py:139-162, fields.py:48-98).  diffrax 0.7.0
// is not part of the reference tree; the step control restated here is its documented
// integral controller (pcoeff = dcoeff = 0, icoeff = 1; safety 0.9, factor in [0.2, 10],
// accepted steps never shrink, error order 5, RMS norm), Hairer's starting-step heuristic and
// its end-of-interval clipping.
//   Tableau: Dormand & Prince, J. Comput. Appl. Math. 6 (1980) 19-26.
//   Dense output: the quartic through y0, y1, h f0, h f1 and Shampine's (1986) midpoint value.
// Time runs in tau = dir * t so that the interval is always increasing.
// ---------------------------------------------------------------------------------------
namespace dp5 {
const double A21 = 1.0 / 5;
const double A31 = 3.0 / 40, A32 = 9.0 / 40;
const double A41 = 44.0 / 45, A42 = -56.0 / 15, A43 = 32.0 / 9;
const double A51 = 19372.0 / 6561, A52 = -25360.0 / 2187, A53 = 64448.0 / 6561, A54 = -212.0 / 729;
const double A61 = 9017.0 / 3168, A62 = -355.0 / 33, A63 = 46732.0 / 5247, A64 = 49.0 / 176,
             A65 = -5103.0 / 18656;
const double C2 = 1.0 / 5, C3 = 3.0 / 10, C4 = 4.0 / 5, C5 = 8.0 / 9;
const double B1 = 35.0 / 384, B3 = 500.0 / 1113, B4 = 125.0 / 192, B5 = -2187.0 / 6784,
             B6 = 11.0 / 84;
// b - bhat (bhat: the embedded 4th-order weights)
const double E1 = 71.0 / 57600, E3 = -71.0 / 16695, E4 = 71.0 / 1920, E5 = -17253.0 / 339200,
             E6 = 22.0 / 525, E7 = -1.0 / 40;
// Shampine's midpoint weights: y(t + h/2) = y0 + h sum_i M_i k_i
const double M1 = 6025192743.0 / 30085553152.0 / 2, M3 = 51252292925.0 / 65400821598.0 / 2,
             M4 = -2691868925.0 / 45128329728.0 / 2, M5 = 187940372067.0 / 1594534317056.0 / 2,
             M6 = -1776094331.0 / 19743644256.0 / 2, M7 = 11237099.0 / 235043384.0 / 2;
}  // namespace dp5

struct Dp5Opts {
  double rtol, atol, dtmin, dtmax;  // dtmin/dtmax <= 0: none
  double dt0;                       // <= 0: starting-step heuristic
  int32_t max_steps;
  const double* stops;  // tau-space, ascending
  int32_t n_stops;
};

void dp5_rhs(const ssc_host_potential* hp, const ssc_sources* src, double dir, double tau,
             const double y[6], double k[6]) {
  double a[3];
  accel(hp, src, y, 0, dir * tau, a);
  for (int c = 0; c < 3; ++c) { k[c] = dir * y[3 + c]; k[3 + c] = dir * a[c]; }
}

double rms6(const double v[6]) {
  double s = 0;
  for (int c = 0; c < 6; ++c) s += v[c] * v[c];
  return std::sqrt(s / 6.0);
}

// End the step at the first stop strictly after tau0 if the step would pass it.
double clip_stops(const Dp5Opts& o, double tau0, double tau1) {
  for (int j = 0; j < o.n_stops; ++j)
    if (o.stops[j] > tau0) return tau1 > o.stops[j] ? o.stops[j] : tau1;
  return tau1;
}

// One Dopri5 step of size h from (tau, y) with f0 = f(tau, y): the 5th-order solution yn, the
// embedded error estimate err = h sum (b - bhat) k, f(tau + h, yn) (first stage of the next
// step) and the dense-output quartic y(s) = y + s (F0 + s (c2 + s (b3 + s a4))), s in [0, 1].
struct Dp5Step {
  double yn[6], err[6], k7[6], F0[6], a4[6], b3[6], c2[6];
};

void dp5_step(const ssc_host_potential* hp, const ssc_sources* src, double dir, double tau,
              double h, const double y[6], const double f0[6], Dp5Step& r) {
  using namespace dp5;
  double k1[6], k2[6], k3[6], k4[6], k5[6], k6[6], ys_[6];
  for (int c = 0; c < 6; ++c) k1[c] = f0[c];
  for (int c = 0; c < 6; ++c) ys_[c] = y[c] + h * (A21 * k1[c]);
  dp5_rhs(hp, src, dir, tau + C2 * h, ys_, k2);
  for (int c = 0; c < 6; ++c) ys_[c] = y[c] + h * (A31 * k1[c] + A32 * k2[c]);
  dp5_rhs(hp, src, dir, tau + C3 * h, ys_, k3);
  for (int c = 0; c < 6; ++c) ys_[c] = y[c] + h * (A41 * k1[c] + A42 * k2[c] + A43 * k3[c]);
  dp5_rhs(hp, src, dir, tau + C4 * h, ys_, k4);
  for (int c = 0; c < 6; ++c)
    ys_[c] = y[c] + h * (A51 * k1[c] + A52 * k2[c] + A53 * k3[c] + A54 * k4[c]);
  dp5_rhs(hp, src, dir, tau + C5 * h, ys_, k5);
  for (int c = 0; c < 6; ++c)
    ys_[c] = y[c] + h * (A61 * k1[c] + A62 * k2[c] + A63 * k3[c] + A64 * k4[c] + A65 * k5[c]);
  dp5_rhs(hp, src, dir, tau + h, ys_, k6);
  for (int c = 0; c < 6; ++c)
    r.yn[c] = y[c] + h * (B1 * k1[c] + B3 * k3[c] + B4 * k4[c] + B5 * k5[c] + B6 * k6[c]);
  dp5_rhs(hp, src, dir, tau + h, r.yn, r.k7);
  const double* k7 = r.k7;
  for (int c = 0; c < 6; ++c) {
    r.err[c] = h * (E1 * k1[c] + E3 * k3[c] + E4 * k4[c] + E5 * k5[c] + E6 * k6[c] + E7 * k7[c]);
    const double ym =
        y[c] + h * (M1 * k1[c] + M3 * k3[c] + M4 * k4[c] + M5 * k5[c] + M6 * k6[c] + M7 * k7[c]);
    // quartic with p(0) = y, p'(0) = F0, p(1) = yn, p'(1) = h k7, p(1/2) = ym
    r.F0[c] = h * k1[c];
    const double D1 = r.yn[c] - y[c] - r.F0[c], D2 = h * k7[c] - r.F0[c],
                 D3 = ym - y[c] - 0.5 * r.F0[c];
    r.a4[c] = 2.0 * D2 - 8.0 * D1 + 16.0 * D3;
    r.b3[c] = D2 - 2.0 * D1 - 2.0 * r.a4[c];
    r.c2[c] = D1 - r.a4[c] - r.b3[c];
  }
}

void dp5_dense(const Dp5Step& r, const double y[6], double s, double out[6]) {
  for (int c = 0; c < 6; ++c)
    out[c] = y[c] + s * (r.F0[c] + s * (r.c2[c] + s * (r.b3[c] + s * r.a4[c])));
}

// One orbit from tau0 to tau1.  Saves: tau_s[j] (ascending) in [tau0, tau1] are written to
// ys[j]; returns 0 on success, 1 when max_steps ran out (then every output is +inf).
int dp5_orbit(const ssc_host_potential* hp, const ssc_sources* src, const Dp5Opts& o, double dir,
              double tau0, double tau1, const double w0[6], const double* tau_s, int32_t nts,
              double* ys, double w_out[6], int32_t stats[3]) {
  using namespace dp5;
  double y[6], f0[6];
  for (int c = 0; c < 6; ++c) y[c] = w0[c];
  int js = 0;
  while (js < nts && tau_s[js] < tau0) {
    for (int c = 0; c < 6; ++c) ys[6 * js + c] = NAN;  // before this orbit's start
    ++js;
  }
  while (js < nts && tau_s[js] == tau0) {
    for (int c = 0; c < 6; ++c) ys[6 * js + c] = y[c];
    ++js;
  }
  int32_t nacc = 0, nrej = 0, nsteps = 0;
  double tprev = tau0;
  if (tau0 < tau1) {
    dp5_rhs(hp, src, dir, tau0, y, f0);
    // starting step (Hairer, Norsett & Wanner, Solving ODEs I, sec. II.4)
    double sc[6], u[6];
    for (int c = 0; c < 6; ++c) sc[c] = o.atol + std::fabs(y[c]) * o.rtol;
    for (int c = 0; c < 6; ++c) u[c] = y[c] / sc[c];
    const double d0 = rms6(u);
    for (int c = 0; c < 6; ++c) u[c] = f0[c] / sc[c];
    const double d1 = rms6(u);
    const double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
    double y1[6], f1[6];
    for (int c = 0; c < 6; ++c) y1[c] = y[c] + h0 * f0[c];
    dp5_rhs(hp, src, dir, tau0 + h0, y1, f1);
    for (int c = 0; c < 6; ++c) u[c] = (f1[c] - f0[c]) / sc[c];
    const double d2 = rms6(u) / h0;
    const double md = d1 > d2 ? d1 : d2;
    const double h1 = md <= 1e-15 ? std::max(1e-6, h0 * 1e-3) : std::pow(0.01 / md, 1.0 / 5.0);
    double dt0 = o.dt0 > 0 ? o.dt0 : std::min(100.0 * h0, h1);
    // PIDController.init clamps the first step to [dtmin, dtmax] and a step raised to dtmin is
    // accepted whatever its error
    bool at_dtmin = false;
    if (o.dtmax > 0) dt0 = std::min(dt0, o.dtmax);
    if (o.dtmin > 0) {
      at_dtmin = dt0 <= o.dtmin;
      dt0 = std::max(dt0, o.dtmin);
    }
    double tnext = clip_stops(o, tau0, tau0 + dt0);
    if (tnext > tau1) tnext = tau1;
    while (tprev < tau1 && nsteps < o.max_steps) {
      const double h = tnext - tprev;
      Dp5Step r;
      dp5_step(hp, src, dir, tprev, h, y, f0, r);
      for (int c = 0; c < 6; ++c)
        u[c] = r.err[c] / (o.atol + o.rtol * std::max(std::fabs(y[c]), std::fabs(r.yn[c])));
      const double e = rms6(u);
      const bool keep = (e < 1.0) || at_dtmin;
      const double fmin = keep ? 1.0 : 0.2;
      double factor = 0.9 * std::pow(1.0 / e, 1.0 / 5.0);
      factor = std::min(std::max(factor, fmin), 10.0);
      double dtn = h * factor;
      if (o.dtmax > 0 && dtn > o.dtmax) dtn = o.dtmax;
      if (o.dtmin > 0) {
        at_dtmin = dtn <= o.dtmin;
        if (dtn < o.dtmin) dtn = o.dtmin;
      }
      double t_new = tprev;
      if (keep) {
        // dense output on (tprev, tnext]
        while (js < nts && tau_s[js] <= tnext) {
          if (tau_s[js] == tnext) {
            for (int c = 0; c < 6; ++c) ys[6 * js + c] = r.yn[c];
          } else {
            dp5_dense(r, y, (tau_s[js] - tprev) / h, ys + 6 * js);
          }
          ++js;
        }
        for (int c = 0; c < 6; ++c) { y[c] = r.yn[c]; f0[c] = r.k7[c]; }
        t_new = tnext;
        ++nacc;
      } else {
        ++nrej;
      }
      double tn1 = clip_stops(o, t_new, t_new + dtn);
      tprev = t_new < tau1 ? t_new : tau1;
      if (tn1 > tau1 - 1e-10) tn1 = keep ? tau1 : tprev + 0.5 * (tau1 - tprev);
      tnext = tn1;
      ++nsteps;
    }
  }
  stats[0] = nacc;
  stats[1] = nrej;
  stats[2] = tprev < tau1 ? 1 : 0;
  if (stats[2]) {
    for (int j = 0; j < nts; ++j)
      for (int c = 0; c < 6; ++c) ys[6 * j + c] = INFINITY;
    for (int c = 0; c < 6; ++c) w_out[c] = INFINITY;
    return 1;
  }
  for (; js < nts; ++js)
    for (int c = 0; c < 6; ++c) ys[6 * js + c] = NAN;  // after t1
  for (int c = 0; c < 6; ++c) w_out[c] = y[c];
  return 0;
}

// Source terms of the first-order responses to perturbing source j at (x, grid row k, t):
// gj = grad Phi_j(x) (mass block) and rj = grad(dPhi_j/da_j)(x) (radius block), both zero when
// the source is inactive.  Plummer: dPhi/da = G m a / s^3, grad = -3 G m a (x - s) / s^5;
// Hernquist (LINE_HERNQUIST): dPhi/da = G m / (r + a)^2, grad = -2 G m (x - s) / (r (r + a)^3).
void pert_terms(const ssc_sources* pert, int j, int k, double t, const double x[3], double gj[3],
                double rj[3]) {
  for (int c = 0; c < 3; ++c) gj[c] = rj[c] = 0.0;
  double sp[3];
  if (!source_pos(pert, j, k, t, sp)) return;
  double dv3[3] = {x[0] - sp[0], x[1] - sp[1], x[2] - sp[2]};
  double A = pert->rs[j];
  double d2 = dv3[0] * dv3[0] + dv3[1] * dv3[1] + dv3[2] * dv3[2];
  double c5;
  if (pert->kind == SSC_SRC_LINE_HERNQUIST) {
    const double p[4] = {pert->gm[j], A, 0.0, 0.0};
    hernquist_eval<double>(p, dv3, nullptr, gj, nullptr);
    double r = std::sqrt(d2), ra = r + A;
    c5 = -2.0 * pert->gm[j] / (r * ra * ra * ra);
  } else {
    plummer_eval<double>(pert->gm[j], A, dv3, nullptr, gj, nullptr);
    double s2 = d2 + A * A;
    c5 = -3.0 * pert->gm[j] * A / (s2 * s2 * std::sqrt(s2));
  }
  for (int c = 0; c < 3; ++c) rj[c] = c5 * dv3[c];
}

// Dopri5 with the controller of dp5_orbit on an N-dimensional state y (tau space; f(tau, y, k)
// gives dy/dtau), the error norm being the RMS over all N components (diffrax's RMS norm over
// every leaf of a pytree state).  Saves at tau_s[0..nts) into ys[nts][N]; returns 0 on success,
// 1 when max_steps ran out (then ys rows and y are +inf).  No jump times.
template <class F>
int dp5_system(F&& f, int N, const Dp5Opts& o, double tau0, double tau1, std::vector<double>& y,
               const double* tau_s, int32_t nts, double* ys, int32_t stats[3]) {
  using namespace dp5;
  std::vector<double> f0(N), k2(N), k3(N), k4(N), k5(N), k6(N), k7(N), yt(N), yn(N), sc(N);
  auto rms = [&](const std::vector<double>& u) {
    double s = 0;
    for (int c = 0; c < N; ++c) s += u[c] * u[c];
    return std::sqrt(s / N);
  };
  int js = 0;
  while (js < nts && tau_s[js] < tau0) {
    for (int c = 0; c < N; ++c) ys[(size_t)N * js + c] = NAN;
    ++js;
  }
  while (js < nts && tau_s[js] == tau0) {
    for (int c = 0; c < N; ++c) ys[(size_t)N * js + c] = y[c];
    ++js;
  }
  int32_t nacc = 0, nrej = 0, nsteps = 0;
  double tprev = tau0;
  if (tau0 < tau1) {
    f(tau0, y.data(), f0.data());
    std::vector<double> u(N);
    for (int c = 0; c < N; ++c) sc[c] = o.atol + std::fabs(y[c]) * o.rtol;
    for (int c = 0; c < N; ++c) u[c] = y[c] / sc[c];
    const double d0 = rms(u);
    for (int c = 0; c < N; ++c) u[c] = f0[c] / sc[c];
    const double d1 = rms(u);
    double dt0 = o.dt0;
    if (!(dt0 > 0)) {
      const double h0 = (d0 < 1e-5 || d1 < 1e-5) ? 1e-6 : 0.01 * (d0 / d1);
      for (int c = 0; c < N; ++c) yt[c] = y[c] + h0 * f0[c];
      f(tau0 + h0, yt.data(), k2.data());
      for (int c = 0; c < N; ++c) u[c] = (k2[c] - f0[c]) / sc[c];
      const double d2 = rms(u) / h0;
      const double md = d1 > d2 ? d1 : d2;
      const double h1 = md <= 1e-15 ? std::max(1e-6, h0 * 1e-3) : std::pow(0.01 / md, 1.0 / 5.0);
      dt0 = std::min(100.0 * h0, h1);
    }
    bool at_dtmin = false;  // first step clamped to [dtmin, dtmax] as in dp5_orbit
    if (o.dtmax > 0) dt0 = std::min(dt0, o.dtmax);
    if (o.dtmin > 0) {
      at_dtmin = dt0 <= o.dtmin;
      dt0 = std::max(dt0, o.dtmin);
    }
    double tnext = tau0 + dt0;
    if (tnext > tau1) tnext = tau1;
    while (tprev < tau1 && nsteps < o.max_steps) {
      const double h = tnext - tprev;
      for (int c = 0; c < N; ++c) yt[c] = y[c] + h * (A21 * f0[c]);
      f(tprev + C2 * h, yt.data(), k2.data());
      for (int c = 0; c < N; ++c) yt[c] = y[c] + h * (A31 * f0[c] + A32 * k2[c]);
      f(tprev + C3 * h, yt.data(), k3.data());
      for (int c = 0; c < N; ++c) yt[c] = y[c] + h * (A41 * f0[c] + A42 * k2[c] + A43 * k3[c]);
      f(tprev + C4 * h, yt.data(), k4.data());
      for (int c = 0; c < N; ++c)
        yt[c] = y[c] + h * (A51 * f0[c] + A52 * k2[c] + A53 * k3[c] + A54 * k4[c]);
      f(tprev + C5 * h, yt.data(), k5.data());
      for (int c = 0; c < N; ++c)
        yt[c] = y[c] + h * (A61 * f0[c] + A62 * k2[c] + A63 * k3[c] + A64 * k4[c] + A65 * k5[c]);
      f(tnext, yt.data(), k6.data());
      for (int c = 0; c < N; ++c)
        yn[c] = y[c] + h * (B1 * f0[c] + B3 * k3[c] + B4 * k4[c] + B5 * k5[c] + B6 * k6[c]);
      f(tnext, yn.data(), k7.data());
      for (int c = 0; c < N; ++c) {
        const double err = h * (E1 * f0[c] + E3 * k3[c] + E4 * k4[c] + E5 * k5[c] + E6 * k6[c] +
                                E7 * k7[c]);
        u[c] = err / (o.atol + o.rtol * std::max(std::fabs(y[c]), std::fabs(yn[c])));
      }
      const double e = rms(u);
      const bool keep = (e < 1.0) || at_dtmin;
      double factor = 0.9 * std::pow(1.0 / e, 1.0 / 5.0);
      factor = std::min(std::max(factor, keep ? 1.0 : 0.2), 10.0);
      double dtn = h * factor;
      if (o.dtmax > 0 && dtn > o.dtmax) dtn = o.dtmax;
      if (o.dtmin > 0) {
        at_dtmin = dtn <= o.dtmin;
        if (dtn < o.dtmin) dtn = o.dtmin;
      }
      double t_new = tprev;
      if (keep) {
        for (; js < nts && tau_s[js] <= tnext; ++js) {
          double* row = ys + (size_t)N * js;
          if (tau_s[js] == tnext) {
            for (int c = 0; c < N; ++c) row[c] = yn[c];
            continue;
          }
          const double sj = (tau_s[js] - tprev) / h;
          for (int c = 0; c < N; ++c) {
            const double ym = y[c] + h * (M1 * f0[c] + M3 * k3[c] + M4 * k4[c] + M5 * k5[c] +
                                          M6 * k6[c] + M7 * k7[c]);
            const double F0 = h * f0[c];
            const double D1 = yn[c] - y[c] - F0, D2 = h * k7[c] - F0, D3 = ym - y[c] - 0.5 * F0;
            const double a4 = 2.0 * D2 - 8.0 * D1 + 16.0 * D3;
            const double b3 = D2 - 2.0 * D1 - 2.0 * a4;
            const double c2 = D1 - a4 - b3;
            row[c] = y[c] + sj * (F0 + sj * (c2 + sj * (b3 + sj * a4)));
          }
        }
        for (int c = 0; c < N; ++c) { y[c] = yn[c]; f0[c] = k7[c]; }
        t_new = tnext;
        ++nacc;
      } else {
        ++nrej;
      }
      double tn1 = t_new + dtn;
      tprev = t_new < tau1 ? t_new : tau1;
      if (tn1 > tau1 - 1e-10) tn1 = keep ? tau1 : tprev + 0.5 * (tau1 - tprev);
      tnext = tn1;
      ++nsteps;
    }
  }
  stats[0] = nacc;
  stats[1] = nrej;
  stats[2] = tprev < tau1 ? 1 : 0;
  if (stats[2]) {
    for (size_t k = 0; k < (size_t)N * nts; ++k) ys[k] = INFINITY;
    for (int c = 0; c < N; ++c) y[c] = INFINITY;
    return 1;
  }
  for (; js < nts; ++js)
    for (int c = 0; c < N; ++c) ys[(size_t)N * js + c] = NAN;
  return 0;
}

}  // namespace

// =========================================================================================
// extern "C" surface of the oracle
// =========================================================================================
extern "C" {

void orc_threefry2x32(uint32_t k0, uint32_t k1, uint32_t x0, uint32_t x1, uint32_t* out) {
  threefry2x32(k0, k1, x0, x1, out);
}
void orc_randint5_1000(int64_t seed, int64_t* out) { randint5_1000(seed, out); }
double orc_prng_normal(int64_t key_seed) { return prng_normal(key_seed); }
double orc_erfinv(double u) { return erfinv(u); }
void orc_release_normals(int64_t seed, int64_t i, double* z) { release_normals(seed, i, z); }

// Potential / gradient / Hessian at n points.
int orc_eval(const ssc_host_potential* hp, const ssc_sources* src, const double* xyz, int64_t n,
             double t, int32_t step, double* phi, double* grad, double* hess) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double p = 0, g[3], H[9];
    total_eval(hp, src, xyz + 3 * i, step, t, &p, g, H);
    if (phi) phi[i] = p;
    if (grad) for (int k = 0; k < 3; ++k) grad[3 * i + k] = g[k];
    if (hess) for (int k = 0; k < 9; ++k) hess[9 * i + k] = H[k];
  }
  return 0;
}

// Host-parameter partials of the gradient at n points: out[i][c] = (d grad Phi_c / d(G m_c),
// d grad Phi_c / d p1_c), 6 doubles per component (the same partials the adjoint accumulates).
int orc_host_dgrad(const ssc_host_potential* hp, const double* xyz, int64_t n, double* out) {
  const int nc = hp->n_components;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    for (int c = 0; c < nc; ++c) {
      double gc[3] = {0, 0, 0}, dp[3];
      comp_eval<double>(hp->c[c].kind, hp->c[c].p, xyz + 3 * i, nullptr, gc, nullptr);
      comp_grad_dp1(hp->c[c].kind, hp->c[c].p, xyz + 3 * i, dp);
      double* o = out + ((int64_t)i * nc + c) * 6;
      for (int k = 0; k < 3; ++k) {
        o[k] = gc[k] / hp->c[c].p[0];
        o[3 + k] = dp[k];
      }
    }
  }
  return 0;
}

// Kick-drift-kick leapfrog on the grid t_k = t0 + k dt (Verlet 1967; Hairer, Lubich & Wanner,
// Geometric Numerical Integration, sec. I.1.4).  The reference has no fixed-step solver: this
// restates the north-star fixed-step replacement of integrate_orbit (main.py:125-163).
int orc_leapfrog(const ssc_host_potential* hp, const ssc_sources* src, const double* w0,
                 const int32_t* start, int64_t n, double t0, double dt, int32_t nsteps,
                 double* w_out, double* traj, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < n; ++i) {
    double x[3], v[3], a[3];
    for (int k = 0; k < 3; ++k) { x[k] = w0[6 * i + k]; v[k] = w0[6 * i + 3 + k]; }
    int s = start ? start[i] : 0;
    if (s < 0) s = 0;
    if (traj)
      for (int k = 0; k <= (s < nsteps ? s : nsteps); ++k)
        for (int c = 0; c < 6; ++c) traj[((int64_t)k * n + i) * 6 + c] = w0[6 * i + c];
    if (s < nsteps) accel(hp, src, x, s, t0 + s * dt, a);
    for (int k = s; k < nsteps; ++k) {
      double h = 0.5 * dt;
      for (int c = 0; c < 3; ++c) v[c] = v[c] + h * a[c];
      for (int c = 0; c < 3; ++c) x[c] = x[c] + dt * v[c];
      accel(hp, src, x, k + 1, t0 + (k + 1) * dt, a);
      for (int c = 0; c < 3; ++c) v[c] = v[c] + h * a[c];
      if (traj)
        for (int c = 0; c < 3; ++c) {
          traj[((int64_t)(k + 1) * n + i) * 6 + c] = x[c];
          traj[((int64_t)(k + 1) * n + i) * 6 + 3 + c] = v[c];
        }
    }
    for (int c = 0; c < 3; ++c) { w_out[6 * i + c] = x[c]; w_out[6 * i + 3 + c] = v[c]; }
  }
  return 0;
}

// Spray release conditions (reference gen_stream_ics, main.py:287-306) in host + sources:
// release i happens at time ts[i] (LINE sources) and grid row steps[i] (TRACK sources; NULL
// when there are none).
int orc_release(const ssc_host_potential* hp, const ssc_sources* src, const double* prog_w,
                const double* ts, const int32_t* steps, const double* g_msat, int32_t nrel,
                int64_t seed, const double* kvals, double* pos_close, double* pos_far,
                double* vel_close, double* vel_far) {
  const double* kv = kvals ? kvals : kDefaultK;
  std::vector<double> hostp(4 * hp->n_components);
  for (int c = 0; c < hp->n_components; ++c)
    for (int k = 0; k < 4; ++k) hostp[4 * c + k] = hp->c[c].p[k];
  const bool with_src = src && src->kind != SSC_SRC_NONE;
  if (with_src && (src->kind == SSC_SRC_TRACK || src->kind == SSC_SRC_TRACK_HERNQUIST) && !steps)
    return -1;
  if (with_src && (src->kind == SSC_SRC_LINE || src->kind == SSC_SRC_LINE_HERNQUIST) && !ts)
    return -1;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nrel; ++i) {
    double z[4], out[12], Hs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, gs[3] = {0, 0, 0};
    release_normals(seed, i, z);
    if (with_src)
      sources_eval(src, prog_w + 6 * i, steps ? steps[i] : 0, ts ? ts[i] : 0.0, nullptr, gs, Hs);
    release_T<double>(hp, hostp.data(), prog_w + 6 * i, g_msat[i], z, kv, out,
                      with_src ? Hs : nullptr);
    for (int k = 0; k < 3; ++k) {
      pos_close[3 * i + k] = out[k];
      pos_far[3 * i + k] = out[3 + k];
      vel_close[3 * i + k] = out[6 + k];
      vel_far[3 * i + k] = out[9 + k];
    }
  }
  return 0;
}

// Jacobian of the release model w.r.t. (x, v) of the progenitor and the Gm / scale parameter
// (p[0], p[1]) of every host component: jac[nrel][12][6 + 2*n_comp] (reference
// release_func_jacobian, perturbative.py:282-296, extended with host parameters).
int orc_release_jacobian(const ssc_host_potential* hp, const double* prog_w, const double* g_msat,
                         int32_t nrel, int64_t seed, const double* kvals, double* jac) {
  const double* kv = kvals ? kvals : kDefaultK;
  const int nc = hp->n_components;
  constexpr int ND = 6 + 2 * SSC_MAX_COMPONENTS;
  const int nin = 6 + 2 * nc;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nrel; ++i) {
    using D = Dual<ND>;
    double z[4];
    release_normals(seed, i, z);
    D in[6], hostp[4 * SSC_MAX_COMPONENTS], out[12];
    for (int k = 0; k < 6; ++k) { in[k] = D(prog_w[6 * i + k]); in[k].d[k] = 1.0; }
    for (int c = 0; c < nc; ++c)
      for (int k = 0; k < 4; ++k) {
        hostp[4 * c + k] = D(hp->c[c].p[k]);
        if (k < 2) hostp[4 * c + k].d[6 + 2 * c + k] = 1.0;
      }
    release_T<D>(hp, hostp, in, D(g_msat[i]), z, kv, out);
    for (int o = 0; o < 12; ++o)
      for (int k = 0; k < nin; ++k) jac[((int64_t)i * 12 + o) * nin + k] = out[o].d[k];
  }
  return 0;
}

// First-order responses to the perturbing sources, integrated as the exact derivative of the
// KDK map of the perturbed system Phi_base + sum_j eps_j Phi_j at eps = 0 (restates
// MassRadiusPerturbation_OTF.term, reference streamsculptor/fields.py:175-206: the base flow
// [v0, a0], the mass block d(dx)/dt = dv, d(dv)/dt = -grad Phi_j(x0) - H(x0) dx, and, when
// with_radius, the radius block driven by -grad(dPhi_j/da_j)); the sources are Plummer spheres
// or, for LINE_HERNQUIST, Hernquist spheres (reference SubhaloLinePotentialCustom_fromFunc and
// SubhaloLinePotentialCustom_dRadius_fromFunc, potential.py:1161-1268).  State per particle: w (6) and
// dw[j] = (dx, dv[, d2x/da, d2v/da]) for j < m.  H is the Hessian of the base potential.
int orc_perturb_leapfrog(const ssc_host_potential* hp, const ssc_sources* base_src,
                         const ssc_sources* pert, const double* w0, const double* dw0,
                         const int32_t* start, int64_t n, double t0, double dt, int32_t nsteps,
                         int with_radius, double* w_out, double* dw_out, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const int m = pert->n;
  const int dd = with_radius ? 12 : 6;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t i = 0; i < n; ++i) {
    double x[3], v[3], a[3], H[9], g[3];
    std::vector<double> d((size_t)m * dd), ad((size_t)m * dd / 2);
    for (int c = 0; c < 3; ++c) { x[c] = w0[6 * i + c]; v[c] = w0[6 * i + 3 + c]; }
    for (int j = 0; j < m * dd; ++j) d[j] = dw0 ? dw0[(size_t)i * m * dd + j] : 0.0;
    int s = start ? start[i] : 0;
    if (s < 0) s = 0;
    // accelerations of base and responses at grid row k
    auto forces = [&](int k) {
      double t = t0 + k * dt;
      total_eval(hp, base_src, x, k, t, nullptr, g, H);
      for (int c = 0; c < 3; ++c) a[c] = -g[c];
      for (int j = 0; j < m; ++j) {
        double gj[3], rj[3];
        pert_terms(pert, j, k, t, x, gj, rj);
        const double* dx = &d[(size_t)j * dd];
        double* o = &ad[(size_t)j * dd / 2];
        for (int r = 0; r < 3; ++r) {
          double hd = H[3 * r] * dx[0] + H[3 * r + 1] * dx[1] + H[3 * r + 2] * dx[2];
          o[r] = -gj[r] - hd;
        }
        if (with_radius) {
          const double* rx = dx + 6;
          for (int r = 0; r < 3; ++r) {
            double hd = H[3 * r] * rx[0] + H[3 * r + 1] * rx[1] + H[3 * r + 2] * rx[2];
            o[3 + r] = -rj[r] - hd;
          }
        }
      }
    };
    if (s < nsteps) forces(s);
    const double h = 0.5 * dt;
    for (int k = s; k < nsteps; ++k) {
      for (int c = 0; c < 3; ++c) v[c] = v[c] + h * a[c];
      for (int j = 0; j < m; ++j) {
        double* q = &d[(size_t)j * dd];
        const double* o = &ad[(size_t)j * dd / 2];
        for (int c = 0; c < 3; ++c) q[3 + c] = q[3 + c] + h * o[c];
        if (with_radius) for (int c = 0; c < 3; ++c) q[9 + c] = q[9 + c] + h * o[3 + c];
      }
      for (int c = 0; c < 3; ++c) x[c] = x[c] + dt * v[c];
      for (int j = 0; j < m; ++j) {
        double* q = &d[(size_t)j * dd];
        for (int c = 0; c < 3; ++c) q[c] = q[c] + dt * q[3 + c];
        if (with_radius) for (int c = 0; c < 3; ++c) q[6 + c] = q[6 + c] + dt * q[9 + c];
      }
      forces(k + 1);
      for (int c = 0; c < 3; ++c) v[c] = v[c] + h * a[c];
      for (int j = 0; j < m; ++j) {
        double* q = &d[(size_t)j * dd];
        const double* o = &ad[(size_t)j * dd / 2];
        for (int c = 0; c < 3; ++c) q[3 + c] = q[3 + c] + h * o[c];
        if (with_radius) for (int c = 0; c < 3; ++c) q[9 + c] = q[9 + c] + h * o[3 + c];
      }
    }
    for (int c = 0; c < 3; ++c) { w_out[6 * i + c] = x[c]; w_out[6 * i + 3 + c] = v[c]; }
    for (int j = 0; j < m * dd; ++j) dw_out[(size_t)i * m * dd + j] = d[j];
  }
  return 0;
}

// First-order responses integrated with adaptive Dopri5 (reference
// GenerateMassRadiusPerturbation with a Runge-Kutta solver: the progenitor's backward field
// solve, perturbative.py:52-60, and compute_perturbation_OTF, :95-135, through
// fields.integrate_field, fields.py:34-99, on MassRadiusPerturbation_OTF.term, :175-206).
// Particle i's state is [w (6), dw (m, D)] (D = 12 with with_radius, else 6), integrated from
// t_start[i] (t0 when NULL) to t1 as one ODE with one step controller (dp5_system: RMS norm
// over all 6 + m D components).  Saves at ts into ys[n][nts][6], dys[n][nts][m][D]; finals into
// w_out[n][6], dw_out[n][m][D]; stats[n][3] as orc_dopri5.
int orc_perturb_dopri5(const ssc_host_potential* hp, const ssc_sources* pert, const double* w0,
                       const double* dw0, const double* t_start, int64_t n, double t0, double t1,
                       const double* ts, int32_t nts, int with_radius, double rtol, double atol,
                       double dtmin, double dtmax, double dt0, int32_t max_steps, double* w_out,
                       double* dw_out, double* ys, double* dys, int32_t* stats, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  if (pert->kind == SSC_SRC_TRACK) return -1;  // sampled on a fixed grid only
  const int m = pert->n;
  const int dd = with_radius ? 12 : 6;
  const int N = 6 + m * dd;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
  for (int64_t i = 0; i < n; ++i) {
    const double ta = t_start ? t_start[i] : t0;
    const double dir = t1 >= ta ? 1.0 : -1.0;
    std::vector<double> tau_s(nts > 0 ? nts : 1), y(N), ysl((size_t)N * (nts > 0 ? nts : 1));
    for (int j = 0; j < nts; ++j) tau_s[j] = dir * ts[j];
    for (int c = 0; c < 6; ++c) y[c] = w0[6 * i + c];
    for (int k = 0; k < m * dd; ++k) y[6 + k] = dw0 ? dw0[(size_t)i * m * dd + k] : 0.0;
    // dy/dtau: base [v, -grad Phi_host]; per source j the mass block [dv, -gj - H dx] and the
    // radius block [rv, -rj - H rx]
    auto f = [&](double tau, const double* yy, double* kk) {
      const double t = dir * tau;
      double g[3], H[9];
      total_eval(hp, nullptr, yy, 0, t, nullptr, g, H);
      for (int c = 0; c < 3; ++c) { kk[c] = dir * yy[3 + c]; kk[3 + c] = -dir * g[c]; }
      for (int j = 0; j < m; ++j) {
        double gj[3], rj[3];
        pert_terms(pert, j, 0, t, yy, gj, rj);
        for (int b = 0; b < (with_radius ? 2 : 1); ++b) {
          const double* q = yy + 6 + (size_t)j * dd + 6 * b;
          double* o = kk + 6 + (size_t)j * dd + 6 * b;
          const double* s3 = b ? rj : gj;
          for (int r = 0; r < 3; ++r) {
            const double hq = H[3 * r] * q[0] + H[3 * r + 1] * q[1] + H[3 * r + 2] * q[2];
            o[r] = dir * q[3 + r];
            o[3 + r] = -dir * (s3[r] + hq);
          }
        }
      }
    };
    Dp5Opts o{rtol, atol, dtmin, dtmax, dt0, max_steps, nullptr, 0};
    int32_t st[3];
    bad += dp5_system(f, N, o, dir * ta, dir * t1, y, tau_s.data(), nts, ysl.data(), st);
    for (int js = 0; js < nts; ++js) {
      const double* row = ysl.data() + (size_t)N * js;
      if (ys) for (int c = 0; c < 6; ++c) ys[((size_t)i * nts + js) * 6 + c] = row[c];
      if (dys)
        for (int k = 0; k < m * dd; ++k) dys[((size_t)i * nts + js) * m * dd + k] = row[6 + k];
    }
    if (w_out) for (int c = 0; c < 6; ++c) w_out[6 * i + c] = y[c];
    if (dw_out) for (int k = 0; k < m * dd; ++k) dw_out[(size_t)i * m * dd + k] = y[6 + k];
    if (stats) for (int c = 0; c < 3; ++c) stats[3 * i + c] = st[c];
  }
  return bad;
}

// Discrete adjoint of the kick-drift-kick leapfrog (no reference counterpart: the reference
// differentiates forward-mode only, streamsculptor/main.py:160).  For every orbit, with the
// forward trajectory stored, the reverse sweep is
//   terminal: lx, lv = lam_T[i] (+ forcing at the last row); W = h lv; apply(x_T, W)
//   k = T-1 .. s: lv_half = lv + dt lx; lv = lv_half + F_v(k); lx += F_x(k);
//                 W = h lv_half + (k > s ? h lv : 0); apply(x_k, W)
//   apply(x, W): lx -= H(x) W (H = Hessian of host + sources); dL/dtheta += (da/dtheta)^T W;
//                per source j: dL/dS_jk += H_j(x - S_jk) W, dL/d(Gm_j) += -(x - S_jk).W / s^3
// Outputs: lam_0[i] = adjoint of the starting row, g_host[2c] = dL/d(G m_c),
// g_host[2c+1] = dL/d(p1_c), g_gm[j], g_pos[k][j][3] (TRACK rows; row 0 for STATIC).
int orc_adjoint_leapfrog(const ssc_host_potential* hp, const ssc_sources* src, const double* w0,
                         const int32_t* start, int64_t n, double t0, double dt, int32_t nsteps,
                         const double* lam_T, const double* forcing, double* lam_0,
                         double* g_host, double* g_gm, double* g_pos, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const int nc = hp->n_components;
  const int m = (src && src->kind != SSC_SRC_NONE) ? src->n : 0;
  const int64_t npos = (src && src->kind == SSC_SRC_TRACK) ? (int64_t)(nsteps + 1) * m * 3 : (int64_t)m * 3;
  for (int k = 0; k < 2 * nc; ++k) g_host[k] = 0.0;
  if (g_gm) for (int j = 0; j < m; ++j) g_gm[j] = 0.0;
  if (g_pos) for (int64_t k = 0; k < npos; ++k) g_pos[k] = 0.0;
  const double h = 0.5 * dt;
#pragma omp parallel
  {
    std::vector<double> gh(2 * nc, 0.0), gg(m, 0.0), gp(g_pos ? npos : 0, 0.0);
    std::vector<double> X;
#pragma omp for schedule(dynamic, 8)
    for (int64_t i = 0; i < n; ++i) {
      int s = start ? start[i] : 0;
      if (s < 0) s = 0;
      if (s > nsteps) s = nsteps;
      // forward, storing positions
      X.assign((size_t)(nsteps - s + 1) * 3, 0.0);
      double x[3], v[3], a[3];
      for (int c = 0; c < 3; ++c) { x[c] = w0[6 * i + c]; v[c] = w0[6 * i + 3 + c]; }
      for (int c = 0; c < 3; ++c) X[c] = x[c];
      if (s < nsteps) accel(hp, src, x, s, t0 + s * dt, a);
      for (int k = s; k < nsteps; ++k) {
        for (int c = 0; c < 3; ++c) v[c] += h * a[c];
        for (int c = 0; c < 3; ++c) x[c] += dt * v[c];
        accel(hp, src, x, k + 1, t0 + (k + 1) * dt, a);
        for (int c = 0; c < 3; ++c) v[c] += h * a[c];
        for (int c = 0; c < 3; ++c) X[(size_t)(k + 1 - s) * 3 + c] = x[c];
      }
      auto F = [&](int k, int c) -> double {
        return forcing ? forcing[((int64_t)k * n + i) * 6 + c] : 0.0;
      };
      auto apply = [&](int k, const double W[3], double lx[3]) {
        const double* xk = &X[(size_t)(k - s) * 3];
        const double t = t0 + k * dt;
        double g[3], H[9];
        total_eval(hp, src, xk, k, t, nullptr, g, H);
        for (int r = 0; r < 3; ++r) lx[r] -= H[3 * r] * W[0] + H[3 * r + 1] * W[1] + H[3 * r + 2] * W[2];
        for (int c = 0; c < nc; ++c) {
          double gc[3] = {0, 0, 0};
          comp_eval<double>(hp->c[c].kind, hp->c[c].p, xk, nullptr, gc, nullptr);
          double dp[3];
          comp_grad_dp1(hp->c[c].kind, hp->c[c].p, xk, dp);
          // a = -grad Phi: da/dGm = -grad Phi_c / Gm, da/dp1 = -d(grad Phi_c)/dp1
          gh[2 * c] += -(gc[0] * W[0] + gc[1] * W[1] + gc[2] * W[2]) / hp->c[c].p[0];
          gh[2 * c + 1] += -(dp[0] * W[0] + dp[1] * W[1] + dp[2] * W[2]);
        }
        for (int j = 0; j < m; ++j) {
          double sp[3];
          if (!source_pos(src, j, k, t, sp)) continue;
          double d[3] = {xk[0] - sp[0], xk[1] - sp[1], xk[2] - sp[2]};
          double rs = src->rs[j];
          double s2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + rs * rs, ss = std::sqrt(s2);
          double i3 = 1.0 / (s2 * ss), i5 = i3 / s2;
          double dW = d[0] * W[0] + d[1] * W[1] + d[2] * W[2];
          gg[j] += -dW * i3;
          if (g_pos) {
            double* o = &gp[(src->kind == SSC_SRC_TRACK ? (int64_t)k * m + j : (int64_t)j) * 3];
            double gmj = src->gm[j];
            for (int r = 0; r < 3; ++r) o[r] += gmj * (i3 * W[r] - 3.0 * i5 * d[r] * dW);
          }
        }
      };
      double lx[3], lv[3];
      for (int c = 0; c < 3; ++c) { lx[c] = lam_T[6 * i + c] + F(nsteps, c); lv[c] = lam_T[6 * i + 3 + c] + F(nsteps, 3 + c); }
      if (s < nsteps) {
        double W[3] = {h * lv[0], h * lv[1], h * lv[2]};
        apply(nsteps, W, lx);
      }
      for (int k = nsteps - 1; k >= s; --k) {
        double lh[3];
        for (int c = 0; c < 3; ++c) lh[c] = lv[c] + dt * lx[c];
        for (int c = 0; c < 3; ++c) { lv[c] = lh[c] + F(k, 3 + c); lx[c] += F(k, c); }
        double W[3];
        for (int c = 0; c < 3; ++c) W[c] = h * lh[c] + (k > s ? h * lv[c] : 0.0);
        apply(k, W, lx);
      }
      for (int c = 0; c < 3; ++c) { lam_0[6 * i + c] = lx[c]; lam_0[6 * i + 3 + c] = lv[c]; }
    }
#pragma omp critical
    {
      for (int k = 0; k < 2 * nc; ++k) g_host[k] += gh[k];
      if (g_gm) for (int j = 0; j < m; ++j) g_gm[j] += gg[j];
      if (g_pos) for (int64_t k = 0; k < npos; ++k) g_pos[k] += gp[k];
    }
  }
  return 0;
}


// Adaptive Dopri5 for n orbits (reference integrate_orbit / gen_stream_vmapped with
// solver=Dopri5, main.py:125-163,343-368).  Orbit i runs from t_start[i] (t0 when t_start is
// NULL) to t1 and is saved at ts[0..nts) (monotone in the direction of integration; entries
// outside the orbit's interval are NaN).  stops[n_stops]: times where a step must end.
// stats[i] = {accepted, rejected, status}; status 1 = max_steps reached (outputs +inf).
int orc_dopri5(const ssc_host_potential* hp, const ssc_sources* src, const double* w0,
               const double* t_start, int64_t n, double t0, double t1, const double* ts,
               int32_t nts, double rtol, double atol, double dtmin, double dtmax, double dt0,
               int32_t max_steps, const double* stops, int32_t n_stops, double* w_out,
               double* ys, int32_t* stats, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  int bad = 0;
#pragma omp parallel
  {
    std::vector<double> tau_s(nts > 0 ? nts : 1), tstop(n_stops > 0 ? n_stops : 1);
    std::vector<double> ysl(6 * (nts > 0 ? nts : 1));
#pragma omp for schedule(dynamic, 8) reduction(+ : bad)
    for (int64_t i = 0; i < n; ++i) {
      const double ta = t_start ? t_start[i] : t0;
      const double dir = t1 >= ta ? 1.0 : -1.0;
      for (int j = 0; j < nts; ++j) tau_s[j] = dir * ts[j];
      for (int j = 0; j < n_stops; ++j) tstop[j] = dir * stops[j];
      std::sort(tstop.begin(), tstop.begin() + n_stops);
      Dp5Opts o{rtol, atol, dtmin, dtmax, dt0, max_steps, tstop.data(), n_stops};
      double wo[6];
      int32_t st[3];
      // grouped sources: orbit i sees only the group_size sources of group[i]
      ssc_sources sub;
      const ssc_sources* s_i = src;
      if (src && src->group_size > 0) {
        const int64_t j0 = (int64_t)src->group[i] * src->group_size;
        sub = *src;
        sub.n = src->group_size;
        sub.group_size = 0;
        sub.group = nullptr;
        sub.gm = src->gm + j0;
        sub.rs = src->rs + j0;
        sub.pos = src->pos + 3 * j0;
        if (src->vel) sub.vel = src->vel + 3 * j0;
        if (src->t0) sub.t0 = src->t0 + j0;
        if (src->t_window) sub.t_window = src->t_window + j0;
        s_i = &sub;
      }
      bad += dp5_orbit(hp, s_i, o, dir, dir * ta, dir * t1, w0 + 6 * i, tau_s.data(), nts,
                       ysl.data(), wo, st);
      if (ys) for (int k = 0; k < 6 * nts; ++k) ys[6 * nts * i + k] = ysl[k];
      if (w_out) for (int c = 0; c < 6; ++c) w_out[6 * i + c] = wo[c];
      if (stats) for (int c = 0; c < 3; ++c) stats[3 * i + c] = st[c];
    }
  }
  return bad;
}

// Test hooks: the Dopri5 coefficients (A lower triangle row-major, c, b, b - bhat, midpoint
// weights; 15 + 4 * 7 doubles) and one step with dense output at s[0..ns).
void orc_dp5_tableau(double* out) {
  using namespace dp5;
  const double t[] = {A21, A31, A32, A41, A42, A43, A51, A52, A53, A54, A61, A62, A63, A64, A65,
                      0.0, C2, C3, C4, C5, 1.0, 1.0,
                      B1, 0.0, B3, B4, B5, B6, 0.0,
                      E1, 0.0, E3, E4, E5, E6, E7,
                      M1, 0.0, M3, M4, M5, M6, M7};
  for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); ++i) out[i] = t[i];
}

void orc_dp5_step(const ssc_host_potential* hp, const ssc_sources* src, double t, double h,
                  const double* y, double* yn, double* err, const double* s, int32_t ns,
                  double* dense) {
  double f0[6];
  dp5_rhs(hp, src, 1.0, t, y, f0);
  Dp5Step r;
  dp5_step(hp, src, 1.0, t, h, y, f0, r);
  for (int c = 0; c < 6; ++c) { yn[c] = r.yn[c]; err[c] = r.err[c]; }
  for (int j = 0; j < ns; ++j) dp5_dense(r, y, s[j], dense + 6 * j);
}

}  // extern "C"
