// Spray-stream release conditions on the GPU: one release per lane.  Restates the reference
// Potential.release_model (streamsculptor/main.py:209-280) with the closed-form host Hessian
// in place of nested autodiff, and the reference's key scheme on a Threefry-2x32-20 counter
// PRNG (Salmon et al., SC'11) for the Gaussian offsets.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ssc_device.h"

namespace ssc {

__device__ __forceinline__ uint32_t rotl32(uint32_t v, int r) { return (v << r) | (v >> (32 - r)); }

// Threefry-2x32, 20 rounds: five blocks of four rounds with a key injection after each.
__device__ void threefry2x32(uint32_t k0, uint32_t k1, uint32_t x0, uint32_t x1, uint32_t& o0,
                             uint32_t& o1) {
  const uint32_t k2 = k0 ^ k1 ^ 0x1BD11BDAu;
  const uint32_t ks[3] = {k0, k1, k2};
  const int rot[2][4] = {{13, 15, 26, 6}, {17, 29, 16, 24}};
  uint32_t a = x0 + k0, b = x1 + k1;
#pragma unroll
  for (int blk = 0; blk < 5; ++blk) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      a += b;
      b = rotl32(b, rot[blk & 1][r]);
      b ^= a;
    }
    a += ks[(blk + 1) % 3];
    b += ks[(blk + 2) % 3] + (uint32_t)(blk + 1);
  }
  o0 = a;
  o1 = b;
}

// One 64-bit word j of a flat draw of n words from key (khi, klo): counters (j, n + j).
__device__ __forceinline__ uint64_t bits64(uint32_t khi, uint32_t klo, uint32_t n, uint32_t j) {
  uint32_t y0, y1;
  threefry2x32(khi, klo, j, n + j, y0, y1);
  return ((uint64_t)y0 << 32) | (uint64_t)y1;
}

// randint(PRNGKey(seed), (5,), 0, 1000), element j (reference streamsculptor/main.py:221).
__device__ int64_t randint1000(int64_t seed, int j) {
  const uint32_t khi = (uint32_t)((uint64_t)seed >> 32), klo = (uint32_t)seed;
  uint32_t a0, a1, b0, b1;
  threefry2x32(khi, klo, 0u, 2u, a0, a1);  // split: k1 = (a0, b0), k2 = (a1, b1)
  threefry2x32(khi, klo, 1u, 3u, b0, b1);
  const uint64_t span = 1000;
  uint64_t mult = ((uint64_t)1 << 32) % span;
  mult = (mult * mult) % span;
  const uint64_t hb = bits64(a0, b0, 5u, (uint32_t)j);
  const uint64_t lb = bits64(a1, b1, 5u, (uint32_t)j);
  return (int64_t)(((hb % span) * mult + (lb % span)) % span);
}

// Inverse error function: Winitzki's closed form as a start, three Halley steps on erf
// (|u| < 0.5) or erfc (tails, where 1 - |u| is exact).
__device__ double erfinv_d(double u) {
  if (u == 0.0) return 0.0;
  if (u >= 1.0) return __builtin_inf();
  if (u <= -1.0) return -__builtin_inf();
  const double sgn = u < 0 ? -1.0 : 1.0;
  const double a = fabs(u);
  const double A = 0.147;
  const double ln = log((1.0 - a) * (1.0 + a));
  const double t1 = 2.0 / (M_PI * A) + 0.5 * ln;
  double x = sqrt(sqrt(t1 * t1 - ln / A) - t1);
  const double c = 1.1283791670955126;  // 2 / sqrt(pi)
  for (int it = 0; it < 3; ++it) {
    const double f = (a < 0.5) ? (erf(x) - a) : ((1.0 - a) - erfc(x));
    const double fp = c * exp(-x * x);
    x = x - f / (fp + x * f);
  }
  return sgn * x;
}

// normal(PRNGKey(key_seed), (1,)) in float64.
__device__ double prng_normal(int64_t key_seed) {
  const uint32_t khi = (uint32_t)((uint64_t)key_seed >> 32), klo = (uint32_t)key_seed;
  const uint64_t bits = bits64(khi, klo, 1u, 0u);
  const double f = __longlong_as_double((long long)((bits >> 12) | 0x3FF0000000000000ull)) - 1.0;
  const double lo = -0.99999999999999988898;  // nextafter(-1, 0)
  double u = f * 2.0 + lo;
  u = u < lo ? lo : u;
  return 1.4142135623730951 * erfinv_d(u);
}

struct KVals {
  double k[8];
};

// Geometry of one release: the progenitor row, its frame (r-hat, z-hat along L, phi-hat along
// the tangential velocity), the tidal radius and the release speed scale Omega r_t.
struct ReleaseFrame {
  double x[3], v[3], rh[3], zh[3], ph[3], rt, relv;
};

// Release i happens at time t_rel[i] (LINE sources) and source-table row step_rel[i] (TRACK
// sources); the point sources add their Hessian to the host's, since the reference takes
// d2phidr2 of the whole potential.  A TRACK row outside the table gives NaN rows.
__device__ void release_frame(const ssc_host_potential& hp, const ssc_sources& src,
                              const double* __restrict__ prog_w, const double* __restrict__ t_rel,
                              const int32_t* __restrict__ step_rel,
                              const double* __restrict__ g_msat, int i, ReleaseFrame& f) {
  const double* w = prog_w + 6 * (int64_t)i;
  const double x[3] = {w[0], w[1], w[2]}, v[3] = {w[3], w[4], w[5]};
  double phi, g[3], H[9];
  host_full(hp, x[0], x[1], x[2], phi, g, H);
  if (src.kind != SSC_SRC_NONE) {
    const int step = step_rel ? step_rel[i] : 0;
    const double t = t_rel ? t_rel[i] : 0.0;
    if ((src.kind == SSC_SRC_TRACK || src.kind == SSC_SRC_TRACK_HERNQUIST) &&
        (step < 0 || step >= src.n_track)) {
#pragma unroll
      for (int k = 0; k < 9; ++k) H[k] = __builtin_nan("");
    } else {
      for (int j = 0; j < src.n; ++j) {
        double sx, sy, sz;
        if (!source_position(src, j, step, t, sx, sy, sz)) continue;
        const double rs = src.rs[j];
        if (src.kind == SSC_SRC_LINE_HERNQUIST || src.kind == SSC_SRC_TRACK_HERNQUIST)
          hernquist_full(x[0] - sx, x[1] - sy, x[2] - sz, src.gm[j], rs, phi, g, H);
        else
          plummer_full(x[0] - sx, x[1] - sy, x[2] - sz, src.gm[j], rs * rs, phi, g, H);
      }
    }
  }
  const double rad = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  const double rh[3] = {x[0] / rad, x[1] / rad, x[2] / rad};
  // d2phi/dr2 = rhat^T H rhat  (reference main.py:77-85)
  double d2 = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) d2 += rh[a] * H[3 * a + b] * rh[b];
  // angular speed (main.py:89-96) and tidal radius (main.py:99-104)
  const double L[3] = {x[1] * v[2] - x[2] * v[1], x[2] * v[0] - x[0] * v[2],
                       x[0] * v[1] - x[1] * v[0]};
  const double r2 = rad * rad;
  const double om = sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]) / r2;
  f.rt = cbrt(g_msat[i] / (om * om - d2));
  f.relv = om * f.rt;
  const double Ln = sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
  const double vr = v[0] * rh[0] + v[1] * rh[1] + v[2] * rh[2];
  const double pv[3] = {v[0] - vr * rh[0], v[1] - vr * rh[1], v[2] - vr * rh[2]};
  const double pn = sqrt(pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2]);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    f.x[a] = x[a];
    f.v[a] = v[a];
    f.rh[a] = rh[a];
    f.zh[a] = L[a] / Ln;
    f.ph[a] = pv[a] / pn;
  }
}

// The two released rows of frame f with release index ridx and seed: Gaussian multipliers
// with keys PRNGKey(ridx * r_k) (main.py:220-228,263-266); trailing arm = far, leading arm =
// close (main.py:269-278).  pc/pf/vc/vf may alias the strided slots of one output row.
__device__ void release_rows(const ReleaseFrame& f, int64_t ridx, int64_t seed, const KVals& kv,
                             double* pc, double* pf, double* vc, double* vf) {
  double zr[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) zr[k] = prng_normal(ridx * randint1000(seed, k));
  const double kr = kv.k[0] + zr[0] * kv.k[4];
  const double kvphi = kr * (kv.k[1] + zr[1] * kv.k[5]);
  const double kz = kv.k[2] + zr[2] * kv.k[6];
  const double kvz = kv.k[3] + zr[3] * kv.k[7];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double ptr = f.x[a] + kr * f.rh[a] * f.rt + f.zh[a] * kz * f.rt;
    const double vtr = f.v[a] + kvphi * f.relv * f.ph[a] + kvz * f.relv * f.zh[a];
    const double pld = f.x[a] - kr * f.rh[a] * f.rt - f.zh[a] * kz * f.rt;
    const double vld = f.v[a] - kvphi * f.relv * f.ph[a] - kvz * f.relv * f.zh[a];
    pc[a] = pld;
    pf[a] = ptr;
    vc[a] = vld;
    vf[a] = vtr;
  }
}

__global__ __launch_bounds__(256) void release_kernel(ssc_host_potential hp, ssc_sources src,
                                                      const double* __restrict__ prog_w,
                                                      const int32_t* __restrict__ rel_idx,
                                                      const double* __restrict__ t_rel,
                                                      const int32_t* __restrict__ step_rel,
                                                      const double* __restrict__ g_msat, int nrel,
                                                      int64_t seed, KVals kv,
                                                      double* __restrict__ pos_close,
                                                      double* __restrict__ pos_far,
                                                      double* __restrict__ vel_close,
                                                      double* __restrict__ vel_far) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrel) return;
  ReleaseFrame f;
  release_frame(hp, src, prog_w, t_rel, step_rel, g_msat, i, f);
  const int64_t ridx = rel_idx ? (int64_t)rel_idx[i] : (int64_t)i;
  const int64_t o = 3 * (int64_t)i;
  release_rows(f, ridx, seed, kv, pos_close + o, pos_far + o, vel_close + o, vel_far + o);
}

// nseed seeds of the same releases: a block forms the frames of RS_REL releases once (the
// tidal radius needs the Hessian of the whole potential, a sum over every point source), then
// its threads draw the rows of every (release, seed) pair from them.
// rows[s][arm][i][6] = (position, velocity), arm 0 = close (leading), 1 = far (trailing).
constexpr int RS_REL = 32;
__global__ __launch_bounds__(256) void release_seeds_kernel(
    ssc_host_potential hp, ssc_sources src, const double* __restrict__ prog_w,
    const int32_t* __restrict__ rel_idx, const double* __restrict__ t_rel,
    const int32_t* __restrict__ step_rel, const double* __restrict__ g_msat, int nrel,
    int64_t seed0, int nseed, KVals kv, double* __restrict__ rows) {
  __shared__ ReleaseFrame fr[RS_REL];
  const int i0 = blockIdx.x * RS_REL;
  const int nb = min(RS_REL, nrel - i0);
  if ((int)threadIdx.x < nb)
    release_frame(hp, src, prog_w, t_rel, step_rel, g_msat, i0 + threadIdx.x, fr[threadIdx.x]);
  __syncthreads();
  const int64_t arm = (int64_t)nrel * 6, seed_stride = 2 * arm;
  for (int it = threadIdx.x; it < nb * nseed; it += blockDim.x) {
    const int r = it % nb, s = it / nb, i = i0 + r;
    const int64_t ridx = rel_idx ? (int64_t)rel_idx[i] : (int64_t)i;
    double* close = rows + s * seed_stride + 6 * (int64_t)i;
    double* far = close + arm;
    release_rows(fr[r], ridx, seed0 + s, kv, close, far, close + 3, far + 3);
  }
}

// ------------------------------------------------------------------------------------------
// Vector-Jacobian product of the release model.  The Jacobian is formed column by column in
// forward mode (one dual tangent per input: the progenitor row and (G m, p1) of every host
// component); the tidal radius depends on the host Hessian, so its derivative carries third
// derivatives of the potential.  Restates release_func_jacobian (reference
// streamsculptor/perturbative.py:282-296) with the host parameters as extra inputs.
// ------------------------------------------------------------------------------------------
struct D1 {
  double v, d;
};
__device__ __forceinline__ D1 mk(double v) { return D1{v, 0.0}; }
__device__ __forceinline__ D1 operator+(D1 a, D1 b) { return D1{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ D1 operator-(D1 a, D1 b) { return D1{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ D1 operator-(D1 a) { return D1{-a.v, -a.d}; }
__device__ __forceinline__ D1 operator*(D1 a, D1 b) { return D1{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ D1 operator/(D1 a, D1 b) {
  const double q = a.v / b.v;
  return D1{q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ D1 operator*(double a, D1 b) { return D1{a * b.v, a * b.d}; }
__device__ __forceinline__ D1 operator+(D1 a, double b) { return D1{a.v + b, a.d}; }
__device__ __forceinline__ D1 dsqrt(D1 a) {
  const double r = sqrt(a.v);
  return D1{r, a.d / (2.0 * r)};
}
__device__ __forceinline__ D1 dlog1p(D1 a) { return D1{log1p(a.v), a.d / (1.0 + a.v)}; }
__device__ __forceinline__ D1 dcbrt(D1 a) {
  const double c = cbrt(a.v);
  return D1{c, a.d / (3.0 * c * c)};
}

// Hessian of one component (full 3x3) for dual arguments.
__device__ void comp_hess_d(int kind, const D1* p, const D1* x, D1* H) {
  const D1 gm = p[0];
  if (kind == SSC_COMP_MIYAMOTO_NAGAI) {
    const D1 a = p[1], b = p[2];
    const D1 zeta = dsqrt(x[2] * x[2] + b * b);
    const D1 S = a + zeta;
    const D1 D2 = x[0] * x[0] + x[1] * x[1] + S * S;
    const D1 D = dsqrt(D2);
    const D1 iD3 = mk(1.0) / (D2 * D);
    const D1 iD5 = iD3 / D2;
    const D1 q[3] = {x[0], x[1], x[2] * S / zeta};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) H[3 * i + j] = H[3 * i + j] - 3.0 * (gm * q[i] * q[j] * iD5);
    H[0] = H[0] + gm * iD3;
    H[4] = H[4] + gm * iD3;
    const D1 z2 = zeta * zeta;
    const D1 ddz = S * b * b / (z2 * zeta) + x[2] * x[2] / z2;
    H[8] = H[8] + gm * ddz * iD3;
    return;
  }
  D1 r, d1, d2;
  if (kind == SSC_COMP_NFW) {
    const D1 rs = p[1];
    r = dsqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    const D1 L = dlog1p(r / rs);
    const D1 rp = r + rs;
    d1 = gm * (L / (r * r) - mk(1.0) / (r * rp));
    d2 = gm * (mk(1.0) / (r * r * rp) - 2.0 * L / (r * r * r) + (2.0 * r + rs) / (r * r * rp * rp));
  } else if (kind == SSC_COMP_HERNQUIST) {
    const D1 c = p[1], soft = p[2];
    r = dsqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + soft);
    const D1 rc = r + c;
    d1 = gm / (rc * rc);
    d2 = -2.0 * gm / (rc * rc * rc);
  } else if (kind == SSC_COMP_PLUMMER) {
    const D1 rs = p[1];
    const D1 s2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + rs * rs;
    const D1 sq = dsqrt(s2);
    const D1 f = gm / (s2 * sq);
    const D1 k = -3.0 * f / s2;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        D1 hh = k * x[i] * x[j];
        if (i == j) hh = hh + f;
        H[3 * i + j] = H[3 * i + j] + hh;
      }
    return;
  } else {
    return;
  }
  const D1 d1r = d1 / r;
  const D1 c = (d2 - d1r) / (r * r);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      D1 hh = c * x[i] * x[j];
      if (i == j) hh = hh + d1r;
      H[3 * i + j] = H[3 * i + j] + hh;
    }
}

// Release model with dual inputs: out = [pos_close(3), pos_far(3), vel_close(3), vel_far(3)].
__device__ void release_d(const ssc_host_potential& hp, const D1* hostp, const D1* in, D1 gms,
                          const double* zr, const KVals& kv, D1* out) {
  const D1 x[3] = {in[0], in[1], in[2]}, v[3] = {in[3], in[4], in[5]};
  D1 H[9];
  for (int k = 0; k < 9; ++k) H[k] = mk(0.0);
  for (int c = 0; c < hp.n_components; ++c) comp_hess_d(hp.c[c].kind, hostp + 4 * c, x, H);
  const D1 rad = dsqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  const D1 rh[3] = {x[0] / rad, x[1] / rad, x[2] / rad};
  D1 d2 = mk(0.0);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) d2 = d2 + rh[a] * H[3 * a + b] * rh[b];
  const D1 L[3] = {x[1] * v[2] - x[2] * v[1], x[2] * v[0] - x[0] * v[2], x[0] * v[1] - x[1] * v[0]};
  const D1 Ln = dsqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
  const D1 om = Ln / (rad * rad);
  const D1 rt = dcbrt(gms / (om * om - d2));
  const D1 relv = om * rt;
  const D1 zh[3] = {L[0] / Ln, L[1] / Ln, L[2] / Ln};
  const D1 vr = v[0] * rh[0] + v[1] * rh[1] + v[2] * rh[2];
  const D1 pv[3] = {v[0] - vr * rh[0], v[1] - vr * rh[1], v[2] - vr * rh[2]};
  const D1 pn = dsqrt(pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2]);
  const double kr = kv.k[0] + zr[0] * kv.k[4];
  const double kvphi = kr * (kv.k[1] + zr[1] * kv.k[5]);
  const double kz = kv.k[2] + zr[2] * kv.k[6];
  const double kvz = kv.k[3] + zr[3] * kv.k[7];
  for (int a = 0; a < 3; ++a) {
    const D1 ph = pv[a] / pn;
    const D1 dpos = kr * (rh[a] * rt) + kz * (zh[a] * rt);
    const D1 dvel = kvphi * (relv * ph) + kvz * (relv * zh[a]);
    out[a] = x[a] - dpos;
    out[3 + a] = x[a] + dpos;
    out[6 + a] = v[a] - dvel;
    out[9 + a] = v[a] + dvel;
  }
}

__global__ __launch_bounds__(128) void release_vjp_kernel(
    ssc_host_potential hp, const double* __restrict__ prog_w, const int32_t* __restrict__ rel_idx,
    const double* __restrict__ g_msat, int nrel, int64_t seed, KVals kv,
    const double* __restrict__ gpc, const double* __restrict__ gpf,
    const double* __restrict__ gvc, const double* __restrict__ gvf, double* __restrict__ g_prog,
    double* __restrict__ g_host_rel) {
  // one thread per (release, input direction): the 6 + 2 n_components dual-number columns of
  // a release are independent evaluations, so they run side by side rather than in sequence
  const int nc = hp.n_components;
  const int nin = 6 + 2 * nc;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = tid / nin;
  const int dir = (int)(tid - i * nin);
  if (i >= nrel) return;
  double zr[4];
  const int64_t ridx = rel_idx ? (int64_t)rel_idx[i] : (int64_t)i;
  for (int k = 0; k < 4; ++k) zr[k] = prng_normal(ridx * randint1000(seed, k));
  double gout[12];
  for (int a = 0; a < 3; ++a) {
    gout[a] = gpc[3 * (int64_t)i + a];
    gout[3 + a] = gpf[3 * (int64_t)i + a];
    gout[6 + a] = gvc[3 * (int64_t)i + a];
    gout[9 + a] = gvf[3 * (int64_t)i + a];
  }
  {
    D1 in[6], hostp[4 * SSC_MAX_COMPONENTS], out[12];
    for (int k = 0; k < 6; ++k) in[k] = D1{prog_w[6 * (int64_t)i + k], dir == k ? 1.0 : 0.0};
    for (int c = 0; c < nc; ++c)
      for (int k = 0; k < 4; ++k)
        hostp[4 * c + k] = D1{hp.c[c].p[k], (k < 2 && dir == 6 + 2 * c + k) ? 1.0 : 0.0};
    release_d(hp, hostp, in, mk(g_msat[i]), zr, kv, out);
    double acc = 0.0;
    for (int o = 0; o < 12; ++o) acc += gout[o] * out[o].d;
    if (dir < 6) g_prog[6 * (int64_t)i + dir] = acc;
    else g_host_rel[(int64_t)i * 2 * nc + (dir - 6)] = acc;
  }
}

}  // namespace ssc

using namespace ssc;

extern "C" int ssc_release(const ssc_host_potential* host, const ssc_sources* src,
                           const double* prog_w, const int32_t* rel_idx, const double* t_rel,
                           const int32_t* step_rel, const double* g_msat, int32_t nrel,
                           int64_t seed, const double* kvals, double* pos_close, double* pos_far,
                           double* vel_close, double* vel_far, void* stream) {
  int rc = validate_host(host);
  if (rc) return rc;
  rc = validate_sources(src, 0);
  if (rc) return rc;
  if (nrel < 0) return fail(SSC_E_ARG, "ssc_release: nrel < 0");
  if (nrel == 0) return SSC_OK;
  if (!prog_w || !g_msat || !pos_close || !pos_far || !vel_close || !vel_far)
    return fail(SSC_E_ARG, "ssc_release: NULL array argument");
  const ssc_sources none = {};
  const bool with_src = src && src->kind != SSC_SRC_NONE && src->n > 0;
  if (with_src && (src->kind == SSC_SRC_TRACK || src->kind == SSC_SRC_TRACK_HERNQUIST) &&
      !step_rel)
    return fail(SSC_E_ARG, "ssc_release: TRACK sources need step_rel");
  if (with_src && (src->kind == SSC_SRC_LINE || src->kind == SSC_SRC_LINE_HERNQUIST) && !t_rel)
    return fail(SSC_E_ARG, "ssc_release: LINE sources need t_rel");
  KVals kv;
  const double def[8] = {2.0, 0.3, 0.0, 0.0, 0.4, 0.4, 0.5, 0.5};
  for (int k = 0; k < 8; ++k) kv.k[k] = kvals ? kvals[k] : def[k];
  hipLaunchKernelGGL(release_kernel, dim3((nrel + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     *host, with_src ? *src : none, prog_w, rel_idx, t_rel, step_rel, g_msat,
                     nrel, seed, kv, pos_close, pos_far, vel_close, vel_far);
  return check_launch("ssc_release");
}

extern "C" int ssc_release_seeds(const ssc_host_potential* host, const ssc_sources* src,
                                 const double* prog_w, const int32_t* rel_idx,
                                 const double* t_rel, const int32_t* step_rel,
                                 const double* g_msat, int32_t nrel, int64_t seed0,
                                 int32_t nseed, const double* kvals, double* rows, void* stream) {
  int rc = validate_host(host);
  if (rc) return rc;
  rc = validate_sources(src, 0);
  if (rc) return rc;
  if (nrel < 0 || nseed < 0) return fail(SSC_E_ARG, "ssc_release_seeds: negative size");
  if (nrel == 0 || nseed == 0) return SSC_OK;
  if ((int64_t)nrel * nseed > (int64_t)1 << 31)
    return fail(SSC_E_LIMIT, "ssc_release_seeds: nrel * nseed above 2^31");
  if (!prog_w || !g_msat || !rows) return fail(SSC_E_ARG, "ssc_release_seeds: NULL array argument");
  const ssc_sources none = {};
  const bool with_src = src && src->kind != SSC_SRC_NONE && src->n > 0;
  if (with_src && (src->kind == SSC_SRC_TRACK || src->kind == SSC_SRC_TRACK_HERNQUIST) &&
      !step_rel)
    return fail(SSC_E_ARG, "ssc_release_seeds: TRACK sources need step_rel");
  if (with_src && (src->kind == SSC_SRC_LINE || src->kind == SSC_SRC_LINE_HERNQUIST) && !t_rel)
    return fail(SSC_E_ARG, "ssc_release_seeds: LINE sources need t_rel");
  KVals kv;
  const double def[8] = {2.0, 0.3, 0.0, 0.0, 0.4, 0.4, 0.5, 0.5};
  for (int k = 0; k < 8; ++k) kv.k[k] = kvals ? kvals[k] : def[k];
  hipLaunchKernelGGL(release_seeds_kernel, dim3((nrel + RS_REL - 1) / RS_REL), dim3(256), 0,
                     (hipStream_t)stream, *host, with_src ? *src : none, prog_w, rel_idx, t_rel,
                     step_rel, g_msat, nrel, seed0, nseed, kv, rows);
  return check_launch("ssc_release_seeds");
}

// The release VJP proper: per-release partials of the host parameters go to `scratch`
// ([nrel][2 n_components] doubles, every entry written before it is read), then one block per
// parameter sums them into g_host.
static int release_vjp_launch(const char* name, const ssc_host_potential* host,
                              const double* prog_w, const int32_t* rel_idx, const double* g_msat,
                              int32_t nrel, int64_t seed, const double* kvals,
                              const double* g_pos_close, const double* g_pos_far,
                              const double* g_vel_close, const double* g_vel_far, double* g_prog,
                              double* g_host, double* scratch, void* stream) {
  KVals kv;
  const double def[8] = {2.0, 0.3, 0.0, 0.0, 0.4, 0.4, 0.5, 0.5};
  for (int k = 0; k < 8; ++k) kv.k[k] = kvals ? kvals[k] : def[k];
  hipStream_t st = (hipStream_t)stream;
  const int64_t nthr = (int64_t)nrel * (6 + 2 * host->n_components);
  hipLaunchKernelGGL(release_vjp_kernel, dim3((unsigned)((nthr + 127) / 128)), dim3(128), 0, st,
                     *host, prog_w,
                     rel_idx, g_msat, nrel, seed, kv, g_pos_close, g_pos_far, g_vel_close,
                     g_vel_far, g_prog, scratch);
  if (host->n_components > 0)
    hipLaunchKernelGGL(column_sum_kernel<256>, dim3(2 * host->n_components), dim3(256), 0, st,
                       scratch, (int64_t)nrel, 2 * host->n_components, g_host);
  return check_launch(name);
}

static int release_vjp_check(const char* name, const ssc_host_potential* host,
                             const double* prog_w, const double* g_msat, int32_t nrel,
                             const double* g_pos_close, const double* g_pos_far,
                             const double* g_vel_close, const double* g_vel_far,
                             const double* g_prog, const double* g_host) {
  int rc = validate_host(host);
  if (rc) return rc;
  if (nrel <= 0) return fail(SSC_E_ARG, std::string(name) + ": nrel must be positive");
  if (!prog_w || !g_msat || !g_pos_close || !g_pos_far || !g_vel_close || !g_vel_far || !g_prog ||
      !g_host)
    return fail(SSC_E_ARG, std::string(name) + ": NULL array argument");
  return SSC_OK;
}

extern "C" int ssc_release_vjp(const ssc_host_potential* host, const double* prog_w,
                               const int32_t* rel_idx, const double* g_msat, int32_t nrel,
                               int64_t seed, const double* kvals, const double* g_pos_close,
                               const double* g_pos_far, const double* g_vel_close,
                               const double* g_vel_far, double* g_prog, double* g_host,
                               double* scratch, void* stream) {
  int rc = release_vjp_check("ssc_release_vjp", host, prog_w, g_msat, nrel, g_pos_close,
                             g_pos_far, g_vel_close, g_vel_far, g_prog, g_host);
  if (rc) return rc;
  if (!scratch) return fail(SSC_E_ARG, "ssc_release_vjp: NULL scratch");
  return release_vjp_launch("ssc_release_vjp", host, prog_w, rel_idx, g_msat, nrel, seed, kvals,
                            g_pos_close, g_pos_far, g_vel_close, g_vel_far, g_prog, g_host,
                            scratch, stream);
}

// ------------------------------------------------------------------------------------------
// ssc_release_vjp without a caller workspace and without any allocation: the per-release
// partials live in a file-scope device array of SSC_RELEASE_VJP_NOALLOC_MAX_NREL rows of
// 2 SSC_MAX_COMPONENTS doubles (8 MiB, held by every GPU the library runs on until the process
// ends).  Its address is looked up once per device, with that device current, and only that
// address is passed to kernels.  Every call on one GPU shares the array: calls that overlap on
// one GPU race.  Nothing in it needs clearing (the kernel writes every entry it later reads).
// ------------------------------------------------------------------------------------------
static __device__ __align__(256) double scratch[(size_t)SSC_RELEASE_VJP_NOALLOC_MAX_NREL * 2 *
                                                SSC_MAX_COMPONENTS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static int scratch_on_current_device(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(SSC_E_HIP, "hipGetDevice failed");
  if (dev < 0 || dev >= kMaxDevices)
    return fail(SSC_E_LIMIT, "device index beyond " + std::to_string(kMaxDevices));
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SSC_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return SSC_OK;
}

extern "C" int ssc_release_vjp_noalloc(const ssc_host_potential* host, const double* prog_w,
                                       const int32_t* rel_idx, const double* g_msat,
                                       int32_t nrel, int64_t seed, const double* kvals,
                                       const double* g_pos_close, const double* g_pos_far,
                                       const double* g_vel_close, const double* g_vel_far,
                                       double* g_prog, double* g_host, void* stream) {
  int rc = release_vjp_check("ssc_release_vjp_noalloc", host, prog_w, g_msat, nrel, g_pos_close,
                             g_pos_far, g_vel_close, g_vel_far, g_prog, g_host);
  if (rc) return rc;
  if (nrel > SSC_RELEASE_VJP_NOALLOC_MAX_NREL)
    return fail(SSC_E_LIMIT, "ssc_release_vjp_noalloc: nrel = " + std::to_string(nrel) +
                                 " exceeds the static scratch of " +
                                 std::to_string(SSC_RELEASE_VJP_NOALLOC_MAX_NREL) +
                                 " releases; use ssc_release_vjp with a workspace");
  double* sp = nullptr;
  rc = scratch_on_current_device(&sp);
  if (rc) return rc;
  return release_vjp_launch("ssc_release_vjp_noalloc", host, prog_w, rel_idx, g_msat, nrel, seed,
                            kvals, g_pos_close, g_pos_far, g_vel_close, g_vel_far, g_prog,
                            g_host, sp, stream);
}
