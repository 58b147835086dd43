// Spray-stream release conditions on the GPU: one release per lane.  Restates the reference
// Potential.release_model (streamsculptor/main.py:209-280) with the closed-form host Hessian
// in place of nested autodiff, and the reference's key scheme on a Threefry-2x32-20 counter
// PRNG (Salmon et al., SC'11) for the Gaussian offsets.
#include <hip/hip_runtime.h>

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

__global__ __launch_bounds__(256) void release_kernel(ssc_host_potential hp,
                                                      const double* __restrict__ prog_w,
                                                      const int32_t* __restrict__ rel_idx,
                                                      const double* __restrict__ g_msat, int nrel,
                                                      int64_t seed, KVals kv,
                                                      double* __restrict__ pos_close,
                                                      double* __restrict__ pos_far,
                                                      double* __restrict__ vel_close,
                                                      double* __restrict__ vel_far) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrel) return;
  const double* w = prog_w + 6 * (int64_t)i;
  const double x[3] = {w[0], w[1], w[2]}, v[3] = {w[3], w[4], w[5]};
  double phi, g[3], H[9];
  host_full(hp, x[0], x[1], x[2], phi, g, H);
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
  const double rt = cbrt(g_msat[i] / (om * om - d2));
  const double relv = om * rt;
  const double Ln = sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
  const double zh[3] = {L[0] / Ln, L[1] / Ln, L[2] / Ln};
  const double vr = v[0] * rh[0] + v[1] * rh[1] + v[2] * rh[2];
  const double pv[3] = {v[0] - vr * rh[0], v[1] - vr * rh[1], v[2] - vr * rh[2]};
  const double pn = sqrt(pv[0] * pv[0] + pv[1] * pv[1] + pv[2] * pv[2]);
  const double ph[3] = {pv[0] / pn, pv[1] / pn, pv[2] / pn};
  // Gaussian multipliers with keys PRNGKey(i * r_k) (main.py:220-228,263-266)
  double zr[4];
  const int64_t ridx = rel_idx ? (int64_t)rel_idx[i] : (int64_t)i;
#pragma unroll
  for (int k = 0; k < 4; ++k) zr[k] = prng_normal(ridx * randint1000(seed, k));
  const double kr = kv.k[0] + zr[0] * kv.k[4];
  const double kvphi = kr * (kv.k[1] + zr[1] * kv.k[5]);
  const double kz = kv.k[2] + zr[2] * kv.k[6];
  const double kvz = kv.k[3] + zr[3] * kv.k[7];
  // trailing arm = far, leading arm = close (main.py:269-278)
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double ptr = x[a] + kr * rh[a] * rt + zh[a] * kz * rt;
    const double vtr = v[a] + kvphi * relv * ph[a] + kvz * relv * zh[a];
    const double pld = x[a] - kr * rh[a] * rt - zh[a] * kz * rt;
    const double vld = v[a] - kvphi * relv * ph[a] - kvz * relv * zh[a];
    pos_close[3 * (int64_t)i + a] = pld;
    pos_far[3 * (int64_t)i + a] = ptr;
    vel_close[3 * (int64_t)i + a] = vld;
    vel_far[3 * (int64_t)i + a] = vtr;
  }
}

}  // namespace ssc

using namespace ssc;

extern "C" int ssc_release(const ssc_host_potential* host, const double* prog_w, const int32_t* rel_idx,
                           const double* g_msat, int32_t nrel, int64_t seed, const double* kvals,
                           double* pos_close, double* pos_far, double* vel_close, double* vel_far,
                           void* stream) {
  int rc = validate_host(host);
  if (rc) return rc;
  if (nrel < 0) return fail(SSC_E_ARG, "ssc_release: nrel < 0");
  if (nrel == 0) return SSC_OK;
  if (!prog_w || !g_msat || !pos_close || !pos_far || !vel_close || !vel_far)
    return fail(SSC_E_ARG, "ssc_release: NULL array argument");
  KVals kv;
  const double def[8] = {2.0, 0.3, 0.0, 0.0, 0.4, 0.4, 0.5, 0.5};
  for (int k = 0; k < 8; ++k) kv.k[k] = kvals ? kvals[k] : def[k];
  hipLaunchKernelGGL(release_kernel, dim3((nrel + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     *host, prog_w, rel_idx, g_msat, nrel, seed, kv, pos_close, pos_far, vel_close, vel_far);
  return check_launch("ssc_release");
}
