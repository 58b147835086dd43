// libssc_mi355x core: error plumbing, potential evaluation and the fixed-step leapfrog
// stream kernel (one particle per lane, FP64 state in VGPRs, Plummer source tables in LDS).
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdio>
#include <cstring>
#include <string>

#include "ssc_device.h"
#include "ssc_sources_lds.h"

namespace ssc {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(SSC_E_LAUNCH, std::string(what) + ": kernel launch failed: " + hipGetErrorString(e));
  return SSC_OK;
}

int validate_host(const ssc_host_potential* host) {
  if (!host) return fail(SSC_E_ARG, "host potential descriptor is NULL");
  if (host->n_components < 0 || host->n_components > SSC_MAX_COMPONENTS)
    return fail(SSC_E_ARG, "host potential: n_components out of range [0, 8]");
  for (int c = 0; c < host->n_components; ++c) {
    const int k = host->c[c].kind;
    if (k < SSC_COMP_MIYAMOTO_NAGAI || k > SSC_COMP_PLUMMER)
      return fail(SSC_E_ARG, "host potential: unknown component kind " + std::to_string(k));
  }
  return SSC_OK;
}

int validate_sources(const ssc_sources* src, int32_t nsteps_needed, bool allow_groups) {
  if (!src || src->kind == SSC_SRC_NONE || src->n == 0) return SSC_OK;
  if (src->group_size != 0) {
    if (!allow_groups)
      return fail(SSC_E_ARG, "sources: grouped sources (group_size > 0) are only supported by "
                             "ssc_dopri5");
    if (src->group_size < 0 || src->group_size > SSC_MAX_SOURCES)
      return fail(SSC_E_LIMIT, "sources: group_size out of range [1, " +
                                   std::to_string(SSC_MAX_SOURCES) + "]");
    if (src->n < 0 || src->n % src->group_size != 0)
      return fail(SSC_E_ARG, "sources: n must be a multiple of group_size");
    if (!src->group) return fail(SSC_E_ARG, "sources: group_size > 0 needs the group array");
    if (src->kind != SSC_SRC_STATIC && src->kind != SSC_SRC_LINE &&
        src->kind != SSC_SRC_LINE_HERNQUIST)
      return fail(SSC_E_ARG, "sources: grouped sources must be STATIC, LINE or LINE_HERNQUIST");
  } else if (src->n < 0 || src->n > SSC_MAX_SOURCES) {
    return fail(SSC_E_LIMIT, "sources: n out of range [0, " + std::to_string(SSC_MAX_SOURCES) + "]");
  }
  if (!src->gm || !src->rs || !src->pos) return fail(SSC_E_ARG, "sources: NULL gm/rs/pos");
  if ((src->kind == SSC_SRC_TRACK || src->kind == SSC_SRC_TRACK_HERNQUIST) &&
      src->n_track < nsteps_needed + 1)
    return fail(SSC_E_ARG, "sources: TRACK table has " + std::to_string(src->n_track) +
                               " rows, the integration needs " + std::to_string(nsteps_needed + 1));
  if ((src->kind == SSC_SRC_LINE || src->kind == SSC_SRC_LINE_HERNQUIST) &&
      (!src->vel || !src->t0 || !src->t_window))
    return fail(SSC_E_ARG, "sources: LINE needs vel, t0 and t_window");
  if (src->kind < SSC_SRC_NONE || src->kind > SSC_SRC_TRACK_HERNQUIST)
    return fail(SSC_E_ARG, "sources: unknown kind " + std::to_string(src->kind));
  return SSC_OK;
}

static ssc_sources empty_sources() {
  ssc_sources s;
  memset(&s, 0, sizeof(s));
  return s;
}

// ------------------------------------------------------------------------------------------
// ssc_eval: one point per lane
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void eval_kernel(ssc_host_potential hp, ssc_sources src,
                                                   const double* __restrict__ xyz, int64_t n,
                                                   double t, int step, double* __restrict__ phi_out,
                                                   double* __restrict__ grad_out,
                                                   double* __restrict__ hess_out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  double phi, g[3], H[9];
  host_full(hp, x, y, z, phi, g, H);
  if (src.kind != SSC_SRC_NONE) {
    for (int j = 0; j < src.n; ++j) {
      double sx, sy, sz;
      if (!source_position(src, j, step, t, sx, sy, sz)) continue;
      const double rs = src.rs[j];
      if (src.kind == SSC_SRC_LINE_HERNQUIST || src.kind == SSC_SRC_TRACK_HERNQUIST)
        hernquist_full(x - sx, y - sy, z - sz, src.gm[j], rs, phi, g, H);
      else
        plummer_full(x - sx, y - sy, z - sz, src.gm[j], rs * rs, phi, g, H);
    }
  }
  if (phi_out) phi_out[i] = phi;
  if (grad_out) {
    grad_out[3 * i] = g[0];
    grad_out[3 * i + 1] = g[1];
    grad_out[3 * i + 2] = g[2];
  }
  if (hess_out) {
#pragma unroll
    for (int k = 0; k < 9; ++k) hess_out[9 * i + k] = H[k];
  }
}

// ------------------------------------------------------------------------------------------
// Leapfrog: all lanes of a block walk the global grid together (k = block's earliest start ..
// nsteps-1); a particle kicks/drifts only once k >= its start.  Each lane carries P particles
// so that one LDS read of a source record serves P force terms.  Source tables live in LDS:
//   STATIC: rec[j] = (x, y, z, G*m), rs2[j]           (loaded once)
//   TRACK:  rows kb..kb+chunk of the trajectory table, rec[r][j] = (x, y, z, G*m)
//   LINE:   (x0, v, G*m, r_s^2, t0, t_window) per source, position formed per step
// add_grad(j0, js) sums sources j0, j0 + js, ... so a wave can also split one orbit's sum.
// ------------------------------------------------------------------------------------------
constexpr int LF_BLOCK = 256;

// Acceleration of P particles at grid row k / time t.
template <int SK, int P, int HS>
__device__ __forceinline__ void accel_P(const ssc_host_potential& hp, const SourceLDS<SK>& lds,
                                        int k, double t, const double* x, const double* y,
                                        const double* z, double* ax, double* ay, double* az) {
  double gx[P], gy[P], gz[P];
#pragma unroll
  for (int p = 0; p < P; ++p)
    host_grad_sig<HS, lean_host(SK)>(hp, x[p], y[p], z[p], gx[p], gy[p], gz[p]);
  lds.template add_grad<P>(k, t, x, y, z, gx, gy, gz, 0, 1);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    ax[p] = -gx[p];
    ay[p] = -gy[p];
    az[p] = -gz[p];
  }
}

template <int SK, int P, int HS>
__global__ __launch_bounds__(LF_BLOCK) void leapfrog_kernel(
    ssc_host_potential hp, ssc_sources src, const double* __restrict__ w0,
    const int32_t* __restrict__ start, int64_t n, double t0, double dt, int nsteps,
    double* __restrict__ w_out, const int32_t* __restrict__ save_idx, int nsave,
    double* __restrict__ ys) {
  extern __shared__ __align__(16) char lds_raw[];
  __shared__ int kmin_sh;
  double x[P], y[P], z[P], vx[P], vy[P], vz[P], ax[P], ay[P], az[P];
  int s[P];
  int64_t idx[P];
  bool valid[P];
  int smin = nsteps;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    idx[p] = (int64_t)blockIdx.x * (LF_BLOCK * P) + p * LF_BLOCK + threadIdx.x;
    valid[p] = idx[p] < n;
    s[p] = nsteps;
    x[p] = 1.0; y[p] = 0.0; z[p] = 0.0; vx[p] = vy[p] = vz[p] = 0.0;
    if (valid[p]) {
      int sp = start ? start[idx[p]] : 0;
      s[p] = sp < 0 ? 0 : (sp > nsteps ? nsteps : sp);
      const double* w = w0 + 6 * idx[p];
      x[p] = w[0]; y[p] = w[1]; z[p] = w[2]; vx[p] = w[3]; vy[p] = w[4]; vz[p] = w[5];
    }
    smin = min(smin, s[p]);
  }
  if (threadIdx.x == 0) kmin_sh = INT_MAX;
  __syncthreads();
  atomicMin(&kmin_sh, smin);
  SourceLDS<SK> lds;
  lds.init(src, lds_raw);
  __syncthreads();
  const int kmin = kmin_sh;
  // Save rows at grid indices <= kmin hold the initial state of every particle.
  int next = 0;
  while (next < nsave && save_idx[next] <= kmin) {
#pragma unroll
    for (int p = 0; p < P; ++p)
      if (valid[p]) {
        double* o = ys + ((int64_t)next * n + idx[p]) * 6;
        o[0] = x[p]; o[1] = y[p]; o[2] = z[p]; o[3] = vx[p]; o[4] = vy[p]; o[5] = vz[p];
      }
    ++next;
  }
  const double h = 0.5 * dt;
  if (kmin < nsteps) {
    lds.stage(src, kmin, nsteps + 1);
    accel_P<SK, P, HS>(hp, lds, kmin, t0 + kmin * dt, x, y, z, ax, ay, az);
  }
  // A particle that has not started keeps its initial row; the force evaluated there at
  // t_{k+1} is exactly the initial force it needs when it starts at k + 1.
  for (int k = kmin; k < nsteps; ++k) {
    lds.stage(src, k, nsteps + 1);
#pragma unroll
    for (int p = 0; p < P; ++p)
      if (k >= s[p]) {
        vx[p] = vx[p] + h * ax[p]; vy[p] = vy[p] + h * ay[p]; vz[p] = vz[p] + h * az[p];
        x[p] = x[p] + dt * vx[p]; y[p] = y[p] + dt * vy[p]; z[p] = z[p] + dt * vz[p];
      }
    accel_P<SK, P, HS>(hp, lds, k + 1, t0 + (k + 1) * dt, x, y, z, ax, ay, az);
#pragma unroll
    for (int p = 0; p < P; ++p)
      if (k >= s[p]) {
        vx[p] = vx[p] + h * ax[p]; vy[p] = vy[p] + h * ay[p]; vz[p] = vz[p] + h * az[p];
      }
    // block-uniform: every particle is now at grid index k + 1 (or holds its initial row)
    while (next < nsave && save_idx[next] == k + 1) {
#pragma unroll
      for (int p = 0; p < P; ++p)
        if (valid[p]) {
          double* o = ys + ((int64_t)next * n + idx[p]) * 6;
          o[0] = x[p]; o[1] = y[p]; o[2] = z[p]; o[3] = vx[p]; o[4] = vy[p]; o[5] = vz[p];
        }
      ++next;
    }
  }
#pragma unroll
  for (int p = 0; p < P; ++p)
    if (valid[p]) {
      double* o = w_out + 6 * idx[p];
      o[0] = x[p]; o[1] = y[p]; o[2] = z[p]; o[3] = vx[p]; o[4] = vy[p]; o[5] = vz[p];
    }
}

// G lanes per orbit: for batches too small to fill the chip one orbit per lane (1e5 particles
// are 1.5 waves per SIMD), G consecutive lanes split the orbit's source sum (lane r takes
// sources r, r + G, ...), combine it with a butterfly over the group and all carry the same
// orbit state.  Block-uniform grid walk and TRACK staging as in leapfrog_kernel.
template <int SK, int G, int HS>
__device__ __forceinline__ void accel_group(const ssc_host_potential& hp, const SourceLDS<SK>& lds,
                                            int k, double t, int r, double x, double y, double z,
                                            double& ax, double& ay, double& az) {
  double gx = 0.0, gy = 0.0, gz = 0.0;
  lds.template add_grad<1>(k, t, &x, &y, &z, &gx, &gy, &gz, r, G);
  gx = group_sum<G>(gx);
  gy = group_sum<G>(gy);
  gz = group_sum<G>(gz);
  double hx, hy, hz;
  host_grad_sig<HS, lean_host(SK)>(hp, x, y, z, hx, hy, hz);
  ax = -(hx + gx);
  ay = -(hy + gy);
  az = -(hz + gz);
}

template <int SK, int G, int HS>
__global__ __launch_bounds__(LF_BLOCK) void leapfrog_group_kernel(
    ssc_host_potential hp, ssc_sources src, const double* __restrict__ w0,
    const int32_t* __restrict__ start, int64_t n, double t0, double dt, int nsteps,
    double* __restrict__ w_out, const int32_t* __restrict__ save_idx, int nsave,
    double* __restrict__ ys) {
  extern __shared__ __align__(16) char lds_raw[];
  __shared__ int kmin_sh;
  const int r = threadIdx.x % G;
  const int64_t i = (int64_t)blockIdx.x * (LF_BLOCK / G) + threadIdx.x / G;
  const bool valid = i < n, lead = valid && r == 0;
  int s = nsteps;
  double x = 1.0, y = 0.0, z = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
  if (valid) {
    const int sp = start ? start[i] : 0;
    s = sp < 0 ? 0 : (sp > nsteps ? nsteps : sp);
    const double* w = w0 + 6 * i;
    x = w[0]; y = w[1]; z = w[2]; vx = w[3]; vy = w[4]; vz = w[5];
  }
  if (threadIdx.x == 0) kmin_sh = INT_MAX;
  __syncthreads();
  atomicMin(&kmin_sh, s);
  SourceLDS<SK> lds;
  lds.init(src, lds_raw);
  __syncthreads();
  const int kmin = kmin_sh;
  int next = 0;
  while (next < nsave && save_idx[next] <= kmin) {
    if (lead) {
      double* o = ys + ((int64_t)next * n + i) * 6;
      o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
    }
    ++next;
  }
  const double h = 0.5 * dt;
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (kmin < nsteps) {
    lds.stage(src, kmin, nsteps + 1);
    accel_group<SK, G, HS>(hp, lds, kmin, t0 + kmin * dt, r, x, y, z, ax, ay, az);
  }
  for (int k = kmin; k < nsteps; ++k) {
    lds.stage(src, k, nsteps + 1);
    const bool act = k >= s;
    if (act) {
      vx = vx + h * ax; vy = vy + h * ay; vz = vz + h * az;
      x = x + dt * vx; y = y + dt * vy; z = z + dt * vz;
    }
    accel_group<SK, G, HS>(hp, lds, k + 1, t0 + (k + 1) * dt, r, x, y, z, ax, ay, az);
    if (act) {
      vx = vx + h * ax; vy = vy + h * ay; vz = vz + h * az;
    }
    while (next < nsave && save_idx[next] == k + 1) {
      if (lead) {
        double* o = ys + ((int64_t)next * n + i) * 6;
        o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
      }
      ++next;
    }
  }
  if (lead) {
    double* o = w_out + 6 * i;
    o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
  }
}

// One wave per orbit: for a handful of orbits in a large source table (the progenitor in a
// clumpy halo) the 64 lanes split the source sum and combine it with a butterfly reduction;
// the orbit state is replicated in every lane.
constexpr int WAVE = 64;

template <int SK, int HS>
__device__ __forceinline__ void accel_wave(const ssc_host_potential& hp, const SourceLDS<SK>& lds,
                                           int k, double t, double x, double y, double z,
                                           double& ax, double& ay, double& az) {
  double gx = 0.0, gy = 0.0, gz = 0.0;
  lds.template add_grad<1>(k, t, &x, &y, &z, &gx, &gy, &gz, threadIdx.x, WAVE);
  gx = group_sum<WAVE>(gx);
  gy = group_sum<WAVE>(gy);
  gz = group_sum<WAVE>(gz);
  double hx, hy, hz;
  host_grad_sig<HS, lean_host(SK)>(hp, x, y, z, hx, hy, hz);
  ax = -(hx + gx);
  ay = -(hy + gy);
  az = -(hz + gz);
}

template <int SK, int HS>
__global__ __launch_bounds__(WAVE) void leapfrog_wave_kernel(
    ssc_host_potential hp, ssc_sources src, const double* __restrict__ w0,
    const int32_t* __restrict__ start, int64_t n, double t0, double dt, int nsteps,
    double* __restrict__ w_out, const int32_t* __restrict__ save_idx, int nsave,
    double* __restrict__ ys) {
  extern __shared__ __align__(16) char lds_raw[];
  const int64_t i = blockIdx.x;
  const bool lead = threadIdx.x == 0;
  int s = start ? start[i] : 0;
  s = s < 0 ? 0 : (s > nsteps ? nsteps : s);
  const double* w = w0 + 6 * i;
  double x = w[0], y = w[1], z = w[2], vx = w[3], vy = w[4], vz = w[5];
  SourceLDS<SK> lds;
  lds.init(src, lds_raw);
  __syncthreads();
  int next = 0;
  while (next < nsave && save_idx[next] <= s) {
    if (lead) {
      double* o = ys + ((int64_t)next * n + i) * 6;
      o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
    }
    ++next;
  }
  const double h = 0.5 * dt;
  double ax = 0, ay = 0, az = 0;
  if (s < nsteps) {
    lds.template stage<8>(src, s, nsteps + 1);
    accel_wave<SK, HS>(hp, lds, s, t0 + s * dt, x, y, z, ax, ay, az);
  }
  for (int k = s; k < nsteps; ++k) {
    lds.template stage<8>(src, k, nsteps + 1);
    vx = vx + h * ax; vy = vy + h * ay; vz = vz + h * az;
    x = x + dt * vx; y = y + dt * vy; z = z + dt * vz;
    accel_wave<SK, HS>(hp, lds, k + 1, t0 + (k + 1) * dt, x, y, z, ax, ay, az);
    vx = vx + h * ax; vy = vy + h * ay; vz = vz + h * az;
    while (next < nsave && save_idx[next] == k + 1) {
      if (lead) {
        double* o = ys + ((int64_t)next * n + i) * 6;
        o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
      }
      ++next;
    }
  }
  if (lead) {
    double* o = w_out + 6 * i;
    o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
  }
}

// Orbits in the analytic host alone, for batches too small to fill the chip, where the step's
// latency sets the rate: the host components are spread over the waves of a block (wave c forms
// component c's gradient for the block's 64 orbits, one per lane; the kinds differ, so lanes of
// one wave could not share the work without running every kind's code), the waves exchange
// their gradient parts through LDS and each accumulates them in component order with the
// operations of host_grad (so the orbits are bit-identical to the one-lane kernels') and
// advances the same orbits.  One barrier per force evaluation (parts double-buffered).
__global__ __launch_bounds__(WAVE * SSC_MAX_COMPONENTS) void leapfrog_comp_kernel(
    ssc_host_potential hp, const double* __restrict__ w0, const int32_t* __restrict__ start,
    int64_t n, double t0, double dt, int nsteps, double* __restrict__ w_out,
    const int32_t* __restrict__ save_idx, int nsave, double* __restrict__ ys) {
  __shared__ double part[2][SSC_MAX_COMPONENTS][WAVE][3];  // f, zm, zadd of each component
  __shared__ int kmin_sh;
  const int l = threadIdx.x & (WAVE - 1), c = threadIdx.x / WAVE;
  const int64_t i = (int64_t)blockIdx.x * WAVE + l;
  const bool valid = i < n;
  const bool writer = valid && c == 0;
  double x = 1.0, y = 0.0, z = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
  int s = nsteps;
  if (valid) {
    const double* w = w0 + 6 * i;
    x = w[0]; y = w[1]; z = w[2]; vx = w[3]; vy = w[4]; vz = w[5];
    const int sp = start ? start[i] : 0;
    s = sp < 0 ? 0 : (sp > nsteps ? nsteps : sp);
  }
  if (threadIdx.x == 0) kmin_sh = INT_MAX;
  __syncthreads();
  atomicMin(&kmin_sh, s);
  __syncthreads();
  const int kmin = kmin_sh;
  const int nc = hp.n_components;
  int buf = 0;
  // block-uniform: every thread calls it the same number of times
  auto accel = [&](double& ax, double& ay, double& az) {
    double f, zm;
    bool zadd;
    comp_grad_parts(hp.c[c], x, y, z, f, zm, zadd);
    part[buf][c][l][0] = f; part[buf][c][l][1] = zm; part[buf][c][l][2] = zadd ? 1.0 : 0.0;
    __syncthreads();
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int q = 0; q < nc; ++q)
      comp_grad_add(part[buf][q][l][0], part[buf][q][l][1], part[buf][q][l][2] != 0.0, x, y, z,
                    gx, gy, gz);
    ax = -gx; ay = -gy; az = -gz;
    buf ^= 1;
  };
  int next = 0;
  while (next < nsave && save_idx[next] <= kmin) {
    if (writer) {
      double* o = ys + ((int64_t)next * n + i) * 6;
      o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
    }
    ++next;
  }
  const double h = 0.5 * dt;
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (kmin < nsteps) accel(ax, ay, az);
  // A particle that has not started keeps its initial row; the force evaluated there is the
  // initial force it needs when it starts.
  for (int k = kmin; k < nsteps; ++k) {
    const bool act = k >= s;
    if (act) {
      vx = vx + h * ax; vy = vy + h * ay; vz = vz + h * az;
      x = x + dt * vx; y = y + dt * vy; z = z + dt * vz;
    }
    accel(ax, ay, az);
    if (act) {
      vx = vx + h * ax; vy = vy + h * ay; vz = vz + h * az;
    }
    while (next < nsave && save_idx[next] == k + 1) {
      if (writer) {
        double* o = ys + ((int64_t)next * n + i) * 6;
        o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
      }
      ++next;
    }
  }
  if (writer) {
    double* o = w_out + 6 * i;
    o[0] = x; o[1] = y; o[2] = z; o[3] = vx; o[4] = vy; o[5] = vz;
  }
}

static int set_lds(const void* fn, size_t lds) {
  if (lds > 160 * 1024) return fail(SSC_E_LIMIT, "ssc_leapfrog: source table exceeds LDS");
  if (lds > 60 * 1024) {
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return fail(SSC_E_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  }
  return SSC_OK;
}

// Launch shape: one wave per orbit for a handful of orbits in a large source table; two
// particles per lane once the batch alone fills the chip; G lanes per orbit for mid-sized
// batches with sources; for up to 65536 orbits in a host of no known signature, host components
// over waves; one particle per lane otherwise (with straight-line host code for the known
// hosts, which is faster still).
template <int SK, int HS>
static int launch_leapfrog(const ssc_host_potential& hp, const ssc_sources& src, const double* w0,
                           const int32_t* start, int64_t n, double t0, double dt, int nsteps,
                           double* w_out, const int32_t* save_idx, int nsave, double* ys,
                           hipStream_t st) {
  const size_t lds = SourceLDS<SK>::bytes(src);
  if (SK != SSC_SRC_NONE && src.n >= 64 && n <= 8192) {
    int rc = set_lds((const void*)leapfrog_wave_kernel<SK, HS>, lds);
    if (rc) return rc;
    hipLaunchKernelGGL((leapfrog_wave_kernel<SK, HS>), dim3((unsigned)n), dim3(WAVE), lds, st, hp, src,
                       w0, start, n, t0, dt, nsteps, w_out, save_idx, nsave, ys);
    return check_launch("ssc_leapfrog");
  }
  if (SK != SSC_SRC_NONE && src.n >= 16 && n >= 393216) {
    int rc = set_lds((const void*)leapfrog_kernel<SK, 2, HS>, lds);
    if (rc) return rc;
    const int64_t blocks = (n + 2 * LF_BLOCK - 1) / (2 * LF_BLOCK);
    hipLaunchKernelGGL((leapfrog_kernel<SK, 2, HS>), dim3((unsigned)blocks), dim3(LF_BLOCK), lds, st,
                       hp, src, w0, start, n, t0, dt, nsteps, w_out, save_idx, nsave, ys);
    return check_launch("ssc_leapfrog");
  }
  if (SK == SSC_SRC_NONE && HS == HS_GENERIC && hp.n_components >= 2 && n <= 65536) {
    const int64_t blocks = (n + WAVE - 1) / WAVE;
    hipLaunchKernelGGL(leapfrog_comp_kernel, dim3((unsigned)blocks),
                       dim3(WAVE * hp.n_components), 0, st, hp, w0, start, n, t0, dt, nsteps,
                       w_out, save_idx, nsave, ys);
    return check_launch("ssc_leapfrog");
  }
  const int g = SK == SSC_SRC_NONE ? 1 : group_lanes(n, src.n);
#define SSC_GROUP(GV)                                                                              \
  if (g == GV) {                                                                                   \
    int rc = set_lds((const void*)leapfrog_group_kernel<SK, GV, HS>, lds);                            \
    if (rc) return rc;                                                                             \
    const int64_t blocks = (n + LF_BLOCK / GV - 1) / (LF_BLOCK / GV);                              \
    hipLaunchKernelGGL((leapfrog_group_kernel<SK, GV, HS>), dim3((unsigned)blocks), dim3(LF_BLOCK),   \
                       lds, st, hp, src, w0, start, n, t0, dt, nsteps, w_out, save_idx, nsave, ys); \
    return check_launch("ssc_leapfrog");                                                          \
  }
  SSC_GROUP(2)
  SSC_GROUP(4)
  SSC_GROUP(8)
  SSC_GROUP(16)
#undef SSC_GROUP
  int rc = set_lds((const void*)leapfrog_kernel<SK, 1, HS>, lds);
  if (rc) return rc;
  const int64_t blocks = (n + LF_BLOCK - 1) / LF_BLOCK;
  hipLaunchKernelGGL((leapfrog_kernel<SK, 1, HS>), dim3((unsigned)blocks), dim3(LF_BLOCK), lds, st, hp,
                     src, w0, start, n, t0, dt, nsteps, w_out, save_idx, nsave, ys);
  return check_launch("ssc_leapfrog");
}

}  // namespace ssc

using namespace ssc;

extern "C" {

const char* ssc_last_error(void) { return g_last_error.c_str(); }

const char* ssc_version(void) { return "libssc_mi355x 0.1 (gfx950)"; }

int ssc_eval(const ssc_host_potential* host, const ssc_sources* src, const double* xyz, int64_t n,
             double t, int32_t step, double* phi, double* grad, double* hess, void* stream) {
  int rc = validate_host(host);
  if (rc) return rc;
  if (n < 0) return fail(SSC_E_ARG, "ssc_eval: n < 0");
  if (n == 0) return SSC_OK;
  if (!xyz) return fail(SSC_E_ARG, "ssc_eval: xyz is NULL");
  rc = validate_sources(src, step < 0 ? 0 : step);
  if (rc) return rc;
  ssc_sources s = src ? *src : empty_sources();
  if (s.n == 0) s.kind = SSC_SRC_NONE;
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(eval_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, *host,
                     s, xyz, n, t, step, phi, grad, hess);
  return check_launch("ssc_eval");
}

int ssc_leapfrog(const ssc_host_potential* host, const ssc_sources* src, const double* w0,
                 const int32_t* start_step, int64_t n, double t0, double dt, int32_t nsteps,
                 double* w_out, const int32_t* save_idx, int32_t nsave, double* ys,
                 void* stream) {
  int rc = validate_host(host);
  if (rc) return rc;
  if (n < 0 || nsteps < 0 || nsave < 0) return fail(SSC_E_ARG, "ssc_leapfrog: negative size");
  if (n == 0) return SSC_OK;
  if (!w0 || !w_out) return fail(SSC_E_ARG, "ssc_leapfrog: w0 or w_out is NULL");
  if (nsave > 0 && (!save_idx || !ys)) return fail(SSC_E_ARG, "ssc_leapfrog: nsave > 0 needs save_idx and ys");
  if (nsave == 0) { save_idx = nullptr; ys = nullptr; }
  if ((n + LF_BLOCK - 1) / LF_BLOCK > (int64_t)INT_MAX)
    return fail(SSC_E_LIMIT, "ssc_leapfrog: too many particles");
  rc = validate_sources(src, nsteps);
  if (rc) return rc;
  ssc_sources s = src ? *src : empty_sources();
  if (s.n == 0) s.kind = SSC_SRC_NONE;
  hipStream_t st = (hipStream_t)stream;
  const int hs = host_signature(*host);
#define SSC_LF3(K, H) launch_leapfrog<K, H>(*host, s, w0, start_step, n, t0, dt, nsteps, w_out, save_idx, nsave, ys, st)
#define SSC_LF(K)                                                                                 \
  (hs == HS_MN_HQ_NFW ? SSC_LF3(K, HS_MN_HQ_NFW)                                                 \
                      : (hs == HS_MN_HQ_HQ_NFW ? SSC_LF3(K, HS_MN_HQ_HQ_NFW) : SSC_LF3(K, HS_GENERIC)))
  switch (s.kind) {
    case SSC_SRC_NONE: return SSC_LF(SSC_SRC_NONE);
    case SSC_SRC_STATIC: return SSC_LF(SSC_SRC_STATIC);
    case SSC_SRC_TRACK: return SSC_LF(SSC_SRC_TRACK);
    case SSC_SRC_LINE: return SSC_LF(SSC_SRC_LINE);
    case SSC_SRC_LINE_HERNQUIST: return SSC_LF(SSC_SRC_LINE_HERNQUIST);
    case SSC_SRC_TRACK_HERNQUIST: return SSC_LF(SSC_SRC_TRACK_HERNQUIST);
    default: return fail(SSC_E_ARG, "ssc_leapfrog: bad source kind");
  }
#undef SSC_LF
#undef SSC_LF3
}

}  // extern "C"
