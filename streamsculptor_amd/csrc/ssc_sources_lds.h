// LDS-resident Plummer source tables shared by the leapfrog, response and adjoint kernels.
//   STATIC: geo[j] = (x, y, z, r_s^2), kc[j] = (-1.5 G*m, 2.5 G*m), gmv[j]   (loaded once)
//   TRACK:  rows kb..kb+chunk of the trajectory table, rec[r][j] = (x, y, z, r_s^2), kc[j], gmv[j]
//   LINE:   (x0, v, G*m, r_s^2, t0, t_window) per source, position formed per step
//   LINE_HERNQUIST: the same with r_s in place of r_s^2 and the Hernquist term
//   TRACK_HERNQUIST: TRACK rows with r_s in place of r_s^2 and the Hernquist term
// add_grad(j0, js) sums sources j0, j0 + js, ... so a wave can also split one orbit's sum;
// get(k, t, j, ...) returns one source's position/G*m/r_s^2 and whether it is active.
#pragma once

#include <climits>

#include "ssc_device.h"

namespace ssc {

struct alignas(16) SrcRec {
  double x, y, z, gm;
};

template <int SK>
struct SourceLDS;

// No sources.
template <>
struct SourceLDS<SSC_SRC_NONE> {
  static constexpr int n = 0;
  __device__ void init(const ssc_sources&, char*) {}
  template <int SB = 1>
  __device__ void stage(const ssc_sources&, int, int) {}
  __device__ void stage_row(const ssc_sources&, int, int) {}
  template <int P>
  __device__ void add_grad(int, double, const double*, const double*, const double*, double*,
                           double*, double*, int, int) const {}
  static size_t bytes(const ssc_sources&) { return 0; }
  __device__ bool get(int, double, int, double&, double&, double&, double&, double&) const {
    return false;
  }
};

// STATIC tables keep the geometry and G m in the forms the pair loop consumes:
// geo[j] = (x, y, z, r_s^2), kc[j] = (-1.5 G m, 2.5 G m) for plummer_grad_k, gmv[j] = G m.
struct alignas(16) SrcGeo {
  double x, y, z, rs2;
};
struct alignas(16) SrcK {
  double c1, c2;
};

template <>
struct SourceLDS<SSC_SRC_STATIC> {
  SrcGeo* geo;
  SrcK* kc;
  double* gmv;
  int n;
  static size_t bytes(const ssc_sources& s) {
    return (size_t)s.n * (sizeof(SrcGeo) + sizeof(SrcK) + 8);
  }
  __device__ void init(const ssc_sources& s, char* lds) {
    n = s.n;
    geo = reinterpret_cast<SrcGeo*>(lds);
    kc = reinterpret_cast<SrcK*>(lds + (size_t)n * sizeof(SrcGeo));
    gmv = reinterpret_cast<double*>(lds + (size_t)n * (sizeof(SrcGeo) + sizeof(SrcK)));
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      const double gm = s.gm[j];
      geo[j] = SrcGeo{s.pos[3 * j], s.pos[3 * j + 1], s.pos[3 * j + 2], s.rs[j] * s.rs[j]};
      kc[j] = SrcK{-1.5 * gm, 2.5 * gm};
      gmv[j] = gm;
    }
    __syncthreads();
  }
  template <int SB = 1>
  __device__ void stage(const ssc_sources&, int, int) {}
  __device__ void stage_row(const ssc_sources&, int, int) {}
  template <int P>
  __device__ void add_grad(int, double, const double* x, const double* y, const double* z,
                           double* gx, double* gy, double* gz, int j0, int js) const {
    for (int j = j0; j < n; j += js) {
      const SrcGeo g = geo[j];
      const SrcK c = kc[j];
#pragma unroll
      for (int p = 0; p < P; ++p)
        plummer_grad_k(x[p] - g.x, y[p] - g.y, z[p] - g.z, g.rs2, c.c1, c.c2, gx[p], gy[p], gz[p]);
    }
  }
  __device__ bool get(int, double, int j, double& sx, double& sy, double& sz, double& gm,
                      double& q) const {
    const SrcGeo g = geo[j];
    sx = g.x; sy = g.y; sz = g.z; gm = gmv[j]; q = g.rs2;
    return true;
  }
  // Source j in the folded form of plummer_grad_k: position, r_s^2, (-1.5 G m, 2.5 G m).
  __device__ void get_k(int, int j, double& sx, double& sy, double& sz, double& rs2, double& c1,
                        double& c2) const {
    const SrcGeo g = geo[j];
    const SrcK c = kc[j];
    sx = g.x; sy = g.y; sz = g.z; rs2 = g.rs2; c1 = c.c1; c2 = c.c2;
  }
};

// Grid steps staged per LDS fill for TRACK sources: (chunk + 1) rows of n records stay under
// 64 KiB (16 steps for 64 subhalos, 1 step at 1024 sources).
__host__ __device__ inline int track_chunk(int n) {
  int c = 2048 / (n > 0 ? n : 1) - 1;
  return c < 1 ? 1 : (c > 16 ? 16 : c);
}

// TRACK tables with Plummer (HQ = false: rec q = r_s^2) or Hernquist (HQ = true: q = r_s)
// profiles.
template <bool HQ>
struct TrackLDS {
  SrcGeo* rec;  // [chunk + 1][n]: (x, y, z, q)
  SrcK* kc;     // [n]: (-1.5 G m, 2.5 G m)
  double* gmv;  // [n]: G m
  int n;
  int chunk;
  int kb;       // grid row held in rec[0]
  static size_t bytes(const ssc_sources& s) {
    return (size_t)(track_chunk(s.n) + 1) * s.n * sizeof(SrcGeo) +
           (size_t)s.n * (sizeof(SrcK) + 8);
  }
  __device__ void init(const ssc_sources& s, char* lds) {
    n = s.n;
    chunk = track_chunk(n);
    rec = reinterpret_cast<SrcGeo*>(lds);
    kc = reinterpret_cast<SrcK*>(lds + (size_t)(chunk + 1) * n * sizeof(SrcGeo));
    gmv = reinterpret_cast<double*>(lds + (size_t)(chunk + 1) * n * sizeof(SrcGeo) +
                                    (size_t)n * sizeof(SrcK));
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      const double gm = s.gm[j];
      kc[j] = SrcK{-1.5 * gm, 2.5 * gm};
      gmv[j] = gm;
    }
    kb = INT_MIN;
  }
  // Make rows k and k+1 resident (block-uniform k; contains barriers).  Each thread issues the
  // loads of SB records before its first LDS store: a block of one wave (one orbit; 17 records
  // per lane at 64 sources) then waits for one memory round trip per SB records instead of one
  // per record.  The SB records take 8 SB VGPRs, so blocks of many orbits keep SB = 1.
  template <int SB = 1>
  __device__ void stage(const ssc_sources& s, int k, int nrows) {
    if (kb != INT_MIN && k >= kb && k + 1 <= kb + chunk) return;
    __syncthreads();
    kb = k;
    const int rows = min(chunk + 1, nrows - k);
    const int tot = rows * n;
    for (int e0 = threadIdx.x; e0 < tot; e0 += SB * blockDim.x) {
      SrcGeo v[SB];
#pragma unroll
      for (int b = 0; b < SB; ++b) {
        const int e = e0 + b * (int)blockDim.x;
        if (e < tot) {
          const int r = e / n, j = e - r * n;
          const double* p = s.pos + ((int64_t)(k + r) * n + j) * 3;
          const double rs = s.rs[j];
          v[b] = SrcGeo{p[0], p[1], p[2], HQ ? rs : rs * rs};
        }
      }
#pragma unroll
      for (int b = 0; b < SB; ++b) {
        const int e = e0 + b * (int)blockDim.x;
        if (e < tot) rec[e] = v[b];  // e = r * n + j
      }
    }
    __syncthreads();
  }
  // Make row k resident for a walk towards lower rows (block-uniform k; barriers inside).
  __device__ void stage_row(const ssc_sources& s, int k, int nrows) {
    if (kb != INT_MIN && k >= kb && k <= kb + chunk) return;
    stage(s, k - chunk > 0 ? k - chunk : 0, nrows);
  }
  template <int P>
  __device__ void add_grad(int k, double, const double* x, const double* y, const double* z,
                           double* gx, double* gy, double* gz, int j0, int js) const {
    const SrcGeo* row = rec + (k - kb) * n;
    for (int j = j0; j < n; j += js) {
      const SrcGeo g = row[j];
      const SrcK c = kc[j];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        if (HQ)
          hernquist_grad(x[p] - g.x, y[p] - g.y, z[p] - g.z, gmv[j], g.rs2, gx[p], gy[p], gz[p]);
        else
          plummer_grad_k(x[p] - g.x, y[p] - g.y, z[p] - g.z, g.rs2, c.c1, c.c2, gx[p], gy[p],
                         gz[p]);
      }
    }
  }
  __device__ bool get(int k, double, int j, double& sx, double& sy, double& sz, double& gm,
                      double& q) const {
    const SrcGeo g = rec[(k - kb) * n + j];
    sx = g.x; sy = g.y; sz = g.z; gm = gmv[j]; q = g.rs2;
    return true;
  }
  // Source j at grid row k in the folded form of plummer_grad_k (Plummer tables).
  __device__ void get_k(int k, int j, double& sx, double& sy, double& sz, double& rs2, double& c1,
                        double& c2) const {
    const SrcGeo g = rec[(k - kb) * n + j];
    const SrcK c = kc[j];
    sx = g.x; sy = g.y; sz = g.z; rs2 = g.rs2; c1 = c.c1; c2 = c.c2;
  }
};

template <>
struct SourceLDS<SSC_SRC_TRACK> : TrackLDS<false> {};

template <>
struct SourceLDS<SSC_SRC_TRACK_HERNQUIST> : TrackLDS<true> {};

struct alignas(16) LineRec {
  double x0, y0, z0, gm, vx, vy, vz, q, t0, tw;  // q: r_s^2 (Plummer) or r_s (Hernquist)
};

// Straight-line sources with Plummer (HQ = false) or Hernquist (HQ = true) profiles.
template <bool HQ>
struct LineLDS {
  LineRec* rec;
  int n;
  static size_t bytes(const ssc_sources& s) { return (size_t)s.n * sizeof(LineRec); }
  __device__ void init(const ssc_sources& s, char* lds) {
    n = s.n;
    rec = reinterpret_cast<LineRec*>(lds);
    for (int j = threadIdx.x; j < n; j += blockDim.x)
      rec[j] = LineRec{s.pos[3 * j], s.pos[3 * j + 1], s.pos[3 * j + 2], s.gm[j],
                       s.vel[3 * j], s.vel[3 * j + 1], s.vel[3 * j + 2],
                       HQ ? s.rs[j] : s.rs[j] * s.rs[j], s.t0[j], s.t_window[j]};
    __syncthreads();
  }
  template <int SB = 1>
  __device__ void stage(const ssc_sources&, int, int) {}
  __device__ void stage_row(const ssc_sources&, int, int) {}
  template <int P>
  __device__ void add_grad(int, double t, const double* x, const double* y, const double* z,
                           double* gx, double* gy, double* gz, int j0, int js) const {
    for (int j = j0; j < n; j += js) {
      const LineRec r = rec[j];
      const double dt = t - r.t0;
      if (!(fabs(dt) < r.tw)) continue;
      const double sx = r.x0 + r.vx * dt, sy = r.y0 + r.vy * dt, sz = r.z0 + r.vz * dt;
#pragma unroll
      for (int p = 0; p < P; ++p) {
        if (HQ)
          hernquist_grad(x[p] - sx, y[p] - sy, z[p] - sz, r.gm, r.q, gx[p], gy[p], gz[p]);
        else
          plummer_grad(x[p] - sx, y[p] - sy, z[p] - sz, r.gm, r.q, gx[p], gy[p], gz[p]);
      }
    }
  }
  // One source at time t (the response and adjoint kernels): position, G m, q (r_s^2 for
  // Plummer, r_s for Hernquist), active.
  __device__ bool get(int, double t, int j, double& sx, double& sy, double& sz, double& gm,
                      double& q) const {
    const LineRec r = rec[j];
    const double dt = t - r.t0;
    sx = r.x0 + r.vx * dt; sy = r.y0 + r.vy * dt; sz = r.z0 + r.vz * dt;
    gm = r.gm; q = r.q;
    return fabs(dt) < r.tw;
  }
};

template <>
struct SourceLDS<SSC_SRC_LINE> : LineLDS<false> {};

template <>
struct SourceLDS<SSC_SRC_LINE_HERNQUIST> : LineLDS<true> {};

}  // namespace ssc
