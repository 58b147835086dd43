/*
 * ssc_mi355x.h -- C-ABI of the MI355X (gfx950) stream-particle integration library.
 *
 * Every entry point takes plain pointers and sizes.  Array arguments are DEVICE pointers
 * (caller-owned, e.g. torch tensors' data_ptr()); descriptor structs are HOST pointers and are
 * copied into the kernel arguments, so they may be freed after the call returns.  Work is
 * enqueued on the caller's HIP stream (`stream`, a hipStream_t; NULL = default stream) and the
 * functions do not synchronise unless stated.  Return value: 0 on success, a negative
 * SSC_E_* code otherwise; ssc_last_error() then returns a thread-local message.
 *
 * Units follow the reference (kpc, Myr, Msun): every mass enters as G*m with
 * G = 4.498502151469553e-12 kpc^3 Msun^-1 Myr^-2, and phase-space rows are AoS
 * [x, y, z, vx, vy, vz] (reference streamsculptor/main.py:117-120).
 *
 * Integration grid: a fixed-step integration runs on the global grid t_k = t0 + k*dt,
 * k = 0..nsteps.  Particle i starts at grid index start_step[i] (NULL = all 0) and ends at
 * k = nsteps; this is how a spray stream integrates each particle from its release time ts[i]
 * to the observation time ts[-1] (reference streamsculptor/main.py:343-368).
 */
#ifndef SSC_MI355X_H
#define SSC_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSC_OK 0
#define SSC_E_ARG (-1)      /* invalid argument (size, pointer, descriptor) */
#define SSC_E_HIP (-2)      /* HIP runtime error */
#define SSC_E_LAUNCH (-3)   /* kernel launch failed */
#define SSC_E_LIMIT (-4)    /* a size exceeds a documented limit */

#define SSC_MAX_COMPONENTS 8

/* Analytic host-potential components.  p[] holds, per kind:
 *   MIYAMOTO_NAGAI: {G*m, a, b, -}     reference streamsculptor/potential.py:66-72
 *   NFW:            {G*m, r_s, -, -}   reference streamsculptor/potential.py:74-84
 *   HERNQUIST:      {G*m, r_s, soft,-} reference streamsculptor/potential.py:132-138
 *   PLUMMER:        {G*m, r_s, -, -}   reference streamsculptor/potential.py:124-130      */
enum {
  SSC_COMP_NONE = 0,
  SSC_COMP_MIYAMOTO_NAGAI = 1,
  SSC_COMP_NFW = 2,
  SSC_COMP_HERNQUIST = 3,
  SSC_COMP_PLUMMER = 4
};

typedef struct ssc_component {
  int32_t kind;
  int32_t reserved;
  double p[4];
} ssc_component;

/* Sum of analytic components: the kernel-side form of Potential_Combine
 * (reference streamsculptor/potential.py:1279-1296). */
typedef struct ssc_host_potential {
  int32_t n_components;
  int32_t reserved;
  ssc_component c[SSC_MAX_COMPONENTS];
} ssc_host_potential;

/* Point sources summed directly (N particles x M sources), Plummer profiles unless stated.
 *   STATIC: pos[n][3], fixed in time (clumps).
 *   TRACK:  pos[n_track][n][3], source positions sampled on the integration grid: row k is the
 *           position at t_k = t0 + k*dt (subhalos on their own orbits).  n_track >= nsteps + 1.
 *   LINE:   straight-line subhalos, position pos[j] + vel[j]*(t - t0[j]), active only while
 *           |t - t0[j]| < t_window[j] (reference SubhaloLinePotential,
 *           streamsculptor/potential.py:802-830).
 *   LINE_HERNQUIST: LINE geometry with Hernquist profiles, Phi = -G m / (r + rs) (reference
 *           SubhaloLinePotentialCustom_fromFunc with func=HernquistPotential,
 *           streamsculptor/potential.py:1161-1213).  Evaluation, leapfrog and Dopri5 only: the
 *           response and adjoint kernels differentiate Plummer sources.
 *   TRACK_HERNQUIST: TRACK geometry with Hernquist profiles: a Hernquist sphere whose centre
 *           follows a track sampled on the integration grid (reference ProgenitorPotential and
 *           TimeDepTranslatingPotential, streamsculptor/potential.py:140-153,448-462, with a
 *           Hernquist profile).  Evaluation, release and leapfrog only.                     */
enum { SSC_SRC_NONE = 0, SSC_SRC_STATIC = 1, SSC_SRC_TRACK = 2, SSC_SRC_LINE = 3,
       SSC_SRC_LINE_HERNQUIST = 4, SSC_SRC_TRACK_HERNQUIST = 5 };

#define SSC_MAX_SOURCES 1024

typedef struct ssc_sources {
  int32_t kind;
  int32_t n;            /* number of sources, 0..SSC_MAX_SOURCES (ungrouped) */
  int32_t n_track;      /* TRACK, TRACK_HERNQUIST: rows in pos */
  int32_t group_size;   /* 0: every orbit sees all n sources.  > 0: the sources form
                           consecutive groups of group_size and orbit i sees only group
                           group[i] -- a batch of independent simulations in one call
                           (ssc_dopri5 only; STATIC, LINE or LINE_HERNQUIST; n may
                           then exceed SSC_MAX_SOURCES) */
  const double* gm;     /* [n] G*m (device) */
  const double* rs;     /* [n] scale radius of the profile (device) */
  const double* pos;    /* STATIC/LINE*: [n][3]; TRACK*: [n_track][n][3] (device) */
  const double* vel;    /* LINE*: [n][3] (device) */
  const double* t0;     /* LINE*: [n] (device) */
  const double* t_window; /* LINE*: [n] (device) */
  const int32_t* group; /* group_size > 0: [orbits] group index of each orbit (device) */
} ssc_sources;

/* Thread-local message for the last failing call on this thread. */
const char* ssc_last_error(void);

/* Library version string. */
const char* ssc_version(void);

/* Potential, gradient and Hessian of host + sources at n points (any output may be NULL).
 *   xyz[n][3], phi[n], grad[n][3], hess[n][9] (row-major, symmetric).
 *   t: evaluation time (LINE sources); step: grid row used for TRACK sources.
 * Replaces Potential.potential / gradient (jax.grad) / jacobian_force (jax.jacfwd of grad)
 * of reference streamsculptor/main.py:37-40,60-65 and the Combine sum potential.py:1285-1296. */
int ssc_eval(const ssc_host_potential* host, const ssc_sources* src,
             const double* xyz, int64_t n, double t, int32_t step,
             double* phi, double* grad, double* hess, void* stream);

/* Fixed-step kick-drift-kick leapfrog for n independent orbits in host + sources.
 *   w0[n][6]       initial rows at grid index start_step[i]
 *   start_step[n]  (NULL = 0); particles with start_step >= nsteps are copied unchanged
 *   w_out[n][6]    rows at grid index nsteps
 *   save_idx[nsave] NULL, or ascending grid indices in [0, nsteps] at which to record the state
 *                  into ys[nsave][n][6]; a particle not yet started records its initial row
 * Particles should be ordered by start_step (a spray stream already is): a block iterates the
 * grid from its earliest start, masking lanes that are not yet released.
 * Fixed-step counterpart of integrate_orbit / integrate_orbit_batch_vmapped /
 * gen_stream_vmapped (reference streamsculptor/main.py:125-202,343-368), which the reference
 * only offers with adaptive diffrax solvers. */
int ssc_leapfrog(const ssc_host_potential* host, const ssc_sources* src,
                 const double* w0, const int32_t* start_step, int64_t n,
                 double t0, double dt, int32_t nsteps, double* w_out,
                 const int32_t* save_idx, int32_t nsave, double* ys, void* stream);

/* Step-size control of ssc_dopri5 (reference PIDController(rtol, atol, dtmin, dtmax,
 * force_dtmin=True, jump_ts), streamsculptor/main.py:146, fields.py:54). */
typedef struct ssc_dopri5_options {
  double rtol;
  double atol;
  double dtmin;         /* <= 0: no floor; steps clamped up to the floor are always accepted */
  double dtmax;         /* <= 0: no ceiling */
  double dt0;           /* <= 0: starting step from the error-based heuristic */
  int32_t max_steps;    /* accepted + rejected steps per orbit before it fails */
  int32_t n_stops;
  const double* stops;  /* [n_stops] device: times at which a step must end (jump_ts) */
} ssc_dopri5_options;

/* Adaptive Dormand-Prince 5(4) for n independent orbits in host + sources (NONE, STATIC or
 * LINE; TRACK is rejected).  Orbit i runs from t_start[i] (device, NULL = t0 for all) to t1,
 * forwards or backwards.
 *   w_out[n][6]      final rows (NULL allowed when ys is given)
 *   ts[nts]          save times (device), monotone in the direction of integration; rows of
 *                    ys[n][nts][6] outside an orbit's interval are NaN
 *   stats[n][3]      NULL or {accepted, rejected, status}; status 1 = max_steps ran out, and
 *                    then every output row of that orbit is +inf (the reference's throw=False);
 *                    status 2 = its source group index is out of range (rows are NaN)
 * Replaces integrate_orbit / integrate_orbit_batch_vmapped / gen_stream_vmapped with
 * solver=Dopri5 (reference streamsculptor/main.py:125-202,343-368). */
int ssc_dopri5(const ssc_host_potential* host, const ssc_sources* src, const double* w0,
               const double* t_start, int64_t n, double t0, double t1, const double* ts,
               int32_t nts, const ssc_dopri5_options* opt, double* w_out, double* ys,
               int32_t* stats, void* stream);

/* ssc_dopri5 with a time interval and save times of its own for every orbit: orbit i runs from
 * t_start[i] to t_end[i] (device, both required; forwards or backwards) and saves at
 * ts[i][0..nts) (device, monotone in its direction of integration) into ys[n][nts][6] (NaN
 * outside the orbit's interval).  Replaces integrate_orbit_batch_vmapped / _batch_scan with a
 * per-orbit (N, M) ts array (reference streamsculptor/main.py:166-202: each orbit integrates
 * from min(ts[i]) to max(ts[i]) unless t0/t1 are given).  opt->n_stops applies to every
 * orbit; other semantics as ssc_dopri5. */
int ssc_dopri5_orbits(const ssc_host_potential* host, const ssc_sources* src, const double* w0,
                      const double* t_start, const double* t_end, int64_t n, const double* ts,
                      int32_t nts, const ssc_dopri5_options* opt, double* w_out, double* ys,
                      int32_t* stats, void* stream);

/* ssc_dopri5 with dense output instead of save times (reference integrate_orbit(dense=True),
 * SaveAt(dense=True), streamsculptor/main.py:126,141, and gen_stream_vmapped_dense,
 * main.py:409-430).  dense[n][dense_cap][20] (device; dense_cap >= max_steps + 1) receives one
 * record per accepted step of orbit i, in tau = sign(t1 - t_start[i]) t:
 *   {tau at the step start, step h, y0[6], f0[6] = dy/dtau at y0, Shampine midpoint value[6]}
 * followed by a closing record {tau1, 0, y1[6], f1[6], 0[6]}; stats[i][0] accepted steps (so
 * stats[i][0] + 1 records).  On step k the quartic through y0, y1 (record k + 1), h f0, h f1 and
 * the midpoint value interpolates the orbit (the same interpolant ssc_dopri5 saves with). */
int ssc_dopri5_dense(const ssc_host_potential* host, const ssc_sources* src, const double* w0,
                     const double* t_start, int64_t n, double t0, double t1,
                     const ssc_dopri5_options* opt, double* w_out, double* dense,
                     int32_t dense_cap, int32_t* stats, void* stream);

/* Spray-stream release conditions (reference Potential.release_model + gen_stream_ics,
 * streamsculptor/main.py:209-306).  For each row i < nrel, from the progenitor row
 * prog_w[i][6] with satellite mass G*Msat = g_msat[i] and release index rel_idx[i]
 * (NULL = i; the index seeds the per-release keys):
 *   pos_close/pos_far/vel_close/vel_far [nrel][3]  (close = leading, far = trailing).
 * The tidal radius uses d2Phi/dr2 of host + src (src NULL or NONE: the host alone), as the
 * reference takes the Hessian of the whole potential (main.py:77-85,230-237); release i is at
 * time t_rel[i] (device; needed by LINE sources) and source-table row step_rel[i] (device;
 * needed by TRACK sources, a row outside the table gives NaN rows).  Both may be NULL
 * otherwise.
 * kvals: NULL selects the reference defaults {2.0, 0.3, 0, 0, 0.4, 0.4, 0.5, 0.5}, else 8
 * host doubles.  Offsets are drawn with the reference's key scheme (PRNGKey(seed) ->
 * randint(5, 0, 1000) -> PRNGKey(i*r_k) -> normal) on a Threefry-2x32-20 counter PRNG.
 * g_msat is G*Msat per release (device). */
int ssc_release(const ssc_host_potential* host, const ssc_sources* src, const double* prog_w,
                const int32_t* rel_idx, const double* t_rel, const int32_t* step_rel,
                const double* g_msat, int32_t nrel, int64_t seed, const double* kvals,
                double* pos_close, double* pos_far, double* vel_close, double* vel_far,
                void* stream);

/* ssc_release for the seeds seed0 .. seed0 + nseed - 1 of the same releases, in one launch: the
 * many-particle spray draws ceil(per_step / 2) seeds per release row, and the reference calls
 * release_model once per seed (gen_stream_ics with seed_num, streamsculptor/main.py:209-306).
 * The tidal radius and frame of a release (the Hessian of the whole potential) are formed once
 * and shared by its seeds.  rows[nseed][2][nrel][6] (device): row i of arm 0 = (pos_close,
 * vel_close), of arm 1 = (pos_far, vel_far) of seed seed0 + s -- the values ssc_release writes
 * for that seed, bit for bit.  Other arguments as in ssc_release; nrel * nseed <= 2^31. */
int ssc_release_seeds(const ssc_host_potential* host, const ssc_sources* src, const double* prog_w,
                      const int32_t* rel_idx, const double* t_rel, const int32_t* step_rel,
                      const double* g_msat, int32_t nrel, int64_t seed0, int32_t nseed,
                      const double* kvals, double* rows, void* stream);

/* Base orbits in the host potential plus first-order responses to each perturbing Plummer
 * source (reference fields.integrate_field with MassRadiusPerturbation_OTF,
 * streamsculptor/fields.py:34-99,159-206, driven per particle as in
 * GenerateMassRadiusPerturbation.compute_perturbation_OTF, perturbative.py:102-135), integrated
 * as the exact derivative of the fixed-step leapfrog map.
 *   pert            the m perturbing sources (STATIC, TRACK or LINE)
 *   w0[n][6]        base rows at grid index start_step[i]
 *   dw0[n][m][D]    initial responses (NULL = zeros), D = 6 (mass block: d(x,v)/d eps_j) or
 *                   12 with with_radius (adds d^2(x,v)/(d a_j d eps_j), a_j = Plummer radius)
 *   w_out[n][6], dw_out[n][m][D] at grid index nsteps (must not alias w0, dw0 when a
 *                   workspace is given)
 *   workspace       NULL, or device memory of workspace_bytes (ssc_perturb_workspace_bytes
 *                   gives a size of at most 2 GiB): the base orbits are then advanced once per
 *                   particle through segments of grid rows, recording position and host Hessian
 *                   per row for the responses; without it every lane that holds responses of a
 *                   particle recomputes them (about twice the arithmetic; the source forces
 *                   are summed in another order, so results agree to the last bits). */
int64_t ssc_perturb_workspace_bytes(int64_t n, int32_t nsteps);
int ssc_perturb_leapfrog(const ssc_host_potential* host, const ssc_sources* pert,
                         const double* w0, const double* dw0, const int32_t* start_step,
                         int64_t n, double t0, double dt, int32_t nsteps, int32_t with_radius,
                         double* w_out, double* dw_out, void* workspace, int64_t workspace_bytes,
                         void* stream);

/* Base orbits plus first-order responses integrated with adaptive Dormand-Prince 5(4): the
 * reference's Runge-Kutta path of GenerateMassRadiusPerturbation (the progenitor's backward
 * field solve, streamsculptor/perturbative.py:52-60, and compute_perturbation_OTF, :95-135,
 * through fields.integrate_field, fields.py:34-99, on MassRadiusPerturbation_OTF.term,
 * :175-206).  Particle i's state [w (6), dw (m, D)] is one ODE with one step controller whose
 * error norm is the RMS over all 6 + m D components (D = 12 with with_radius, else 6); step
 * control as ssc_dopri5 (opt->n_stops must be 0).
 *   pert            the m perturbing sources: STATIC, LINE or LINE_HERNQUIST, not grouped;
 *                   m <= 1024 (512 with with_radius)
 *   w0[n][6], dw0[n][m][D] (NULL = zeros) at t_start[i] (device, NULL = t0 for all), run to t1
 *   w_out[n][6], dw_out[n][m][D]  final values (NULL allowed when saving)
 *   ts[nts], ys[n][nts][6], dys[n][nts][m][D]  saves (quartic dense output; NaN outside an
 *                   orbit's interval)
 *   stats[n][3]     NULL or {accepted, rejected, status}; status 1 = max_steps ran out, and then
 *                   every output of that particle is +inf. */
int ssc_perturb_dopri5(const ssc_host_potential* host, const ssc_sources* pert,
                       const double* w0, const double* dw0, const double* t_start, int64_t n,
                       double t0, double t1, const double* ts, int32_t nts, int32_t with_radius,
                       const ssc_dopri5_options* opt, double* w_out, double* dw_out, double* ys,
                       double* dys, int32_t* stats, void* stream);

/* Vector-Jacobian product of the release model in the host potential alone (no point sources:
 * the release of the reference's perturbed streams, streamhelpers.py:57-66): given
 * dL/d(pos_close, pos_far, vel_close, vel_far) [nrel][3] each, writes dL/d(progenitor row)
 * g_prog[nrel][6] and dL/d(G m_c, p1_c) of
 * the host components into g_host[n_components][2] (summed over releases).  scratch:
 * [nrel][2 * n_components] doubles.  Extends the reference's release-model Jacobian
 * (perturbative.py:282-296) with the host parameters. */
int ssc_release_vjp(const ssc_host_potential* host, const double* prog_w, const int32_t* rel_idx,
                    const double* g_msat, int32_t nrel, int64_t seed, const double* kvals,
                    const double* g_pos_close, const double* g_pos_far,
                    const double* g_vel_close, const double* g_vel_far, double* g_prog,
                    double* g_host, double* scratch, void* stream);

/* ssc_release_vjp with no workspace argument and no allocation of any kind.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The per-release partials live in
 * a static device array of SSC_RELEASE_VJP_NOALLOC_MAX_NREL x 2 SSC_MAX_COMPONENTS doubles
 * (8 MiB per GPU, held until the process ends); nrel above the maximum returns SSC_E_LIMIT.
 * Same arguments and results as ssc_release_vjp otherwise.  Calls that overlap on one GPU
 * share the array and race. */
#define SSC_RELEASE_VJP_NOALLOC_MAX_NREL 65536
int ssc_release_vjp_noalloc(const ssc_host_potential* host, const double* prog_w,
                            const int32_t* rel_idx, const double* g_msat, int32_t nrel,
                            int64_t seed, const double* kvals, const double* g_pos_close,
                            const double* g_pos_far, const double* g_vel_close,
                            const double* g_vel_far, double* g_prog, double* g_host,
                            void* stream);

/* ---------------------------------------------------------------------------------------
 * Discrete adjoint of ssc_leapfrog (no reference counterpart: the reference differentiates
 * forward-mode only, streamsculptor/main.py:160).  Checkpointed: the forward stores the state
 * every checkpoint_every grid steps in the caller's workspace; the backward re-runs each
 * segment and sweeps it in reverse.  checkpoint_every = 1 stores every row (48 B per orbit per
 * grid row: 19.2 GB for 1e5 orbits x 4000 steps) and the backward re-runs nothing.  Sources:
 * NONE, STATIC or TRACK.
 * --------------------------------------------------------------------------------------- */
int64_t ssc_adjoint_workspace_bytes(int64_t n, int32_t nsteps, int32_t checkpoint_every,
                                    int32_t n_src);

/* Forward pass (same result as ssc_leapfrog) keeping the checkpoints in workspace. */
int ssc_adjoint_forward(const ssc_host_potential* host, const ssc_sources* src, const double* w0,
                        const int32_t* start_step, int64_t n, double t0, double dt, int32_t nsteps,
                        int32_t checkpoint_every, double* w_out, void* workspace,
                        int64_t workspace_bytes, void* stream);

/* The checkpoints ssc_adjoint_forward would store, taken from orbits the caller already
 * integrated with every grid row saved: rows[nsteps + 1][n][6] (device; ssc_leapfrog's ys with
 * save_idx = 0..nsteps, bit-identical to what ssc_adjoint_forward computes); n_src as for
 * ssc_adjoint_workspace_bytes.  ssc_adjoint_backward may follow without a forward pass. */
int ssc_adjoint_checkpoints_from_rows(const double* rows, int64_t n, int32_t nsteps,
                                      int32_t checkpoint_every, int32_t n_src, void* workspace,
                                      int64_t workspace_bytes, void* stream);

/* Backward pass after ssc_adjoint_forward with the same arguments and workspace.
 *   lam_T[n][6]    dL/d(final row)
 *   forcing        NULL or [nsteps+1][n][6]: dL/d(row k) of every grid row a particle reaches
 *                  (rows before its start are ignored)
 *   lam_0[n][6]    dL/d(initial row)
 *   g_host[n_components][2]  dL/d(G m_c), dL/d(p1_c) (p1: MN a, NFW r_s, Hernquist c, Plummer r_s)
 *   g_gm[n_src]    NULL or dL/d(G m_j) of the sources
 *   g_pos          NULL or dL/d(source positions): [nsteps+1][n_src][3] (TRACK), [n_src][3] (STATIC)
 * Requires nsteps >= 1. */
int ssc_adjoint_backward(const ssc_host_potential* host, const ssc_sources* src,
                         const int32_t* start_step, int64_t n, double t0, double dt,
                         int32_t nsteps, int32_t checkpoint_every, const double* lam_T,
                         const double* forcing, double* lam_0, double* g_host, double* g_gm,
                         double* g_pos, void* workspace, int64_t workspace_bytes, void* stream);

/* The backward pass in pieces: the grid rows [k_lo, k_hi) only, from k_hi down (k_hi = nsteps
 * includes the final row).  Call with k_hi = nsteps first, then each next call's k_hi = the
 * previous k_lo, down to k_lo = 0 (the adjoint state between calls stays in the workspace);
 * the whole range in one call is ssc_adjoint_backward.  With checkpoint_every > 1, k_lo and
 * k_hi must be multiples of it (or k_hi = nsteps).  Each call writes its range's share: g_host
 * and g_gm (the sums over the range), g_pos (zero outside the range's rows), and lam_0, whose
 * rows are final for the orbits that start at a row >= k_lo.  Forcing rows outside the range
 * are not read.  Lets a caller start work that depends on the late rows' adjoints (a release
 * model, the source orbits) while the early rows are still being swept. */
int ssc_adjoint_backward_range(const ssc_host_potential* host, const ssc_sources* src,
                               const int32_t* start_step, int64_t n, double t0, double dt,
                               int32_t nsteps, int32_t checkpoint_every, int32_t k_lo,
                               int32_t k_hi, const double* lam_T, const double* forcing,
                               double* lam_0, double* g_host, double* g_gm, double* g_pos,
                               void* workspace, int64_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------------------------
 * Clump-force MLP (no counterpart in the reference, which has no network): 3 -> 256 -> 256 ->
 * 256 -> 4 with SiLU, fp32 master parameters, BF16 MFMA products with fp32 accumulation.
 * Parameter vector (floats): W1[256][3] b1[256] W2[256][256] b2[256] W3[256][256] b3[256]
 * W4[4][256] b4[4].  x is fp32 [batch][3] (scaled coordinates), y fp32 [batch][4].
 * --------------------------------------------------------------------------------------- */
int64_t ssc_mlp_num_params(void);
int64_t ssc_mlp_packed_bytes(void);
int64_t ssc_mlp_workspace_bytes(int64_t batch);

/* Training-form forward (activations kept in the caller's workspace). */
int ssc_mlp_forward(const float* params, const float* x, int64_t batch, float* y,
                    void* workspace, int64_t workspace_bytes, void* stream);

/* Forward, weighted squared-error loss sum_b w_b |y_b - t_b|^2 / weight_sum (weight NULL =
 * ones) into loss[0], and the gradient of that loss into grads[num_params]. */
int ssc_mlp_loss_backward(const float* params, const float* x, const float* target,
                          const float* weight, float weight_sum, int64_t batch, float* y,
                          float* grads, float* loss, void* workspace, int64_t workspace_bytes,
                          void* stream);

/* ssc_mlp_loss_backward with the weight sum read from device memory (weight_sum: one float on
 * the device), so a training loop need not bring each minibatch's sum to the host -- a copy
 * that waits for the stream and leaves the launches of every iteration exposed. */
int ssc_mlp_loss_backward_dev(const float* params, const float* x, const float* target,
                              const float* weight, const float* weight_sum, int64_t batch,
                              float* y, float* grads, float* loss, void* workspace,
                              int64_t workspace_bytes, void* stream);

/* Gradient of sum_b dy_b . y_b with respect to the parameters (vector-Jacobian product). */
int ssc_mlp_backward(const float* params, const float* x, const float* dy, int64_t batch,
                     float* grads, void* workspace, int64_t workspace_bytes, void* stream);

/* Adam step with bias correction (step >= 1) and decoupled weight decay. */
int ssc_mlp_adam(float* params, const float* grads, float* m, float* v, int64_t n, float lr,
                 float beta1, float beta2, float eps, int32_t step, float weight_decay,
                 void* stream);

/* Pack parameters into the inference form (ssc_mlp_packed_bytes() bytes, device). */
int ssc_mlp_pack(const float* params, void* packed, void* stream);

/* Inference with the packed form: one wave evaluates 16 points, activations stay in VGPRs. */
int ssc_mlp_infer(const void* packed, const float* x, int64_t batch, float* y, void* stream);

/* Packed transposes W4^T, W3^T, W2^T for the input vector-Jacobian product
 * (ssc_mlp_packed_vjp_bytes() bytes, device). */
int64_t ssc_mlp_packed_vjp_bytes(void);
int ssc_mlp_pack_vjp(const float* params, void* packed_vjp, void* stream);

/* Input vector-Jacobian product dx[b] = d(dy[b] . y(x[b])) / dx[b] (fp32 [batch][3]) with the
 * packed forms: the forward is recomputed in registers and the backward products run on the
 * same BF16 MFMA chain (no counterpart in the reference: the derivative of the learned clump
 * force that the discrete adjoint of the surrogate integrator needs). */
int ssc_mlp_input_vjp(const void* packed, const void* packed_vjp, const float* x, const float* dy,
                      int64_t batch, float* dx, void* stream);

/* Leapfrog of n particles in host + clumps with the clump force from the packed MLP
 * (no counterpart in the reference).  scaling[17] (host doubles): input centre[3], input
 * half-width[3], phi scale, grad-phi scale, residual-grid origin[3], cell size[3], cells per
 * axis[3].  flags[nx][ny][nz] (device, 1 = exact sum): positions in flagged cells or outside the
 * grid use the exact sum over `clumps` (STATIC).  fallback_count (device, may be NULL)
 * accumulates the number of particle-steps that fell back.  Same grid/start/output semantics
 * as ssc_leapfrog. */
int ssc_surrogate_leapfrog(const ssc_host_potential* host, const ssc_sources* clumps,
                           const double* scaling, const void* packed, const uint8_t* flags,
                           const double* w0, const int32_t* start_step, int64_t n, double t0,
                           double dt, int32_t nsteps, double* w_out,
                           unsigned long long* fallback_count, void* stream);

/* Discrete adjoint of ssc_surrogate_leapfrog (no counterpart in the reference: it replaces
 * jax.grad through the stream integration when the clump force is the learned surrogate).
 * Same model arguments as ssc_surrogate_leapfrog; the caller's device workspace holds the
 * checkpoints (ceil(nsteps / checkpoint_every) + 1 rows of (x, v, a) per particle) between the two
 * calls.  The forward also returns w_out.  The backward walks the leapfrog backwards from each
 * checkpoint (the map is explicitly invertible) and applies the clump Hessian through the MLP
 * input VJP (packed_vjp from ssc_mlp_pack_vjp), or the exact clump sum in flagged cells:
 * lam_T[n][6] = dL/dw_out, forcing[(nsteps + 1)][n][6] (may be NULL) = dL/d(row k), outputs
 * lam_0[n][6] = dL/d(w0) and g_host[n_components][2] = dL/d(G m_c), dL/d(p1_c) (host
 * parameters).  fallback_count (device, may be NULL) accumulates the rows that used the exact
 * sum. */
int64_t ssc_surrogate_adjoint_workspace_bytes(int64_t n, int32_t nsteps, int32_t checkpoint_every);
int ssc_surrogate_adjoint_forward(const ssc_host_potential* host, const ssc_sources* clumps,
                                  const double* scaling, const void* packed, const uint8_t* flags,
                                  const double* w0, const int32_t* start_step, int64_t n,
                                  double t0, double dt, int32_t nsteps, int32_t checkpoint_every,
                                  double* w_out, unsigned long long* fallback_count,
                                  void* workspace, int64_t workspace_bytes, void* stream);
int ssc_surrogate_adjoint_backward(const ssc_host_potential* host, const ssc_sources* clumps,
                                   const double* scaling, const void* packed,
                                   const void* packed_vjp, const uint8_t* flags,
                                   const int32_t* start_step, int64_t n, double dt,
                                   int32_t nsteps, int32_t checkpoint_every, const double* lam_T,
                                   const double* forcing, double* lam_0, double* g_host,
                                   unsigned long long* fallback_count, void* workspace,
                                   int64_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SSC_MI355X_H */
