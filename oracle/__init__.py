"""ORACLE -- test infrastructure only.

CPU FP64 restatement of the reference's stream-particle integration path (C++ in
``oracle/ssc_oracle.cpp``, wrapped here with ctypes + numpy).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import this package, and
only as the reference results are compared against.  The product package (``streamsculptor_amd``) never imports it.

What pins each function (independently of this file) -- see tests/test_oracle_*.py:

* G and units: closed form from IAU/CODATA constants; the reference notebook value
  ``custom_potential.ipynb:132`` to 15 digits.
* Component potentials / gradients / Hessians (MN, NFW, Hernquist, Plummer): sympy
  differentiation of the reference formulas (potential.py:66-84,124-138), central finite
  differences, Laplacian = 4 pi G rho for the analytic densities, and the reference notebook
  NFW acceleration ``tests.ipynb:1671``.
* Leapfrog: time reversibility, the closed-form discrete harmonic oscillator (Plummer core),
  second-order convergence, bounded energy error.
* Threefry-2x32-20: the Random123 known-answer vectors.
* erfinv: scipy.special.erfinv; normal / randint: bit-level restatement of the key scheme
  (JAX bit compatibility of the draws themselves is parity-unpinned: jax is not installable).
* Release model: closed-form geometry on hand-built cases; its Jacobian against central
  finite differences of the release model itself.
* Dopri5: the tableau and dense-output weights against scipy's RK45 arrays; one step and its
  dense output against scipy's own RK step on the same right-hand side; converged orbits
  against scipy DOP853 and the reference notebooks' printed orbits (tests.ipynb:302,1717,1747;
  StreamSubhaloExample.ipynb:35).  The step-size controller (diffrax 0.7.0 PIDController,
  not in the reference tree) is restated from its documentation: its step sequence is
  parity-unpinned; its invariants (dtmin floor, max_steps failure, tolerance convergence)
  are tested.
* Subhalo sums (STATIC, TRACK, LINE Plummer; LINE_HERNQUIST): superposition of single-source
  sums, the Plummer and Hernquist components at the source position, time-window semantics.
* First-order responses (leapfrog): central differences in eps of the oracle leapfrog with the
  subhalo potential scaled by eps (the reference's own test method, tests.ipynb:2959,3182);
  radius block = difference of the mass block in the scale radius; linearity.
* Adaptive responses (perturb_dopri5): inactive sources make it take exactly the steps of the
  orbit oracle with tolerances divided by sqrt(6 / (6 + m D)) (the RMS norm over the whole
  state); at tight tolerances the mass block = central differences of the nonlinear Dopri5
  flow, the radius block = difference of the mass block in the scale radius; saved rows =
  integrations ending at the save times.
* Adjoint: every output against central differences of the forward loss.
* MLP: forward/backward against torch CPU autograd in float64, BF16 rounding against torch's
  bfloat16; the input VJP against torch autograd and central differences.
* Surrogate leapfrog and adjoint (oracle/surrogate.py): with every cell falling back they equal
  the plain oracle leapfrog and adjoint; with a smooth float64 network the adjoint equals
  central differences of the forward map.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libssc_oracle.so")
_FAST_PATH = os.path.join(_HERE, "libssc_oracle_fast.so")

G = 4.498502151469553e-12  # kpc^3 / (Msun Myr^2)

COMP_MN, COMP_NFW, COMP_HERNQUIST, COMP_PLUMMER = 1, 2, 3, 4
SRC_NONE, SRC_STATIC, SRC_TRACK, SRC_LINE, SRC_LINE_HERNQUIST, SRC_TRACK_HERNQUIST = 0, 1, 2, 3, 4, 5
MAX_COMPONENTS = 8


class _Component(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("p", ctypes.c_double * 4)]


class _Host(ctypes.Structure):
    _fields_ = [("n_components", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("c", _Component * MAX_COMPONENTS)]


class _Sources(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("n", ctypes.c_int32), ("n_track", ctypes.c_int32),
                ("group_size", ctypes.c_int32),
                ("gm", ctypes.c_void_p), ("rs", ctypes.c_void_p), ("pos", ctypes.c_void_p),
                ("vel", ctypes.c_void_p), ("t0", ctypes.c_void_p), ("t_window", ctypes.c_void_p),
                ("group", ctypes.c_void_p)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++ + OpenMP)."""
    src = os.path.join(_HERE, "ssc_oracle.cpp")
    stale = [p for p in (_LIB_PATH, _FAST_PATH)
             if not os.path.exists(p) or os.path.getmtime(p) < os.path.getmtime(src)]
    if force or stale:
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


_lib = None
_fast = None


def lib(fast: bool = False):
    """The -O2 build, or (fast=True) the -O3 x86-64-v3 build used to time the CPU
    baseline."""
    global _lib, _fast
    if fast:
        if _fast is None:
            build()
            _fast = _bind(ctypes.CDLL(_FAST_PATH))
        return _fast
    if _lib is None:
        build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def _bind(L):
    vp, i64, i32, d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    L.orc_threefry2x32.argtypes = [ctypes.c_uint32] * 4 + [vp]
    L.orc_randint5_1000.argtypes = [i64, vp]
    L.orc_prng_normal.argtypes = [i64]
    L.orc_prng_normal.restype = d
    L.orc_erfinv.argtypes = [d]
    L.orc_erfinv.restype = d
    L.orc_release_normals.argtypes = [i64, i64, vp]
    L.orc_eval.argtypes = [vp, vp, vp, i64, d, i32, vp, vp, vp]
    L.orc_leapfrog.argtypes = [vp, vp, vp, vp, i64, d, d, i32, vp, vp, ctypes.c_int]
    L.orc_release.argtypes = [vp, vp, vp, vp, vp, vp, i32, i64, vp, vp, vp, vp, vp]
    L.orc_release_jacobian.argtypes = [vp, vp, vp, i32, i64, vp, vp]
    L.orc_adjoint_leapfrog.argtypes = [vp, vp, vp, vp, i64, d, d, i32, vp, vp, vp, vp, vp, vp,
                                        ctypes.c_int]
    L.orc_perturb_leapfrog.argtypes = [vp, vp, vp, vp, vp, vp, i64, d, d, i32, ctypes.c_int,
                                       vp, vp, ctypes.c_int]
    L.orc_dopri5.argtypes = [vp, vp, vp, vp, i64, d, d, vp, i32, d, d, d, d, d, i32, vp, i32, vp,
                             vp, vp, ctypes.c_int]
    L.orc_perturb_dopri5.argtypes = [vp, vp, vp, vp, vp, i64, d, d, vp, i32, ctypes.c_int, d, d,
                                     d, d, d, i32, vp, vp, vp, vp, vp, ctypes.c_int]
    L.orc_dp5_tableau.argtypes = [vp]
    L.orc_dp5_step.argtypes = [vp, vp, d, d, vp, vp, vp, vp, i32, vp]
    L.orc_host_dgrad.argtypes = [vp, vp, i64, vp]
    return L


# ------------------------------------------------------------------------------------------
# descriptors
# ------------------------------------------------------------------------------------------
def host(components):
    """components: list of (kind, [p0, p1, p2, p3]) with p0 = G*m."""
    h = _Host()
    h.n_components = len(components)
    for i, (kind, p) in enumerate(components):
        h.c[i].kind = kind
        pp = list(p) + [0.0] * (4 - len(p))
        for k in range(4):
            h.c[i].p[k] = float(pp[k])
    return h


def milky_way_host(m_halo=5.4e11, rs_halo=15.62):
    """Config host: MN disk (6.8e10, 3, 0.28) + Hernquist bulge (5e9, 1) + NFW halo
    (reference GalaMilkyWayPotential minus its nucleus, potential.py:390-414)."""
    return host([(COMP_MN, [G * 6.8e10, 3.0, 0.28]),
                 (COMP_HERNQUIST, [G * 5e9, 1.0, 0.0]),
                 (COMP_NFW, [G * m_halo, rs_halo])])


class Sources:
    """Holds numpy arrays alive for a _Sources descriptor."""

    def __init__(self, kind=SRC_NONE, gm=None, rs=None, pos=None, vel=None, t0=None,
                 t_window=None, group_size=0, group=None):
        self.kind = kind
        self.arrays = {}
        s = _Sources()
        s.kind = kind
        if kind != SRC_NONE:
            gm = np.ascontiguousarray(gm, dtype=np.float64)
            rs = np.ascontiguousarray(rs, dtype=np.float64)
            pos = np.ascontiguousarray(pos, dtype=np.float64)
            s.n = gm.shape[0]
            s.n_track = pos.shape[0] if kind in (SRC_TRACK, SRC_TRACK_HERNQUIST) else 0
            self.arrays.update(gm=gm, rs=rs, pos=pos)
            s.gm, s.rs, s.pos = gm.ctypes.data, rs.ctypes.data, pos.ctypes.data
            if kind in (SRC_LINE, SRC_LINE_HERNQUIST):
                vel = np.ascontiguousarray(vel, dtype=np.float64)
                t0 = np.ascontiguousarray(t0, dtype=np.float64)
                t_window = np.ascontiguousarray(t_window, dtype=np.float64)
                self.arrays.update(vel=vel, t0=t0, t_window=t_window)
                s.vel, s.t0, s.t_window = vel.ctypes.data, t0.ctypes.data, t_window.ctypes.data
            if group_size:
                group = np.ascontiguousarray(group, dtype=np.int32)
                self.arrays.update(group=group)
                s.group_size, s.group = int(group_size), group.ctypes.data
        self.desc = s


NO_SOURCES = None


def _src_ptr(src):
    if src is None:
        src = Sources()
    return src, ctypes.addressof(src.desc)


def _p(a):
    return None if a is None else a.ctypes.data


# ------------------------------------------------------------------------------------------
# functions
# ------------------------------------------------------------------------------------------
def threefry2x32(k0, k1, x0, x1):
    out = np.zeros(2, dtype=np.uint32)
    lib().orc_threefry2x32(k0, k1, x0, x1, out.ctypes.data)
    return int(out[0]), int(out[1])


def randint5_1000(seed):
    out = np.zeros(5, dtype=np.int64)
    lib().orc_randint5_1000(seed, out.ctypes.data)
    return out


def prng_normal(key_seed):
    return lib().orc_prng_normal(key_seed)


def erfinv(u):
    return lib().orc_erfinv(float(u))


def release_normals(seed, i):
    z = np.zeros(4)
    lib().orc_release_normals(seed, i, z.ctypes.data)
    return z


def evaluate(h, xyz, t=0.0, step=0, src=None):
    xyz = np.ascontiguousarray(np.atleast_2d(xyz), dtype=np.float64)
    n = xyz.shape[0]
    phi, grad, hess = np.zeros(n), np.zeros((n, 3)), np.zeros((n, 9))
    src, sp = _src_ptr(src)
    lib().orc_eval(ctypes.addressof(h), sp, xyz.ctypes.data, n, t, step,
                   phi.ctypes.data, grad.ctypes.data, hess.ctypes.data)
    return phi, grad, hess.reshape(n, 3, 3)


def host_dgrad(h, xyz):
    """(n, n_comp, 2, 3): d(grad Phi_c)/d(G m_c) and d(grad Phi_c)/d(p1_c) at each point."""
    xyz = np.ascontiguousarray(np.atleast_2d(xyz), dtype=np.float64)
    n = xyz.shape[0]
    out = np.zeros((n, max(h.n_components, 1), 2, 3))
    lib().orc_host_dgrad(ctypes.addressof(h), xyz.ctypes.data, n, out.ctypes.data)
    return out[:, :h.n_components]


def leapfrog(h, w0, t0, dt, nsteps, start=None, src=None, traj=False, nthreads=0, fast=False):
    w0 = np.ascontiguousarray(np.atleast_2d(w0), dtype=np.float64)
    n = w0.shape[0]
    out = np.zeros_like(w0)
    st = None if start is None else np.ascontiguousarray(start, dtype=np.int32)
    tr = np.zeros((nsteps + 1, n, 6)) if traj else None
    src, sp = _src_ptr(src)
    lib(fast).orc_leapfrog(ctypes.addressof(h), sp, w0.ctypes.data, _p(st), n, t0, dt, nsteps,
                       out.ctypes.data, _p(tr), nthreads)
    return (out, tr) if traj else out


def dopri5(h, w0, t1, ts=(), t0=None, t_start=None, src=None, rtol=1e-7, atol=1e-7, dtmin=0.3,
           dtmax=None, max_steps=10_000, stops=(), dt0=None, nthreads=0, fast=False):
    """Adaptive Dormand-Prince 5(4) with diffrax's PID step control (restates integrate_orbit
    with solver=Dopri5, main.py:125-163).  Returns (w_final (n, 6), ys (n, len(ts), 6),
    stats (n, 3) = accepted, rejected, status)."""
    w0 = np.ascontiguousarray(np.atleast_2d(w0), dtype=np.float64)
    n = w0.shape[0]
    ts = np.ascontiguousarray(np.atleast_1d(np.asarray(ts, dtype=np.float64)))
    nts = ts.size
    tsa = None if t_start is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(t_start, dtype=np.float64), (n,)))
    t0v = float(t0) if t0 is not None else (float(ts.min()) if nts else 0.0)
    st = np.ascontiguousarray(np.asarray(stops, dtype=np.float64).reshape(-1))
    out = np.zeros((n, 6))
    ys = np.zeros((n, max(nts, 1), 6))
    stats = np.zeros((n, 3), dtype=np.int32)
    src, sp = _src_ptr(src)
    lib(fast).orc_dopri5(ctypes.addressof(h), sp, w0.ctypes.data, _p(tsa), n, t0v, float(t1),
                         ts.ctypes.data if nts else None, nts, rtol, atol,
                         -1.0 if dtmin is None else float(dtmin),
                         -1.0 if dtmax is None else float(dtmax),
                         -1.0 if dt0 is None else float(dt0), int(max_steps),
                         st.ctypes.data if st.size else None, st.size, out.ctypes.data,
                         ys.ctypes.data, stats.ctypes.data, nthreads)
    return out, ys[:, :nts], stats


def stream_dopri5(h, prog_w0, ts, msat, seed, src=None, kvals=None, release_host=None,
                  nthreads=0, fast=False, **tol):
    """Spray stream with adaptive Dopri5 (restates gen_stream_vmapped with solver=Dopri5,
    main.py:343-368): progenitor saved at every ts, release at every ts, particle i < T
    integrated from ts[i] to ts[-1].  Returns (lead, trail, stats)."""
    ts = np.asarray(ts, dtype=np.float64)
    T = len(ts) - 1
    _, prog, _ = dopri5(h, prog_w0, float(ts.max()), t0=float(ts.min()), ts=ts, src=src,
                        nthreads=nthreads, fast=fast, **tol)
    if release_host is None:  # release in the whole potential (gen_stream_vmapped)
        pc, pf, vc, vf = release(h, prog[0], ts, msat, seed, kvals, src=src)
    else:                     # release in a base potential (gen_stream_vmapped_with_pert)
        pc, pf, vc, vf = release(release_host, prog[0], ts, msat, seed, kvals)
    w0 = np.empty((2 * T, 6))
    w0[0::2, :3], w0[0::2, 3:] = pc[:T], vc[:T]
    w0[1::2, :3], w0[1::2, 3:] = pf[:T], vf[:T]
    out, _, stats = dopri5(h, w0, float(ts[-1]), t_start=np.repeat(ts[:T], 2), src=src,
                           nthreads=nthreads, fast=fast, **tol)
    return out[0::2], out[1::2], stats


def dp5_tableau():
    """(A (7, 7) lower triangle, c (7,), b (7,), b - bhat (7,), midpoint weights (7,))."""
    t = np.zeros(15 + 28)
    lib().orc_dp5_tableau(t.ctypes.data)
    A = np.zeros((7, 7))
    k = 0
    for i in range(1, 6):
        for j in range(i):
            A[i, j] = t[k]
            k += 1
    return A, t[15:22], t[22:29], t[29:36], t[36:43]


def dp5_step(h_pot, t, h, y, s=(), src=None):
    """One Dopri5 step: (y1, error estimate, dense output at fractions s)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    s = np.ascontiguousarray(np.atleast_1d(np.asarray(s, dtype=np.float64)))
    yn, err = np.zeros(6), np.zeros(6)
    dense = np.zeros((max(s.size, 1), 6))
    src, sp = _src_ptr(src)
    lib().orc_dp5_step(ctypes.addressof(h_pot), sp, float(t), float(h), y.ctypes.data,
                       yn.ctypes.data, err.ctypes.data, s.ctypes.data if s.size else None,
                       s.size, dense.ctypes.data)
    return yn, err, dense[:s.size]


def release(h, prog_w, ts, msat, seed, kvals=None, src=None, steps=None):
    """Release model at every progenitor row (restates gen_stream_ics' vmapped release_model,
    main.py:287-306) in host + sources: the tidal radius uses the Hessian of the whole potential
    (main.py:77-85).  Release i is at time ts[i]; TRACK sources need the grid row steps[i]."""
    prog_w = np.ascontiguousarray(prog_w, dtype=np.float64)
    nrel = prog_w.shape[0]
    gms = np.ascontiguousarray(G * np.broadcast_to(np.asarray(msat, dtype=np.float64), (nrel,)))
    ts = np.ascontiguousarray(np.broadcast_to(np.asarray(ts, dtype=np.float64), (nrel,)))
    st = None if steps is None else np.ascontiguousarray(steps, dtype=np.int32)
    kv = None if kvals is None else np.ascontiguousarray(kvals, dtype=np.float64)
    outs = [np.zeros((nrel, 3)) for _ in range(4)]
    src, sp = _src_ptr(src)
    rc = lib().orc_release(ctypes.addressof(h), sp, prog_w.ctypes.data, ts.ctypes.data, _p(st),
                           gms.ctypes.data, nrel, seed, _p(kv), *[o.ctypes.data for o in outs])
    if rc < 0:
        raise ValueError("release: TRACK sources need steps")
    return tuple(outs)  # pos_close, pos_far, vel_close, vel_far


def release_jacobian(h, prog_w, msat, seed, kvals=None):
    prog_w = np.ascontiguousarray(prog_w, dtype=np.float64)
    nrel = prog_w.shape[0]
    gms = np.ascontiguousarray(G * np.broadcast_to(np.asarray(msat, dtype=np.float64), (nrel,)))
    kv = None if kvals is None else np.ascontiguousarray(kvals, dtype=np.float64)
    nin = 6 + 2 * h.n_components
    jac = np.zeros((nrel, 12, nin))
    lib().orc_release_jacobian(ctypes.addressof(h), prog_w.ctypes.data, gms.ctypes.data, nrel,
                               seed, _p(kv), jac.ctypes.data)
    return jac


def stream_leapfrog(h, prog_w0, ts, msat, seed, dt, src=None, kvals=None, nthreads=0,
                    release_host=None):
    """Plain spray stream with fixed-step leapfrog (restates gen_stream_vmapped,
    main.py:343-368): progenitor orbit saved at every ts, release model at every ts, then each
    particle i < len(ts)-1 integrated from ts[i] to ts[-1].  ts must lie on the dt grid."""
    ts = np.asarray(ts, dtype=np.float64)
    T = len(ts) - 1
    kidx = np.rint((ts - ts[0]) / dt).astype(np.int64)
    nsteps = int(kidx[-1])
    _, prog = leapfrog(h, prog_w0, ts[0], dt, nsteps, src=src, traj=True)
    prog_at = prog[kidx, 0, :]
    if release_host is None:  # release in the whole potential (gen_stream_vmapped)
        pc, pf, vc, vf = release(h, prog_at, ts, msat, seed, kvals, src=src, steps=kidx)
    else:                     # release in a base potential (gen_stream_vmapped_with_pert)
        pc, pf, vc, vf = release(release_host, prog_at, ts, msat, seed, kvals)
    w0 = np.empty((2 * T, 6))
    w0[0::2, :3], w0[0::2, 3:] = pc[:T], vc[:T]
    w0[1::2, :3], w0[1::2, 3:] = pf[:T], vf[:T]
    start = np.repeat(kidx[:T], 2).astype(np.int32)
    out = leapfrog(h, w0, ts[0], dt, nsteps, start=start, src=src, nthreads=nthreads)
    return out[0::2], out[1::2]


def perturb_leapfrog(h, pert, w0, t0, dt, nsteps, dw0=None, start=None, base_src=None,
                     with_radius=False, nthreads=0):
    """Base orbits + first-order responses to each source of ``pert`` (a Sources).
    Returns (w (n, 6), dw (n, m, 6 or 12))."""
    w0 = np.ascontiguousarray(np.atleast_2d(w0), dtype=np.float64)
    n = w0.shape[0]
    m = pert.desc.n
    dd = 12 if with_radius else 6
    out = np.zeros_like(w0)
    dw = np.zeros((n, m, dd))
    d0 = None if dw0 is None else np.ascontiguousarray(dw0, dtype=np.float64).reshape(n, m, dd)
    st = None if start is None else np.ascontiguousarray(start, dtype=np.int32)
    base_src, bp = _src_ptr(base_src)
    lib().orc_perturb_leapfrog(ctypes.addressof(h), bp, ctypes.addressof(pert.desc),
                               w0.ctypes.data, _p(d0), _p(st), n, t0, dt, nsteps,
                               int(with_radius), out.ctypes.data, dw.ctypes.data, nthreads)
    return out, dw


def perturb_dopri5(h, pert, w0, t1, ts=(), t0=None, t_start=None, dw0=None, with_radius=False,
                   rtol=1e-6, atol=1e-6, dtmin=0.05, dtmax=None, max_steps=10_000, dt0=None,
                   nthreads=0):
    """Base orbits + first-order responses with adaptive Dopri5 on the whole pytree state (one
    controller, RMS norm over all 6 + m D components).  Returns (w (n, 6), dw (n, m, D),
    ys (n, len(ts), 6), dys (n, len(ts), m, D), stats (n, 3))."""
    w0 = np.ascontiguousarray(np.atleast_2d(w0), dtype=np.float64)
    n = w0.shape[0]
    m = pert.desc.n
    dd = 12 if with_radius else 6
    ts = np.ascontiguousarray(np.atleast_1d(np.asarray(ts, dtype=np.float64)))
    nts = ts.size
    tsa = None if t_start is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(t_start, dtype=np.float64), (n,)))
    t0v = float(t0) if t0 is not None else (float(ts.min()) if nts else 0.0)
    d0 = None if dw0 is None else np.ascontiguousarray(dw0, dtype=np.float64).reshape(n, m, dd)
    out = np.zeros((n, 6))
    dw = np.zeros((n, m, dd))
    ys = np.zeros((n, max(nts, 1), 6))
    dys = np.zeros((n, max(nts, 1), m, dd))
    stats = np.zeros((n, 3), dtype=np.int32)
    rc = lib().orc_perturb_dopri5(ctypes.addressof(h), ctypes.addressof(pert.desc),
                                  w0.ctypes.data, _p(d0), _p(tsa), n, t0v, float(t1),
                                  ts.ctypes.data if nts else None, nts, int(with_radius),
                                  rtol, atol, -1.0 if dtmin is None else float(dtmin),
                                  -1.0 if dtmax is None else float(dtmax),
                                  -1.0 if dt0 is None else float(dt0), int(max_steps),
                                  out.ctypes.data, dw.ctypes.data, ys.ctypes.data,
                                  dys.ctypes.data, stats.ctypes.data, nthreads)
    if rc < 0:
        raise ValueError("perturb_dopri5: TRACK sources are not supported")
    return out, dw, ys[:, :nts], dys[:, :nts], stats


def adjoint_leapfrog(h, w0, t0, dt, nsteps, lam_T, start=None, src=None, forcing=None,
                     nthreads=0):
    """Discrete adjoint of the leapfrog.  Returns dict(lam0 (n, 6), g_host (n_comp, 2):
    dL/d(G m_c), dL/d(p1_c), g_gm (m,), g_pos (rows, m, 3))."""
    w0 = np.ascontiguousarray(np.atleast_2d(w0), dtype=np.float64)
    n = w0.shape[0]
    lam_T = np.ascontiguousarray(lam_T, dtype=np.float64).reshape(n, 6)
    st = None if start is None else np.ascontiguousarray(start, dtype=np.int32)
    fo = None if forcing is None else np.ascontiguousarray(forcing, dtype=np.float64)
    src, sp = _src_ptr(src)
    m = src.desc.n if src.kind != SRC_NONE else 0
    rows = nsteps + 1 if src.kind == SRC_TRACK else 1
    lam0 = np.zeros((n, 6))
    gh = np.zeros((max(h.n_components, 1), 2))
    gg = np.zeros(max(m, 1))
    gp = np.zeros((rows, max(m, 1), 3))
    lib().orc_adjoint_leapfrog(ctypes.addressof(h), sp, w0.ctypes.data, _p(st), n, t0, dt, nsteps,
                               lam_T.ctypes.data, _p(fo), lam0.ctypes.data, gh.ctypes.data,
                               gg.ctypes.data, gp.ctypes.data, nthreads)
    return dict(lam0=lam0, g_host=gh[:h.n_components], g_gm=gg[:m], g_pos=gp[:, :m])
