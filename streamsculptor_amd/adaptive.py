"""Adaptive Dormand-Prince 5(4) integration through the gfx950 kernel ``ssc_dopri5``.

Drop-in for integrate_orbit / gen_stream_vmapped with ``solver=Dopri5()`` (reference
streamsculptor/main.py:125-163,343-368): diffrax's PIDController(rtol, atol, dtmin, dtmax,
force_dtmin=True, jump_ts) with dt0=None and SaveAt(ts=ts).  One orbit per lane, each with its
own step-size controller; saved rows between step ends come from the Dopri5 dense output, and
``dense=True`` returns the step polynomials themselves (``DenseSolution``).  Failures
(max_steps exhausted) give +inf rows, as the reference does with throw=False, or raise with
throw=True.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


def _options(rtol, atol, dtmin, dtmax, max_steps, stops_t, dt0=None):
    o = _lib.Dopri5Options()
    o.rtol, o.atol = float(rtol), float(atol)
    o.dtmin = -1.0 if dtmin is None else float(dtmin)
    o.dtmax = -1.0 if dtmax is None else float(dtmax)
    o.dt0 = -1.0 if dt0 is None else float(dt0)
    o.max_steps = int(max_steps)
    o.n_stops = 0 if stops_t is None else int(stops_t.numel())
    o.stops = None if stops_t is None else stops_t.data_ptr()
    return o


def _sources(pot):
    h, src = pot._descriptor()
    if src is not None and src.kind in _lib.TRACK_KINDS:
        raise NotImplementedError("adaptive steps need source positions at arbitrary times; "
                                  "sources sampled on a fixed grid run with solver='leapfrog'")
    return h, src


def run(h, sd, w0, t1, t0=0.0, t_start=None, ts=None, rtol=1e-7, atol=1e-7, dtmin=0.3,
        dtmax=None, max_steps=10_000, jump_ts=None, dt0=None, throw=False):
    """ssc_dopri5 on rows w0 (n, 6) with a host descriptor ``h`` and a sources descriptor ``sd``
    (or None).  Orbit i runs from t_start[i] (default t0) to t1 and is saved at ts (monotone in
    the direction of integration).  Returns (w_final (n, 6), ys (n, len(ts), 6) or None,
    stats (n, 3) int32 = accepted, rejected, status)."""
    w0 = _lib.f64(w0).reshape(-1, 6).contiguous()
    n = w0.shape[0]
    dev = w0.device
    tst = None if t_start is None else _lib.f64(t_start).reshape(-1).contiguous()
    if tst is not None and tst.numel() != n:
        raise ValueError("t_start needs one time per orbit")
    tsd = None if ts is None else _lib.f64(ts).reshape(-1).contiguous()
    nts = 0 if tsd is None else tsd.numel()
    stops = None
    if jump_ts is not None:
        stops = _lib.f64(np.sort(np.asarray(jump_ts, dtype=np.float64).reshape(-1)))
    out = torch.empty_like(w0)
    ys = torch.empty(n, nts, 6, dtype=torch.float64, device=dev) if nts else None
    stats = torch.empty(n, 3, dtype=torch.int32, device=dev)
    opt = _options(rtol, atol, dtmin, dtmax, max_steps, stops, dt0)
    _lib.call("ssc_dopri5", _lib.ctypes.byref(h), None if sd is None else _lib.ctypes.byref(sd),
              _lib.ptr(w0), _lib.ptr(tst), n, float(t0), float(t1), _lib.ptr(tsd), nts,
              _lib.ctypes.byref(opt), _lib.ptr(out), _lib.ptr(ys), _lib.ptr(stats))
    if throw:
        failed = int((stats[:, 2] != 0).sum().item())
        if failed:
            raise RuntimeError(f"Dopri5: {failed} of {n} orbits reached max_steps={max_steps} "
                               "before t1")
    return out, ys, stats


def run_orbits(h, sd, w0, t_start, t_end, ts, rtol=1e-7, atol=1e-7, dtmin=0.3, dtmax=None,
               max_steps=10_000, jump_ts=None, dt0=None, throw=False):
    """ssc_dopri5_orbits: orbit i runs from t_start[i] to t_end[i] and saves at its own row of
    ts (n, m), monotone in its direction of integration.  Returns (w_final (n, 6),
    ys (n, m, 6), stats (n, 3))."""
    w0 = _lib.f64(w0).reshape(-1, 6).contiguous()
    n = w0.shape[0]
    dev = w0.device
    tst = _lib.f64(t_start).reshape(-1).contiguous()
    ten = _lib.f64(t_end).reshape(-1).contiguous()
    tsd = _lib.f64(ts).reshape(n, -1).contiguous()
    if tst.numel() != n or ten.numel() != n:
        raise ValueError("t_start and t_end need one time per orbit")
    nts = tsd.shape[1]
    stops = None
    if jump_ts is not None:
        stops = _lib.f64(np.sort(np.asarray(jump_ts, dtype=np.float64).reshape(-1)))
    out = torch.empty_like(w0)
    ys = torch.empty(n, nts, 6, dtype=torch.float64, device=dev)
    stats = torch.empty(n, 3, dtype=torch.int32, device=dev)
    opt = _options(rtol, atol, dtmin, dtmax, max_steps, stops, dt0)
    _lib.call("ssc_dopri5_orbits", _lib.ctypes.byref(h),
              None if sd is None else _lib.ctypes.byref(sd), _lib.ptr(w0), _lib.ptr(tst),
              _lib.ptr(ten), n, _lib.ptr(tsd) if nts else None, nts, _lib.ctypes.byref(opt),
              _lib.ptr(out), _lib.ptr(ys) if nts else None, _lib.ptr(stats))
    if throw:
        failed = int((stats[:, 2] != 0).sum().item())
        if failed:
            raise RuntimeError(f"Dopri5: {failed} of {n} orbits reached max_steps={max_steps} "
                               "before their end time")
    return out, ys, stats


REC = 20  # doubles per dense-output record (tau, h, y0[6], f0[6], ymid[6])


class DenseSolution:
    """Result of ``integrate_orbit(dense=True)`` (the reference's diffrax Solution with
    SaveAt(dense=True), streamsculptor/main.py:141): ``evaluate(t)`` and ``derivative(t)`` through
    the Dopri5 step polynomials -- on each accepted step the quartic through y0, y1, h f0, h f1
    and Shampine's midpoint value, the interpolant of the kernel's saved rows.  ``ts`` and ``ys``
    are None as in the reference.  Times outside an orbit's interval use its first or last
    step's polynomial.  Orbits that ran out of steps evaluate to +inf.

    Single orbit: evaluate(t) is (6,) for a scalar t and (k, 6) for k times; a batch of n orbits
    gives (n, 6) and (n, k, 6)."""

    def __init__(self, rec, stats, tau_dir, batched, t0, t1):
        self.rec = rec                    # (n, cap, REC)
        self.stats = stats                # (n, 3) int32: accepted, rejected, status
        self.dir = tau_dir                # (n,) +-1
        self.batched = batched
        self.t0, self.t1 = t0, t1
        self.ts = None
        self.ys = None
        n, cap, _ = rec.shape
        cnt = (stats[:, 0].long() + 1).clamp(max=cap)
        starts = rec[:, :, 0].clone()
        starts[torch.arange(cap, device=rec.device)[None, :] >= cnt[:, None]] = float("inf")
        self._starts = starts
        self._cnt = cnt

    def _pieces(self, t):
        t = _lib.f64(t)
        scalar = t.ndim == 0
        t = t.reshape(-1)
        n = self.rec.shape[0]
        tau = self.dir[:, None] * t[None, :]                               # (n, k)
        k = torch.searchsorted(self._starts, tau.contiguous(), right=True) - 1
        last = (self._cnt - 2).clamp(min=0)[:, None]
        k = torch.minimum(k.clamp(min=0), last)
        k1 = torch.minimum(k + 1, (self._cnt - 1)[:, None])
        rows = torch.arange(n, device=t.device)[:, None]
        r0, r1 = self.rec[rows, k], self.rec[rows, k1]                     # (n, k, REC)
        h = r0[..., 1:2]
        s = torch.where(h != 0, (tau[..., None] - r0[..., 0:1]) / torch.where(h != 0, h, 1.0), 0.0)
        y0, f0, ym = r0[..., 2:8], r0[..., 8:14], r0[..., 14:20]
        y1, f1 = r1[..., 2:8], r1[..., 8:14]
        F0, F1 = h * f0, h * f1
        D1 = y1 - y0 - F0
        D2 = F1 - F0
        D3 = ym - y0 - 0.5 * F0
        a4 = 2.0 * D2 - 8.0 * D1 + 16.0 * D3
        b3 = D2 - 2.0 * D1 - 2.0 * a4
        c2 = D1 - a4 - b3
        return scalar, s, h, y0, F0, c2, b3, a4, f0

    def _finish(self, v, scalar):
        v = torch.where((self.stats[:, 2] != 0)[:, None, None], float("inf"), v)
        if scalar:
            v = v[:, 0]
        return v if self.batched else v[0]

    def evaluate(self, t0, t1=None, left=True):
        """y(t0), or y(t1) - y(t0) when t1 is given (diffrax's evaluate(t0, t1))."""
        if t1 is not None:
            return self.evaluate(t1) - self.evaluate(t0)
        scalar, s, h, y0, F0, c2, b3, a4, _ = self._pieces(t0)
        y = y0 + s * (F0 + s * (c2 + s * (b3 + s * a4)))
        return self._finish(y, scalar)

    def derivative(self, t, left=True):
        """dy/dt of the interpolant."""
        scalar, s, h, y0, F0, c2, b3, a4, f0 = self._pieces(t)
        dtau = (F0 + s * (2.0 * c2 + s * (3.0 * b3 + s * 4.0 * a4))) / torch.where(h != 0, h, 1.0)
        dtau = torch.where(h != 0, dtau, f0)
        return self._finish(self.dir[:, None, None] * dtau, scalar)


def run_dense(h, sd, w0, t1, t0=0.0, t_start=None, rtol=1e-7, atol=1e-7, dtmin=0.3, dtmax=None,
              max_steps=10_000, jump_ts=None, dt0=None, throw=False, batched=True):
    """ssc_dopri5_dense: like ``run`` but returns a DenseSolution over [t_start[i], t1]."""
    w0 = _lib.f64(w0).reshape(-1, 6).contiguous()
    n = w0.shape[0]
    dev = w0.device
    tst = None if t_start is None else _lib.f64(t_start).reshape(-1).contiguous()
    if tst is not None and tst.numel() != n:
        raise ValueError("t_start needs one time per orbit")
    stops = None
    if jump_ts is not None:
        stops = _lib.f64(np.sort(np.asarray(jump_ts, dtype=np.float64).reshape(-1)))
    cap = int(max_steps) + 1
    rec = torch.empty(n, cap, REC, dtype=torch.float64, device=dev)
    out = torch.empty_like(w0)
    stats = torch.empty(n, 3, dtype=torch.int32, device=dev)
    opt = _options(rtol, atol, dtmin, dtmax, max_steps, stops, dt0)
    _lib.call("ssc_dopri5_dense", _lib.ctypes.byref(h),
              None if sd is None else _lib.ctypes.byref(sd), _lib.ptr(w0), _lib.ptr(tst), n,
              float(t0), float(t1), _lib.ctypes.byref(opt), _lib.ptr(out), _lib.ptr(rec), cap,
              _lib.ptr(stats))
    if throw:
        failed = int((stats[:, 2] != 0).sum().item())
        if failed:
            raise RuntimeError(f"Dopri5: {failed} of {n} orbits reached max_steps={max_steps} "
                               "before t1")
    ta = tst if tst is not None else torch.full((n,), float(t0), dtype=torch.float64, device=dev)
    tau_dir = torch.where(float(t1) >= ta, 1.0, -1.0).to(torch.float64)
    return DenseSolution(rec, stats, tau_dir, batched, t0, t1)


def dopri5(pot, w0, t1, t0=0.0, t_start=None, ts=None, rtol=1e-7, atol=1e-7, dtmin=0.3,
           dtmax=None, max_steps=10_000, jump_ts=None, dt0=None, throw=False):
    """``run`` with the descriptors of potential ``pot``."""
    h, src = _sources(pot)
    sd = src.desc() if src is not None else None
    return run(h, sd, w0, t1, t0=t0, t_start=t_start, ts=ts, rtol=rtol, atol=atol, dtmin=dtmin,
               dtmax=dtmax, max_steps=max_steps, jump_ts=jump_ts, dt0=dt0, throw=throw)


def integrate_orbit(pot, w, tsn, t0v, t1v, rtol=1e-7, atol=1e-7, dtmin=0.3, dtmax=None,
                    max_steps=10_000, jump_ts=None, throw=False, dt0=None, batched=False):
    """Solution for Potential.integrate_orbit with solver=Dopri5 (reference main.py:125-163)."""
    from .main import Solution
    tsn = np.asarray(tsn, dtype=np.float64).reshape(-1)
    direction = 1.0 if t1v >= t0v else -1.0
    order = np.argsort(direction * tsn, kind="stable")
    sorted_ts = tsn[order]
    _, ys, _ = dopri5(pot, w.reshape(-1, 6), t1v, t0=t0v, ts=sorted_ts, rtol=rtol, atol=atol,
                      dtmin=dtmin, dtmax=dtmax, max_steps=max_steps, jump_ts=jump_ts, dt0=dt0,
                      throw=throw)
    if not np.array_equal(order, np.arange(order.size)):
        inv = torch.as_tensor(np.argsort(order), device=ys.device)
        ys = ys[:, inv]
    ts_t = _lib.f64(tsn)
    return Solution(ts_t, ys if batched else ys[0])


def integrate_orbits_own_ts(pot, w, ts2, t0s, t1s, rtol=1e-7, atol=1e-7, dtmin=0.3,
                            dtmax=None, max_steps=10_000, jump_ts=None, throw=False, dt0=None):
    """Solution for integrate_orbit_batch_vmapped with a per-orbit ts (N, M) and solver=Dopri5
    (reference main.py:186-202): orbit i from t0s[i] to t1s[i], saved at ts2[i]."""
    from .main import Solution
    h, src = _sources(pot)
    sd = src.desc() if src is not None else None
    direction = np.where(t1s >= t0s, 1.0, -1.0)[:, None]
    order = np.argsort(direction * ts2, axis=1, kind="stable")
    srt = np.take_along_axis(ts2, order, axis=1)
    _, ys, _ = run_orbits(h, sd, w, t0s, t1s, srt, rtol=rtol, atol=atol, dtmin=dtmin, dtmax=dtmax,
                          max_steps=max_steps, jump_ts=jump_ts, dt0=dt0, throw=throw)
    inv = torch.as_tensor(np.argsort(order, axis=1), device=ys.device)
    ys = torch.take_along_dim(ys, inv[:, :, None], dim=1)
    return Solution(_lib.f64(ts2), ys)


def integrate_orbit_dense(pot, w, t0v, t1v, rtol=1e-7, atol=1e-7, dtmin=0.3, dtmax=None,
                          max_steps=10_000, jump_ts=None, throw=False, dt0=None, batched=False,
                          t_start=None):
    """DenseSolution for Potential.integrate_orbit(dense=True) with solver=Dopri5."""
    h, src = _sources(pot)
    sd = src.desc() if src is not None else None
    return run_dense(h, sd, w.reshape(-1, 6), t1v, t0=t0v, t_start=t_start, rtol=rtol, atol=atol,
                     dtmin=dtmin, dtmax=dtmax, max_steps=max_steps, jump_ts=jump_ts, dt0=dt0,
                     throw=throw, batched=batched)


def integrate_to(pot, w0, t_start, t1, rtol=1e-7, atol=1e-7, dtmin=0.3, dtmax=None,
                 max_steps=10_000, throw=False, jump_ts=None):
    """Final rows of orbits started at t_start[i] and integrated to t1 (the per-particle solve
    of gen_stream_vmapped, reference main.py:349-360)."""
    out, _, _ = dopri5(pot, w0, t1, t_start=t_start, rtol=rtol, atol=atol, dtmin=dtmin,
                       dtmax=dtmax, max_steps=max_steps, jump_ts=jump_ts, throw=throw)
    return out
