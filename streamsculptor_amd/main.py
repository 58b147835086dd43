"""Drop-in ``Potential`` base class: evaluation, orbit integration and spray streams.

Mirrors the public surface of reference ``streamsculptor/main.py`` (class ``Potential``,
lines 21-430) on hand-written gfx950 kernels reached through the C-ABI in
``include/ssc_mi355x.h``.  Arrays may be numpy arrays, torch tensors or lists; results are
float64 torch tensors on the current ROCm device (the analogue of the reference's device
arrays).  Every call enqueues work on torch's current stream.

Differences from the reference that a user sees:

* ``solver=`` takes ``solvers.Leapfrog()`` / ``"leapfrog"`` (fixed step, needs ``dt0=``) or
  ``solvers.Dopri5()``; the reference's ``integrate_orbit`` default Dopri8 is not built, so the
  default here is Dopri5.
* A potential is lowered to a kernel descriptor (analytic components + at most one set of
  Plummer sources).  A ``CustomPotential`` holding arbitrary Python cannot be lowered and
  raises ``NotImplementedError``.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib
from . import solvers as _solvers


class UnitSystem:
    """Minimal stand-in for gala's UnitSystem: only the (kpc, Myr, Msun, radian) system and the
    dimensionless system are supported (reference streamsculptor/main.py:18)."""

    def __init__(self, *units, G=None):
        self.units = tuple(str(u) for u in units)
        self.G = G

    def __eq__(self, other):
        return isinstance(other, UnitSystem) and self.units == other.units

    def __hash__(self):
        return hash(self.units)

    def __repr__(self):
        return f"UnitSystem{self.units}"


usys = UnitSystem("kpc", "Myr", "solMass", "radian", G=_lib.G_USYS)
dimensionless = UnitSystem("dimensionless", G=1.0)


def _as_float(v):
    if isinstance(v, torch.Tensor):
        return v.detach().cpu().numpy() if v.ndim else float(v)
    return v


class Solution:
    """What ``integrate_orbit`` returns: ``.ts`` (saved times) and ``.ys`` (states), like the
    fields of ``diffrax.Solution`` the reference reads (``.ys``, e.g. main.py:289,349)."""

    def __init__(self, ts, ys, stats=None):
        self.ts = ts
        self.ys = ys
        self.stats = stats or {}

    def evaluate(self, t):
        idx = torch.nonzero(torch.isclose(self.ts, torch.as_tensor(float(t), dtype=self.ts.dtype,
                                                                 device=self.ts.device),
                                          rtol=0, atol=1e-9)).flatten()
        if idx.numel() == 0:
            raise NotImplementedError("dense interpolation is not available; evaluate at a saved time")
        return self.ys[..., int(idx[0]), :]


class SourceSet:
    """Plummer point sources in kernel form (kind STATIC, TRACK or LINE)."""

    def __init__(self, kind, gm, rs, pos, vel=None, t0=None, t_window=None, grid=None):
        self.kind = kind
        self.host = dict(gm=gm, rs=rs, pos=pos, vel=vel, t0=t0, t_window=t_window)
        self.grid = grid  # TRACK: (t0, dt) of row 0 and the row spacing
        self._dev = None

    def tensors(self):
        if self._dev is None:
            self._dev = {k: (None if v is None else _lib.f64(v)) for k, v in self.host.items()}
        return self._dev

    def desc(self):
        t = self.tensors()
        s = _lib.Sources()
        s.kind = self.kind
        s.n = t["gm"].numel()
        if s.n > _lib.MAX_SOURCES:
            raise ValueError(f"at most {_lib.MAX_SOURCES} sources per potential")
        s.n_track = t["pos"].shape[0] if self.kind in _lib.TRACK_KINDS else 0
        for k in ("gm", "rs", "pos", "vel", "t0", "t_window"):
            setattr(s, k, None if t[k] is None else t[k].data_ptr())
        return s

    def desc_grouped(self, group_size, group):
        """Descriptor of these sources as consecutive groups of ``group_size``; orbit i sees
        group ``group[i]`` (a device int32 tensor the caller keeps alive for the call)."""
        if self.kind not in (_lib.SRC_STATIC, _lib.SRC_LINE, _lib.SRC_LINE_HERNQUIST):
            raise NotImplementedError("grouped sources must be STATIC or LINE")
        t = self.tensors()
        s = _lib.Sources()
        s.kind = self.kind
        s.n = t["gm"].numel()
        if group_size < 1 or group_size > _lib.MAX_SOURCES or s.n % group_size:
            raise ValueError("group_size must divide the number of sources and be <= "
                             f"{_lib.MAX_SOURCES}")
        s.group_size = int(group_size)
        if group.dtype != torch.int32 or not group.is_contiguous():
            raise ValueError("group must be a contiguous int32 tensor")
        s.group = group.data_ptr()
        for k in ("gm", "rs", "pos", "vel", "t0", "t_window"):
            setattr(s, k, None if t[k] is None else t[k].data_ptr())
        return s

    def on_grid(self, t0, dt, nsteps):
        """This TRACK set re-indexed so that row k is the position at t0 + k*dt, k = 0..nsteps
        (the integration step must be +- the track spacing and t0 a track row)."""
        g0, gdt = self.grid
        if abs(g0 - t0) <= 1e-9 * max(1.0, abs(t0)) and abs(gdt - dt) <= 1e-12 * max(1.0, abs(dt)):
            return self
        ratio = dt / gdt
        r0f = (t0 - g0) / gdt
        r0 = int(round(r0f))
        sgn = int(round(ratio))
        rows = self.tensors()["pos"].shape[0]
        if (abs(abs(ratio) - 1.0) > 1e-12 or abs(r0f - r0) > 1e-9 * max(1.0, abs(r0f))
                or not (0 <= r0 < rows) or not (0 <= r0 + sgn * nsteps < rows)):
            raise ValueError("a source track must cover the integration grid with the same "
                             f"spacing (track t0={g0}, dt={gdt}; integration t0={t0}, dt={dt}, "
                             f"nsteps={nsteps})")
        t = self.tensors()
        idx = torch.arange(nsteps + 1, device=t["pos"].device) * sgn + r0
        return SourceSet(self.kind, gm=t["gm"], rs=t["rs"], pos=t["pos"][idx].contiguous(),
                         grid=(float(t0), float(dt)))

    def rows_for(self, times):
        """(source set, table row of each time) for evaluations at several times: TRACK tables
        give their rows, other kinds (None)."""
        if self.kind not in _lib.TRACK_KINDS:
            return self, None
        return self, np.array([self.step_of(t) for t in np.asarray(times).reshape(-1)],
                              dtype=np.int32)

    def at_time(self, t):
        """(source set, table row) for an evaluation at time t."""
        return self, self.step_of(t)

    def step_of(self, t):
        """Grid row of time t for TRACK sources."""
        if self.kind not in _lib.TRACK_KINDS:
            return 0
        g0, gdt = self.grid
        kf = (float(t) - g0) / gdt
        k = int(round(kf))
        if abs(kf - k) > 1e-9 * max(1.0, abs(kf)) or k < 0 or k >= self.tensors()["pos"].shape[0]:
            raise ValueError(f"time {t} is not a row of the source track (t0={g0}, dt={gdt})")
        return k


class SampledTrack(SourceSet):
    """TRACK sources (Plummer or Hernquist profiles) whose positions come from a function of
    time, ``centre(times) -> (len(times), n, 3)``: the table is sampled on the integration grid
    when the grid is known (``on_grid``), or at the times of an evaluation."""

    def __init__(self, kind, gm, rs, centre):
        super().__init__(kind, gm=gm, rs=rs, pos=None)
        self.centre = centre

    def _table(self, times):
        t = self.tensors()
        pos = _lib.f64(self.centre(np.asarray(times, dtype=np.float64).reshape(-1)))
        return SourceSet(self.kind, gm=t["gm"], rs=t["rs"], pos=pos.reshape(len(times), -1, 3))

    def tensors(self):
        if self._dev is None:
            self._dev = {k: (None if v is None else _lib.f64(v)) for k, v in self.host.items()}
        return self._dev

    def desc(self):
        raise RuntimeError("a sampled track has no table until it is sampled (on_grid/at_time)")

    def on_grid(self, t0, dt, nsteps):
        src = self._table(t0 + dt * np.arange(nsteps + 1))
        src.grid = (float(t0), float(dt))
        return src

    def at_time(self, t):
        return self._table([float(_as_float(t))]), 0

    def rows_for(self, times):
        times = np.asarray(times, dtype=np.float64).reshape(-1)
        return self._table(times), np.arange(times.size, dtype=np.int32)

    def step_of(self, t):
        raise RuntimeError("a sampled track is evaluated through at_time / rows_for")


def _merge_sources(sets):
    sets = [s for s in sets if s is not None]
    if not sets:
        return None
    if len(sets) > 1:
        raise NotImplementedError("a potential may hold at most one set of Plummer sources")
    return sets[0]


class Potential:
    """Base class (reference streamsculptor/main.py:21-35): stores parameters and G.

    Subclasses describe themselves to the kernels through ``_kernel_components()`` (a list of
    ``(kind, [G*m, p1, p2, p3])``) and ``_kernel_sources()`` (a ``SourceSet`` or None)."""

    def __init__(self, units, params):
        if units is None:
            units = dimensionless
        self.units = units
        self._G = units.G if isinstance(units, UnitSystem) and units.G is not None else _lib.G_USYS
        for name, param in params.items():
            if hasattr(param, "unit") and hasattr(param, "value"):
                raise NotImplementedError("astropy quantities are not supported; pass values in "
                                          "kpc, Myr, Msun")
            setattr(self, name, param)
        self._desc_cache = None

    # ------------------------------------------------------------------ kernel description
    def _kernel_components(self):
        raise NotImplementedError(
            f"{type(self).__name__} has no kernel form: arbitrary Python potentials cannot be "
            "lowered to the gfx950 kernels")

    def _kernel_sources(self):
        return None

    def _descriptor(self):
        if self._desc_cache is None:
            comps = self._kernel_components()
            if len(comps) > _lib.MAX_COMPONENTS:
                raise ValueError(f"at most {_lib.MAX_COMPONENTS} analytic components")
            h = _lib.HostPotential()
            h.n_components = len(comps)
            for i, (kind, p) in enumerate(comps):
                h.c[i].kind = int(kind)
                pp = [float(v) for v in p] + [0.0] * (4 - len(p))
                for k in range(4):
                    h.c[i].p[k] = pp[k]
            self._desc_cache = (h, self._kernel_sources())
        return self._desc_cache

    def _host_only(self):
        """Descriptor of the analytic components alone (no point sources)."""
        h, _ = self._descriptor()
        return h

    # ------------------------------------------------------------------ evaluation
    def _eval(self, xyz, t, want_phi=False, want_grad=False, want_hess=False):
        h, src = self._descriptor()
        x = _lib.f64(xyz)
        shape = x.shape[:-1]
        if x.shape[-1] != 3:
            raise ValueError("xyz must have a trailing dimension of 3")
        x = x.reshape(-1, 3).contiguous()
        n = x.shape[0]
        dev = x.device
        phi = torch.empty(n, dtype=torch.float64, device=dev) if want_phi else None
        grad = torch.empty(n, 3, dtype=torch.float64, device=dev) if want_grad else None
        hess = torch.empty(n, 3, 3, dtype=torch.float64, device=dev) if want_hess else None
        step = 0
        if src is not None:
            src, step = src.at_time(t)
        sd = src.desc() if src is not None else None
        _lib.call("ssc_eval", _lib.ctypes.byref(h), None if sd is None else _lib.ctypes.byref(sd),
                  _lib.ptr(x), n, float(_as_float(t)), step, _lib.ptr(phi), _lib.ptr(grad),
                  _lib.ptr(hess))
        out = []
        if want_phi:
            out.append(phi.reshape(shape))
        if want_grad:
            out.append(grad.reshape(*shape, 3))
        if want_hess:
            out.append(hess.reshape(*shape, 3, 3))
        return out[0] if len(out) == 1 else tuple(out)

    def potential(self, xyz, t):
        return self._eval(xyz, t, want_phi=True)

    def gradient(self, xyz, t):
        """grad Phi (reference main.py:37-40 takes it by jax.grad; here closed forms)."""
        return self._eval(xyz, t, want_grad=True)

    def acceleration(self, xyz, t):
        return -self.gradient(xyz, t)

    def density(self, xyz, t):
        """Laplacian / (4 pi G) (reference main.py:42-45)."""
        H = self._eval(xyz, t, want_hess=True)
        return torch.diagonal(H, dim1=-2, dim2=-1).sum(-1) / (4 * math.pi * self._G)

    def jacobian_force(self, xyz, t):
        """Jacobian of the gradient = Hessian of Phi (reference main.py:59-65)."""
        return self._eval(xyz, t, want_hess=True)

    def local_circular_velocity(self, xyz, t):
        x = _lib.f64(xyz)
        r = torch.linalg.norm(x, dim=-1)
        g = self.gradient(x, t)
        return torch.sqrt(r * (g * x).sum(-1) / r)

    def dphidr(self, x, t):
        x = _lib.f64(x)
        r = torch.linalg.norm(x, dim=-1, keepdim=True)
        return (self.gradient(x, t) * x / r).sum(-1)

    def d2phidr2(self, x, t):
        """rhat^T H rhat with rhat held fixed (reference main.py:76-85)."""
        x = _lib.f64(x)
        rh = x / torch.linalg.norm(x, dim=-1, keepdim=True)
        H = self.jacobian_force(x, t)
        return torch.einsum("...i,...ij,...j->...", rh, H, rh)

    def omega(self, x, v):
        x, v = _lib.f64(x), _lib.f64(v)
        r2 = (x * x).sum(-1, keepdim=True)
        return torch.linalg.norm(torch.linalg.cross(x, v, dim=-1) / r2, dim=-1)

    def tidalr(self, x, v, Msat, t):
        return (self._G * _lib.f64(Msat) / (self.omega(x, v) ** 2 - self.d2phidr2(x, t))) ** (1.0 / 3.0)

    def lagrange_pts(self, x, v, Msat, t):
        x = _lib.f64(x)
        rt = self.tidalr(x, v, Msat, t)
        rh = x / torch.linalg.norm(x, dim=-1, keepdim=True)
        return x - rh * rt[..., None], x + rh * rt[..., None]

    def velocity_acceleration(self, t, xv, args=None):
        """The orbit RHS [v, -grad Phi] (reference main.py:116-120)."""
        xv = _lib.f64(xv)
        return torch.cat([xv[..., 3:], -self.gradient(xv[..., :3], t)], dim=-1)

    # ------------------------------------------------------------------ integration
    def _leapfrog(self, w0, t0, dt, nsteps, start=None, save_idx=None, src_override=False):
        """Run ssc_leapfrog on rows w0 (n, 6); returns (w_final, ys or None)."""
        h, src = self._descriptor()
        w0 = _lib.f64(w0).reshape(-1, 6).contiguous()
        n = w0.shape[0]
        dev = w0.device
        out = torch.empty_like(w0)
        st = None if start is None else _lib.i32(start)
        si = None if save_idx is None else _lib.i32(save_idx)
        nsave = 0 if si is None else si.numel()
        ys = torch.empty(nsave, n, 6, dtype=torch.float64, device=dev) if nsave else None
        if src is not None and src.kind in _lib.TRACK_KINDS:
            src = src.on_grid(t0, dt, nsteps)
        sd = src.desc() if src is not None else None
        _lib.call("ssc_leapfrog", _lib.ctypes.byref(h),
                  None if sd is None else _lib.ctypes.byref(sd), _lib.ptr(w0), _lib.ptr(st), n,
                  float(t0), float(dt), int(nsteps), _lib.ptr(out), _lib.ptr(si), nsave,
                  _lib.ptr(ys))
        return out, ys

    def integrate_orbit(self, w0=None, ts=None, dense=False, solver=None, rtol=1e-7, atol=1e-7,
                        dtmin=0.3, dtmax=None, max_steps=10_000, t0=None, t1=None, steps=False,
                        jump_ts=None, throw=False, dt0=None):
        """Integrate w0 (6,) -- or a batch (N, 6) -- from t0 (default min(ts)) to t1 (default
        max(ts)), saving at ts (reference main.py:125-163).  Returns a ``Solution`` with ys of
        shape (len(ts), 6) (or (N, len(ts), 6)); with dense=True (Dopri5) a DenseSolution whose
        ``evaluate(t)`` interpolates the orbit anywhere in [t0, t1]."""
        method = _solvers.resolve(solver)
        ts_t = _lib.f64(ts).reshape(-1)
        tsn = ts_t.cpu().numpy()
        t0v = float(tsn.min()) if t0 is None else float(_as_float(t0))
        t1v = float(tsn.max()) if t1 is None else float(_as_float(t1))
        w = _lib.f64(w0)
        batched = w.ndim == 2
        if dense:
            if method == "leapfrog":
                raise NotImplementedError("dense=True needs an adaptive solver (Dopri5): the "
                                          "fixed-step leapfrog has no dense interpolant")
            from . import adaptive
            return adaptive.integrate_orbit_dense(self, w, t0v, t1v, rtol=rtol, atol=atol,
                                                  dtmin=dtmin, dtmax=dtmax, max_steps=max_steps,
                                                  jump_ts=jump_ts, throw=throw, dt0=dt0,
                                                  batched=batched)
        if method == "leapfrog":
            from ._grid import leapfrog_grid
            dt, nsteps, kidx = leapfrog_grid(tsn, t0v, t1v, dt0, steps=steps)
            _, ys = self._leapfrog(w.reshape(-1, 6), t0v, dt, nsteps, save_idx=kidx)
            if steps:
                ts_t = _lib.f64(t0v + dt * np.arange(nsteps + 1))
            ys = ys.permute(1, 0, 2)  # (n, nsave, 6)
            return Solution(ts_t, ys if batched else ys[0])
        from . import adaptive
        return adaptive.integrate_orbit(self, w, tsn, t0v, t1v, rtol=rtol, atol=atol, dtmin=dtmin,
                                        dtmax=dtmax, max_steps=max_steps, jump_ts=jump_ts,
                                        throw=throw, dt0=dt0, batched=batched)

    def integrate_orbit_batch_vmapped(self, w0=None, ts=None, dense=False, solver=None, rtol=1e-7,
                                      atol=1e-7, dtmin=0.3, dtmax=None, max_steps=10_000, t0=None,
                                      t1=None, steps=False, jump_ts=None, dt0=None):
        """Batch of orbits (reference main.py:186-202).  ts is shared (1-D) or per orbit
        (N, M): orbit i then runs from min(ts[i]) to max(ts[i]) -- or from t0 to t1 when given --
        and is saved at ts[i]; ys has shape (N, M, 6)."""
        ts_a = np.asarray(_as_float(ts), dtype=np.float64)
        if ts_a.ndim == 1:
            return self.integrate_orbit(w0=_lib.f64(w0).reshape(-1, 6), ts=ts, dense=dense,
                                        solver=solver, rtol=rtol, atol=atol, dtmin=dtmin,
                                        dtmax=dtmax, max_steps=max_steps, t0=t0, t1=t1,
                                        steps=steps, jump_ts=jump_ts, dt0=dt0)
        w = _lib.f64(w0).reshape(-1, 6)
        if ts_a.ndim != 2 or ts_a.shape[0] != w.shape[0]:
            raise ValueError("ts must be 1-D or have one row per orbit, shape (N, M)")
        if dense or steps:
            raise NotImplementedError("dense=True and steps=True take a shared 1-D ts")
        n = w.shape[0]
        t0s = ts_a.min(1) if t0 is None else np.full(n, float(_as_float(t0)))
        t1s = ts_a.max(1) if t1 is None else np.full(n, float(_as_float(t1)))
        if _solvers.resolve(solver) == "leapfrog":
            return self._leapfrog_own_ts(w, ts_a, t0s, t1s, dt0)
        from . import adaptive
        return adaptive.integrate_orbits_own_ts(self, w, ts_a, t0s, t1s, rtol=rtol, atol=atol,
                                                dtmin=dtmin, dtmax=dtmax, max_steps=max_steps,
                                                jump_ts=jump_ts, dt0=dt0)

    def _leapfrog_own_ts(self, w, ts2, t0s, t1s, dt0):
        """Fixed-step orbits with their own intervals and save times on one grid: every orbit
        starts at its own grid row and the kernel records the union of all save rows, from
        which each orbit's rows are gathered (rows after an orbit's end time are not used)."""
        from ._grid import _round_checked
        if dt0 is None:
            raise ValueError("the leapfrog solver needs a step size: pass dt0=")
        nz = np.sign(t1s - t0s)
        nz = nz[nz != 0]
        if nz.size and np.any(nz != nz[0]):
            raise NotImplementedError("per-orbit intervals must all run in one direction")
        d = float(nz[0]) if nz.size else 1.0
        dt = abs(float(dt0)) * d
        tg0 = float((d * t0s).min() * d)
        k_start = _round_checked((t0s - tg0) / dt, "orbit start times")
        k_end = _round_checked((t1s - tg0) / dt, "orbit end times")
        k_ts = _round_checked((ts2 - tg0) / dt, "save times ts")
        if np.any(k_ts < k_start[:, None]) or np.any(k_ts > k_end[:, None]):
            raise ValueError("save times must lie within each orbit's [t0, t1]")
        nsteps = int(k_end.max())
        rows = np.unique(k_ts)
        order = np.argsort(k_start, kind="stable")  # the kernel walks blocks from their earliest start
        _, ys = self._leapfrog(w[torch.as_tensor(order, device=w.device)], tg0, dt, nsteps,
                               start=k_start[order].astype(np.int32),
                               save_idx=rows.astype(np.int32))
        inv = np.empty_like(order)
        inv[order] = np.arange(order.size)
        pos = np.searchsorted(rows, k_ts)                      # (N, M) rows of ys
        dev = ys.device
        out = ys[torch.as_tensor(pos, device=dev),
                 torch.as_tensor(inv, device=dev)[:, None]]    # (N, M, 6)
        return Solution(_lib.f64(ts2), out)

    integrate_orbit_batch_scan = integrate_orbit_batch_vmapped

    # ------------------------------------------------------------------ streams
    def release_model(self, x=None, v=None, Msat=None, i=None, t=None, seed_num=None,
                      kval_arr=1.0):
        """Spray release of one (or a batch of) progenitor points (reference main.py:209-280).
        Returns (pos_lead, pos_trail, v_lead, v_trail)."""
        x, v = _lib.f64(x), _lib.f64(v)
        shape = x.shape[:-1]
        w = torch.cat([x.reshape(-1, 3), v.reshape(-1, 3)], dim=1).contiguous()
        n = w.shape[0]
        idx = _lib.i32(np.broadcast_to(np.asarray(_as_float(i), dtype=np.int64), (n,)).copy())
        gms = self._G * _lib.f64(np.broadcast_to(np.asarray(_as_float(Msat), dtype=np.float64), (n,)).copy())
        tt = None if t is None else np.broadcast_to(np.asarray(_as_float(t), dtype=np.float64), (n,))
        outs = self._release(w, idx, gms, seed_num, kval_arr, t_rel=tt)
        return tuple(o.reshape(*shape, 3) for o in outs)

    def _release(self, prog_w, rel_idx, g_msat, seed_num, kval_arr, t_rel=None):
        """Release model of rows prog_w (n, 6) at times t_rel (n,) in this whole potential: its
        point sources enter the tidal radius through d2phidr2 (reference main.py:77-85,
        230-237).  t_rel is needed when the potential has point sources (LINE: their positions;
        TRACK: the table row of each release)."""
        h, src = self._descriptor()
        n = prog_w.shape[0]
        kv = np.asarray(_as_float(kval_arr), dtype=np.float64)
        kvp = None
        if kv.ndim > 0:
            kv = np.ascontiguousarray(np.ones(8) * kv)
            kvp = kv.ctypes.data_as(_lib.ctypes.c_void_p)
        dev = prog_w.device
        sd, t_dev, step_dev = None, None, None
        if src is not None:
            if t_rel is None:
                raise ValueError("the release in a potential with point sources needs the "
                                 "release times")
            tr = np.asarray(_as_float(t_rel), dtype=np.float64).reshape(-1)
            src, steps = src.rows_for(tr)
            if steps is not None:
                step_dev = _lib.i32(steps)
            t_dev = _lib.f64(tr)
            sd = src.desc()
        outs = [torch.empty(n, 3, dtype=torch.float64, device=dev) for _ in range(4)]
        _lib.call("ssc_release", _lib.ctypes.byref(h),
                  None if sd is None else _lib.ctypes.byref(sd), _lib.ptr(prog_w.contiguous()),
                  _lib.ptr(rel_idx), _lib.ptr(t_dev), _lib.ptr(step_dev),
                  _lib.ptr(g_msat.contiguous()), n, int(seed_num), kvp,
                  *[_lib.ptr(o) for o in outs])
        return outs  # pos_close (lead), pos_far (trail), vel_close, vel_far

    def _release_seeds(self, prog_w, rel_idx, g_msat, seed0, nseed, kval_arr, t_rel=None):
        """``_release`` for seeds seed0 .. seed0 + nseed - 1 in one launch (ssc_release_seeds):
        returns rows (nseed, 2, n, 6), arm 0 = (pos_close, vel_close) (lead), arm 1 =
        (pos_far, vel_far) (trail), equal to the per-seed calls bit for bit."""
        h, src = self._descriptor()
        n = prog_w.shape[0]
        kv = np.asarray(_as_float(kval_arr), dtype=np.float64)
        kvp = None
        if kv.ndim > 0:
            kv = np.ascontiguousarray(np.ones(8) * kv)
            kvp = kv.ctypes.data_as(_lib.ctypes.c_void_p)
        dev = prog_w.device
        sd, t_dev, step_dev = None, None, None
        if src is not None:
            if t_rel is None:
                raise ValueError("the release in a potential with point sources needs the "
                                 "release times")
            tr = np.asarray(_as_float(t_rel), dtype=np.float64).reshape(-1)
            src, steps = src.rows_for(tr)
            if steps is not None:
                step_dev = _lib.i32(steps)
            t_dev = _lib.f64(tr)
            sd = src.desc()
        rows = torch.empty(int(nseed), 2, n, 6, dtype=torch.float64, device=dev)
        _lib.call("ssc_release_seeds", _lib.ctypes.byref(h),
                  None if sd is None else _lib.ctypes.byref(sd), _lib.ptr(prog_w.contiguous()),
                  _lib.ptr(rel_idx), _lib.ptr(t_dev), _lib.ptr(step_dev),
                  _lib.ptr(g_msat.contiguous()), n, int(seed0), int(nseed), kvp, _lib.ptr(rows))
        return rows

    def gen_stream_ics(self, ts=None, prog_w0=None, Msat=None, seed_num=None, solver=None,
                       kval_arr=1.0, rtol=1e-7, atol=1e-7, dtmin=0.3, dtmax=None,
                       max_steps=10_000, dt0=None):
        """Progenitor orbit saved at every ts, then the release model at every ts (reference
        main.py:287-306).  Returns (pos_close, pos_far, vel_close, vel_far), each (len(ts), 3)."""
        ts_t = _lib.f64(ts).reshape(-1)
        ws = self.integrate_orbit(w0=prog_w0, ts=ts_t, solver=solver, rtol=rtol, atol=atol,
                                  dtmin=dtmin, dtmax=dtmax, max_steps=max_steps, dt0=dt0).ys
        return self._release_along(ws, ts_t.cpu().numpy(), Msat, seed_num, kval_arr)

    def _release_along(self, ws, ts, Msat, seed_num, kval_arr):
        """Release at every row of the progenitor orbit ws (T+1, 6) saved at times ts (T+1,)."""
        ts = np.asarray(_as_float(ts), dtype=np.float64).reshape(-1)
        nts = ts.size
        gms = self._G * _lib.f64(np.broadcast_to(np.asarray(_as_float(Msat), dtype=np.float64), (nts,)).copy())
        idx = torch.arange(nts, dtype=torch.int32, device=ws.device)
        return self._release(ws.contiguous(), idx, gms, seed_num, kval_arr, t_rel=ts)

    def gen_stream_vmapped(self, ts=None, prog_w0=None, Msat=None, seed_num=None, solver=None,
                           kval_arr=1.0, rtol=1e-7, atol=1e-7, dtmin=0.3, dtmax=None,
                           max_steps=10_000, throw=False, dt0=None):
        """Spray stream: particle i < len(ts)-1 of each arm is released at ts[i] and integrated
        to ts[-1] (reference main.py:343-368).  Returns (lead (T, 6), trail (T, 6))."""
        from .streams import gen_stream
        return gen_stream(self, self, ts, prog_w0, Msat, seed_num, solver=solver,
                          kval_arr=kval_arr, rtol=rtol, atol=atol, dtmin=dtmin, dtmax=dtmax,
                          max_steps=max_steps, throw=throw, dt0=dt0)

    gen_stream_scan = gen_stream_vmapped

    def gen_stream_vmapped_dense(self, ts=None, prog_w0=None, Msat=None, seed_num=None,
                                 solver=None, kval_arr=1.0, rtol=1e-7, atol=1e-7, dtmin=0.3,
                                 dtmax=None, max_steps=10_000):
        """Dense spray stream (reference main.py:409-430): every particle released at ts[i],
        i < len(ts)-1, integrated to ts[-1] with dense output.  Returns a batched
        DenseSolution over the 2 (len(ts)-1) particles ordered (release, arm) with arm 0 =
        lead; evaluate it with ``streamhelpers.eval_dense_stream``."""
        from .streams import gen_stream_dense
        return gen_stream_dense(self, self, ts, prog_w0, Msat, seed_num, solver=solver,
                                kval_arr=kval_arr, rtol=rtol, atol=atol, dtmin=dtmin, dtmax=dtmax,
                                max_steps=max_steps)

    gen_stream_scan_dense = gen_stream_vmapped_dense
