"""Auxiliary-ODE ("fields") interface (reference streamsculptor/fields.py).

A field is a class with ``term(t, coords, args)``.  ``integrate_field`` lowers the field types
whose right-hand side is known to the gfx950 kernels:

* ``hamiltonian_field(pot)``            -> the orbit integrator (``[v, -grad Phi]``);
* ``MassRadiusPerturbation_OTF(pertgen)`` -> the response kernels: base orbit in
  ``pertgen.potential_base`` plus first-order responses to each perturbing subhalo of
  ``pertgen.potential_perturbation`` (mass block) and, when the generator carries a
  structural potential, the radius block (coords[1] of shape (Nsh, 12)); ``solver=Dopri5()``
  runs ``ssc_perturb_dopri5`` (adaptive, one step controller over the whole state),
  ``solver="leapfrog"`` with ``dt0=`` runs ``ssc_perturb_leapfrog`` (fixed step);
* ``CustomField(term)`` holds arbitrary Python and cannot be lowered: NotImplementedError.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import solvers as _solvers
from .main import Solution, _as_float


class hamiltonian_field:
    """Standard Hamiltonian field (reference fields.py:101-113)."""

    def __init__(self, pot):
        self.pot = pot

    def term(self, t, xv, args=None):
        return self.pot.velocity_acceleration(t, xv, args)


class MassRadiusPerturbation_OTF:
    """First-order perturbations in subhalo mass (and radius) computed on the fly (reference
    fields.py:159-206).  ``perturbation_generator`` exposes ``potential_base``,
    ``potential_perturbation`` and optionally ``potential_structural`` (any non-None value
    enables the radius block)."""

    def __init__(self, perturbation_generator):
        self.pertgen = perturbation_generator

    def term(self, t, coords, args=None):
        raise NotImplementedError("MassRadiusPerturbation_OTF is integrated by the fused response "
                                  "kernel; evaluate the pieces through the potentials instead")


class CustomField:
    """User field with arbitrary Python ``term`` (reference fields.py:362-378): not lowerable."""

    def __init__(self, term=None):
        self._term = term

    def term(self, t, coords, args=None):
        return self._term(t, coords, args)


def _times(ts, t0, t1, backwards_int):
    """t0/t1 resolution of reference fields.py:58-82."""
    tsn = np.asarray(_as_float(ts), dtype=np.float64).reshape(-1)
    t0v, t1v = float(_as_float(t0)), float(_as_float(t1))
    if t0v != t1v:
        return tsn, t0v, t1v
    if backwards_int:
        return tsn, float(tsn.max()), float(tsn.min())
    return tsn, float(tsn.min()), float(tsn.max())


def integrate_field(w0=None, ts=None, dense=False, solver=None, field=None, args=None, rtol=1e-7,
                    atol=1e-7, dtmin=0.05, dtmax=None, max_steps=1_000, jump_ts=None,
                    backwards_int=False, t0=0.0, t1=0.0, step_ts=None, dt0=None):
    """Integrate ``w0`` along ``field`` and save at ``ts`` (reference fields.py:34-99).

    ``step_ts`` (times at which the controller must end a step, reference fields.py:54) are step
    ends of the adaptive orbit solve, like ``jump_ts``; the adaptive perturbation solve takes
    neither, and the fixed-step solver ends a step at every grid time, so there step_ts must
    lie on the grid.  ``args`` is what the reference hands to ``field.term``; the fields with a
    kernel form take no arguments."""
    if dense:
        raise NotImplementedError("dense=True is not available")
    tsn, t0v, t1v = _times(ts, t0, t1, backwards_int)
    method = _solvers.resolve(solver)
    if step_ts is not None and method == "leapfrog":
        from ._grid import leapfrog_grid
        leapfrog_grid(np.sort(np.asarray(_as_float(step_ts), dtype=np.float64).reshape(-1))
                      [::1 if t1v >= t0v else -1], t0v, t1v, dt0)  # raises when off the grid
        step_ts = None
    if isinstance(field, hamiltonian_field):
        stops = jump_ts
        if step_ts is not None:
            stops = np.concatenate([np.asarray(_as_float(x), dtype=np.float64).reshape(-1)
                                    for x in (jump_ts, step_ts) if x is not None])
        return field.pot.integrate_orbit(w0=w0, ts=tsn, solver=solver, rtol=rtol, atol=atol,
                                         dtmin=dtmin, dtmax=dtmax, max_steps=max_steps, t0=t0v,
                                         t1=t1v, jump_ts=stops, dt0=dt0)
    if step_ts is not None:
        raise NotImplementedError("step_ts is not supported by the adaptive perturbation solve")
    if isinstance(field, MassRadiusPerturbation_OTF):
        return _integrate_perturbation(field.pertgen, w0, tsn, t0v, t1v, solver, dt0,
                                       dict(rtol=rtol, atol=atol, dtmin=dtmin, dtmax=dtmax,
                                            max_steps=max_steps, jump_ts=jump_ts))
    if isinstance(field, CustomField):
        raise NotImplementedError("CustomField holds arbitrary Python and cannot be lowered to "
                                  "the gfx950 kernels")
    raise NotImplementedError(f"field {type(field).__name__} has no kernel form")


def _pert_parts(pertgen):
    base = pertgen.potential_base
    pert = pertgen.potential_perturbation
    if base._kernel_sources() is not None:
        raise NotImplementedError("the base potential of a perturbation must be analytic")
    if pert._kernel_components():
        raise NotImplementedError("the perturbing potential must consist of subhalo sources")
    src = pert._kernel_sources()
    struct = getattr(pertgen, "potential_structural", None)
    radius = struct is not None
    if radius and hasattr(struct, "matches") and not struct.matches(pert):
        raise ValueError("potential_structural must describe the same subhalos as "
                         "potential_perturbation (the radius block is d/da of those subhalos)")
    return base, src, radius


def perturb_leapfrog(base, src, w0, dw0, start, t0, dt, nsteps, with_radius, split=True,
                     segment_rows=None):
    """Run ssc_perturb_leapfrog: base rows w0 (n, 6), responses dw0 (n, m, 6|12) or None.
    split: advance the base orbits once per particle through segments of rows kept in a
    workspace (otherwise every lane of a particle recomputes them); segment_rows: rows per
    segment (default: the library's choice, at most 2 GiB of records)."""
    h = base._host_only()
    w0 = _lib.f64(w0).reshape(-1, 6).contiguous()
    n = w0.shape[0]
    m = src.tensors()["gm"].numel()
    dd = 12 if with_radius else 6
    d0 = None if dw0 is None else _lib.f64(dw0).reshape(n, m, dd).contiguous()
    st = None if start is None else _lib.i32(start)
    out = torch.empty_like(w0)
    dout = torch.empty(n, m, dd, dtype=torch.float64, device=w0.device)
    if src.kind == _lib.SRC_TRACK:
        src = src.on_grid(t0, dt, nsteps)
    sd = src.desc()
    ws = None
    if split and n > 0:
        nbytes = (_lib.load().ssc_perturb_workspace_bytes(n, int(nsteps)) if segment_rows is None
                  else int(segment_rows) * n * 9 * 8)
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=w0.device)
    _lib.call("ssc_perturb_leapfrog", _lib.ctypes.byref(h), _lib.ctypes.byref(sd), _lib.ptr(w0),
              _lib.ptr(d0), _lib.ptr(st), n, float(t0), float(dt), int(nsteps), int(with_radius),
              _lib.ptr(out), _lib.ptr(dout), _lib.ptr(ws), 0 if ws is None else ws.numel())
    return out, dout


def perturb_dopri5(base, src, w0, dw0, t1, t0=0.0, t_start=None, ts=None, with_radius=False,
                   rtol=1e-6, atol=1e-6, dtmin=0.05, dtmax=None, max_steps=10_000, dt0=None,
                   throw=True):
    """Run ssc_perturb_dopri5: base rows w0 (n, 6) and responses dw0 (n, m, 6|12) or None at
    t_start (default t0), integrated to t1 with one adaptive step controller per particle over
    its whole state; saved at ts (monotone in the direction of integration).  Returns
    (w (n, 6), dw (n, m, D), ys (n, len(ts), 6) or None, dys (n, len(ts), m, D) or None,
    stats (n, 3) int32 = accepted, rejected, status).  throw=True raises when a particle runs
    out of steps (the reference's diffeqsolve default); otherwise its outputs are +inf."""
    from .adaptive import _options
    if src.kind == _lib.SRC_TRACK:
        raise NotImplementedError("adaptive responses need subhalo positions at arbitrary "
                                  "times; sources sampled on a fixed grid run with "
                                  "solver='leapfrog'")
    h = base._host_only()
    w0 = _lib.f64(w0).reshape(-1, 6).contiguous()
    n = w0.shape[0]
    dev = w0.device
    m = src.tensors()["gm"].numel()
    dd = 12 if with_radius else 6
    d0 = None if dw0 is None else _lib.f64(dw0).reshape(n, m, dd).contiguous()
    tst = None if t_start is None else _lib.f64(t_start).reshape(-1).contiguous()
    if tst is not None and tst.numel() != n:
        raise ValueError("t_start needs one time per particle")
    tsd = None if ts is None else _lib.f64(ts).reshape(-1).contiguous()
    nts = 0 if tsd is None else tsd.numel()
    out = torch.empty_like(w0)
    dout = torch.empty(n, m, dd, dtype=torch.float64, device=dev)
    ys = torch.empty(n, nts, 6, dtype=torch.float64, device=dev) if nts else None
    dys = torch.empty(n, nts, m, dd, dtype=torch.float64, device=dev) if nts else None
    stats = torch.empty(n, 3, dtype=torch.int32, device=dev)
    opt = _options(rtol, atol, dtmin, dtmax, max_steps, None, dt0)
    sd = src.desc()
    _lib.call("ssc_perturb_dopri5", _lib.ctypes.byref(h), _lib.ctypes.byref(sd), _lib.ptr(w0),
              _lib.ptr(d0), _lib.ptr(tst), n, float(t0), float(t1), _lib.ptr(tsd), nts,
              int(with_radius), _lib.ctypes.byref(opt), _lib.ptr(out), _lib.ptr(dout),
              _lib.ptr(ys), _lib.ptr(dys), _lib.ptr(stats))
    if throw:
        failed = int((stats[:, 2] != 0).sum().item())
        if failed:
            raise RuntimeError(f"Dopri5 responses: {failed} of {n} particles reached "
                               f"max_steps={max_steps} before t1")
    return out, dout, ys, dys, stats


def _integrate_perturbation(pertgen, w0, tsn, t0v, t1v, solver, dt0, tol):
    method = _solvers.resolve(solver, default="dopri5")
    base, src, radius = _pert_parts(pertgen)
    w, d = w0
    w = _lib.f64(w).reshape(1, 6)
    m = src.tensors()["gm"].numel()
    d = _lib.f64(d).reshape(1, m, -1)
    dd = d.shape[-1]
    with_radius = radius and dd == 12
    if not with_radius:
        d = d[..., :6].contiguous()
    if method == "dopri5":
        if tol.get("jump_ts") is not None:
            raise NotImplementedError("jump_ts is not supported for perturbation fields")
        tol = {k: v for k, v in tol.items() if k != "jump_ts"}
        _, _, ys_w, ys_d, _ = perturb_dopri5(base, src, w, d, t1v, t0=t0v, ts=tsn,
                                             with_radius=with_radius, dt0=dt0, **tol)
        ys_d = ys_d[0]
        if dd == 12 and not with_radius:  # radius block held at zero, as the fixed-step path
            ys_d = torch.cat([ys_d, torch.zeros_like(ys_d)], dim=-1)
        return Solution(_lib.f64(tsn), [ys_w[0], ys_d])
    from ._grid import leapfrog_grid
    dt, nsteps, kidx = leapfrog_grid(tsn, t0v, t1v, dt0)
    if src.kind == _lib.SRC_TRACK:
        src = src.on_grid(t0v, dt, nsteps)
    ys_w, ys_d = [], []
    cur_w, cur_d, kprev = w, d, 0
    for k in kidx:
        if k > kprev:
            cur_w, cur_d = perturb_leapfrog(base, src, cur_w, cur_d, None, t0v + kprev * dt, dt,
                                            int(k - kprev), with_radius)
            kprev = int(k)
        ys_w.append(cur_w[0])
        full = torch.zeros(m, 12, dtype=torch.float64, device=cur_d.device)
        full[:, :cur_d.shape[-1]] = cur_d[0]
        ys_d.append(full if dd == 12 else full[:, :dd])
    return Solution(_lib.f64(tsn), [torch.stack(ys_w), torch.stack(ys_d)])
