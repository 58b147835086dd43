"""Potential zoo on the stream-particle path (reference streamsculptor/potential.py).

Analytic classes lower to kernel components; point-source classes lower to a Plummer
``SourceSet``.  Parameters follow the reference constructors; masses are in Msun and enter the
kernels as G*m.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .main import Potential, SampledTrack, SourceSet, usys, _as_float


def _arr(x):
    return np.asarray(_as_float(x), dtype=np.float64)


class MiyamotoNagaiDisk(Potential):
    """Reference streamsculptor/potential.py:66-72."""

    def __init__(self, m, a, b, units=None):
        super().__init__(units, {"m": m, "a": a, "b": b})

    def _kernel_components(self):
        return [(_lib.COMP_MIYAMOTO_NAGAI, [self._G * float(self.m), float(self.a), float(self.b)])]


class NFWPotential(Potential):
    """Reference streamsculptor/potential.py:74-84 (no softening: singular at r = 0)."""

    def __init__(self, m, r_s, units=None):
        super().__init__(units, {"m": m, "r_s": r_s})

    def _kernel_components(self):
        return [(_lib.COMP_NFW, [self._G * float(self.m), float(self.r_s)])]


class PlummerPotential(Potential):
    """Reference streamsculptor/potential.py:124-130."""

    def __init__(self, m, r_s, units=None):
        super().__init__(units, {"m": m, "r_s": r_s})

    def _kernel_components(self):
        return [(_lib.COMP_PLUMMER, [self._G * float(self.m), float(self.r_s)])]


class HernquistPotential(Potential):
    """Reference streamsculptor/potential.py:132-138; soft is added to r^2."""

    def __init__(self, m, r_s, soft=0.0, units=None):
        super().__init__(units, {"m": m, "r_s": r_s, "soft": soft})

    def _kernel_components(self):
        return [(_lib.COMP_HERNQUIST, [self._G * float(self.m), float(self.r_s), float(self.soft)])]


class Potential_Combine(Potential):
    """Sum of potentials (reference streamsculptor/potential.py:1279-1296)."""

    def __init__(self, potential_list, units=None):
        super().__init__(units, {"potential_list": potential_list})

    def _kernel_components(self):
        comps = []
        for p in self.potential_list:
            comps += p._kernel_components()
        return comps

    def _kernel_sources(self):
        from .main import _merge_sources
        return _merge_sources([p._kernel_sources() for p in self.potential_list])


class GalaMilkyWayPotential(Potential):
    """Disk + bulge + nucleus + halo (reference streamsculptor/potential.py:390-414)."""

    def __init__(self, units=None):
        super().__init__(units, {"params": None})
        self.m_disk, self.a_disk, self.b_disk = 6.80e10, 3.0, 0.28
        self.m_bulge, self.c_bulge = 5e9, 1.0
        self.m_nucleus, self.c_nucleus = 1.71e9, 0.07
        self.m_halo, self.r_s_halo = 5.4e11, 15.62
        self.pot = Potential_Combine([
            MiyamotoNagaiDisk(self.m_disk, self.a_disk, self.b_disk, units=self.units),
            HernquistPotential(self.m_bulge, self.c_bulge, units=self.units),
            HernquistPotential(self.m_nucleus, self.c_nucleus, units=self.units),
            NFWPotential(self.m_halo, self.r_s_halo, units=self.units)], units=self.units)

    def _kernel_components(self):
        return self.pot._kernel_components()


class CustomPotential(Potential):
    """Arbitrary Python potential (reference streamsculptor/potential.py:787-799).  It cannot
    be lowered to the gfx950 kernels: every kernel-backed method raises NotImplementedError."""

    def __init__(self, potential_func=None, units=None):
        super().__init__(units, {"potential_func": potential_func})


class SubhaloLinePotential(Potential):
    """Plummer subhalos on straight lines x0 + v (t - t0), active while |t - t0| < t_window
    (reference streamsculptor/potential.py:802-830)."""

    def __init__(self, m, a, subhalo_x0, subhalo_v, subhalo_t0, t_window, units=None):
        super().__init__(units, {"m": m, "a": a, "subhalo_x0": subhalo_x0, "subhalo_v": subhalo_v,
                                 "subhalo_t0": subhalo_t0, "t_window": t_window})

    def _kernel_components(self):
        return []

    def _kernel_sources(self):
        m = np.atleast_1d(_arr(self.m))
        n = m.shape[0]
        return SourceSet(_lib.SRC_LINE, gm=self._G * m,
                         rs=np.broadcast_to(np.atleast_1d(_arr(self.a)), (n,)).copy(),
                         pos=_arr(self.subhalo_x0).reshape(n, 3),
                         vel=_arr(self.subhalo_v).reshape(n, 3),
                         t0=np.broadcast_to(np.atleast_1d(_arr(self.subhalo_t0)), (n,)).copy(),
                         t_window=np.broadcast_to(np.atleast_1d(_arr(self.t_window)), (n,)).copy())

    def potential_per_SH(self, xyz, t):
        """Per-subhalo potential at xyz (reference streamsculptor/potential.py:833-850)."""
        return _per_source(self, xyz, t, "phi")


def _profile_kind(func):
    """Source kind of a subhalo profile class (the profiles with a kernel form)."""
    if func is PlummerPotential:
        return _lib.SRC_LINE
    if func is HernquistPotential:
        return _lib.SRC_LINE_HERNQUIST
    raise NotImplementedError(
        f"subhalo profile {getattr(func, '__name__', func)!r} has no kernel form; "
        "use PlummerPotential or HernquistPotential")


class SubhaloLinePotentialCustom_fromFunc(Potential):
    """Straight-line subhalos with the profile of a potential class (reference
    streamsculptor/potential.py:1161-1213): ``func`` is ``PlummerPotential`` or
    ``HernquistPotential`` (the profiles with a kernel form; others raise), instantiated per
    subhalo as func(m=m_j, r_s=r_s_j) at x - (x0_j + v_j (t - t0_j)) while |t - t0_j| < t_window.
    Hernquist subhalos run in the evaluation, leapfrog, Dopri5 and first-order response
    kernels; the adjoint kernels differentiate Plummer subhalos only."""

    def __init__(self, func, m, r_s, subhalo_x0, subhalo_v, subhalo_t0, t_window, units=None):
        super().__init__(units, {"func": func, "m": m, "r_s": r_s, "subhalo_x0": subhalo_x0,
                                 "subhalo_v": subhalo_v, "subhalo_t0": subhalo_t0,
                                 "t_window": t_window})
        self._kind = _profile_kind(func)

    def _kernel_components(self):
        return []

    def _kernel_sources(self):
        m = np.atleast_1d(_arr(self.m))
        n = m.shape[0]
        return SourceSet(self._kind, gm=self._G * m,
                         rs=np.broadcast_to(np.atleast_1d(_arr(self.r_s)), (n,)).copy(),
                         pos=_arr(self.subhalo_x0).reshape(n, 3),
                         vel=_arr(self.subhalo_v).reshape(n, 3),
                         t0=np.broadcast_to(np.atleast_1d(_arr(self.subhalo_t0)), (n,)).copy(),
                         t_window=np.broadcast_to(np.atleast_1d(_arr(self.t_window)), (n,)).copy())

    def potential_per_SH(self, xyz, t):
        """Per-subhalo potential at xyz (reference streamsculptor/potential.py:1193-1213)."""
        return _per_source(self, xyz, t, "phi")


class SubhaloLinePotential_Custom(Potential):
    """Straight-line subhalos sharing one profile instance (reference
    streamsculptor/potential.py:908-956): ``pot`` is a ``PlummerPotential`` or
    ``HernquistPotential`` (soft = 0) with a single (m, r_s) -- the reference's lax.cond needs
    the profile's value to have the shape (1,) of its inactive branch -- and subhalo j is that
    profile at x - (x0_j + v_j (t - t0_j)) while |t - t0_j| < t_window_j.  Other profiles raise
    NotImplementedError (no kernel form)."""

    def __init__(self, pot, subhalo_x0, subhalo_v, subhalo_t0, t_window, units=None):
        super().__init__(units, {"pot": pot, "subhalo_x0": subhalo_x0, "subhalo_v": subhalo_v,
                                 "subhalo_t0": subhalo_t0, "t_window": t_window})
        self._kind = _profile_kind(type(pot))
        if self._kind == _lib.SRC_LINE_HERNQUIST and np.any(_arr(getattr(pot, "soft", 0.0)) != 0):
            raise NotImplementedError("softened Hernquist subhalos have no kernel form")

    def _kernel_components(self):
        return []

    def _kernel_sources(self):
        m = np.atleast_1d(_arr(self.pot.m)).reshape(-1)
        r_s = np.atleast_1d(_arr(self.pot.r_s)).reshape(-1)
        if m.size != 1 or r_s.size != 1:
            raise ValueError("SubhaloLinePotential_Custom takes a profile with one mass and one "
                             "scale radius, shared by every subhalo")
        n = _arr(self.subhalo_x0).reshape(-1, 3).shape[0]
        return SourceSet(self._kind, gm=np.full(n, self.pot._G * m[0]), rs=np.full(n, r_s[0]),
                         pos=_arr(self.subhalo_x0).reshape(n, 3),
                         vel=_arr(self.subhalo_v).reshape(n, 3),
                         t0=np.broadcast_to(np.atleast_1d(_arr(self.subhalo_t0)), (n,)).copy(),
                         t_window=np.broadcast_to(np.atleast_1d(_arr(self.t_window)), (n,)).copy())

    def potential_per_SH(self, xyz, t):
        """Per-subhalo potential at xyz (reference streamsculptor/potential.py:939-956)."""
        return _per_source(self, xyz, t, "phi")


class SubhaloLinePotential_dRadius(Potential):
    """d Phi / d a of the straight-line Plummer subhalos of a SubhaloLinePotential (reference
    streamsculptor/potential.py:852-904): the structural potential whose gradient drives the
    radius block of the first-order responses.  The response kernel forms
    -grad(d Phi_j / d a_j) = 3 G m_j a_j (x - s_j) / (|x - s_j|^2 + a_j^2)^(5/2) in closed form,
    so this class carries the same subhalo description and is consumed by the perturbation
    field; evaluating it on its own has no kernel form."""

    def __init__(self, m, a, subhalo_x0, subhalo_v, subhalo_t0, t_window, units=None):
        super().__init__(units, {"m": m, "a": a, "subhalo_x0": subhalo_x0, "subhalo_v": subhalo_v,
                                 "subhalo_t0": subhalo_t0, "t_window": t_window})

    def matches(self, pert):
        """True when this describes the same subhalos as ``pert``."""
        names = ("m", "a", "subhalo_x0", "subhalo_v", "subhalo_t0", "t_window")
        try:
            return all(np.allclose(np.asarray(_arr(getattr(self, k)), dtype=np.float64),
                                   np.asarray(_arr(getattr(pert, k)), dtype=np.float64),
                                   rtol=0, atol=0) for k in names)
        except (AttributeError, ValueError):
            return False


class SubhaloLinePotentialCustom_dRadius_fromFunc(Potential):
    """d Phi / d r_s of straight-line subhalos with the profile of a potential class (reference
    streamsculptor/potential.py:1215-1268, which takes jax.grad of func(m, r_s).potential in
    r_s): the structural potential of the radius block of the first-order responses.  The
    response kernel forms its gradient in closed form -- Plummer
    -3 G m r_s (x - s) / (|x - s|^2 + r_s^2)^(5/2), Hernquist -2 G m (x - s) / (r (r + r_s)^3)
    with r = |x - s| -- so this class carries the subhalo description for the perturbation
    field; evaluating it on its own has no kernel form."""

    def __init__(self, func, m, r_s, subhalo_x0, subhalo_v, subhalo_t0, t_window, units=None):
        super().__init__(units, {"func": func, "m": m, "r_s": r_s, "subhalo_x0": subhalo_x0,
                                 "subhalo_v": subhalo_v, "subhalo_t0": subhalo_t0,
                                 "t_window": t_window})
        self._kind = _profile_kind(func)

    def potential_per_SH(self, xyz, t):
        self._kernel_components()

    def matches(self, pert):
        """True when this describes the same subhalos (and profile) as ``pert``."""
        if getattr(pert, "_kind", None) != self._kind:
            return False
        names = ("m", "r_s", "subhalo_x0", "subhalo_v", "subhalo_t0", "t_window")
        try:
            return all(np.array_equal(np.asarray(_arr(getattr(self, k)), dtype=np.float64),
                                      np.asarray(_arr(getattr(pert, k)), dtype=np.float64))
                       for k in names)
        except (AttributeError, ValueError):
            return False


class PlummerClumps(Potential):
    """Static Plummer clumps summed directly (the clumpy-halo substructure of the surrogate
    workload).  Equivalent to a SubhaloLinePotential with zero velocity and an infinite
    window."""

    def __init__(self, m, r_s, positions, units=None):
        super().__init__(units, {"m": m, "r_s": r_s, "positions": positions})

    def _kernel_components(self):
        return []

    def _kernel_sources(self):
        m = np.atleast_1d(_arr(self.m))
        n = m.shape[0]
        return SourceSet(_lib.SRC_STATIC, gm=self._G * m,
                         rs=np.broadcast_to(np.atleast_1d(_arr(self.r_s)), (n,)).copy(),
                         pos=_arr(self.positions).reshape(n, 3))


class LiveSubhaloPotential(Potential):
    """Plummer subhalos moving on their own orbits, given as a track sampled on the integration
    grid: ``track[k, j]`` is the position of subhalo j at t0 + k*dt.  The nearest reference
    construct is a potential translating along an interpolated centre
    (TimeDepTranslatingPotential, streamsculptor/potential.py:448-462)."""

    def __init__(self, m, r_s, track, t0, dt, units=None):
        super().__init__(units, {"m": m, "r_s": r_s, "track": track, "t0": t0, "dt": dt})

    @classmethod
    def from_orbits(cls, host, w0, m, r_s, t0, dt, nsteps, units=None):
        """Integrate subhalo initial conditions w0 (M, 6) in ``host`` with the leapfrog kernel and
        keep every grid row as the track."""
        w0 = _lib.f64(w0).reshape(-1, 6)
        _, ys = host._leapfrog(w0, t0, dt, nsteps, save_idx=np.arange(nsteps + 1, dtype=np.int32))
        obj = cls(m, r_s, ys[:, :, :3].contiguous(), t0, dt, units=units or host.units)
        obj.orbits = ys
        obj.w0 = w0
        return obj

    def _kernel_components(self):
        return []

    def _kernel_sources(self):
        m = self.m
        gm = (self._G * m.to(torch.float64)) if isinstance(m, torch.Tensor) else self._G * np.atleast_1d(_arr(m))
        n = gm.shape[0]
        rs = self.r_s if isinstance(self.r_s, torch.Tensor) else np.broadcast_to(np.atleast_1d(_arr(self.r_s)), (n,)).copy()
        return SourceSet(_lib.SRC_TRACK, gm=gm, rs=rs, pos=self.track,
                         grid=(float(self.t0), float(self.dt)))

    def potential_per_SH(self, xyz, t):
        return _per_source(self, xyz, t, "phi")


def _track_profile(cls):
    """Source kind of a profile centred on a moving track."""
    if cls is PlummerPotential:
        return _lib.SRC_TRACK
    if cls is HernquistPotential:
        return _lib.SRC_TRACK_HERNQUIST
    raise NotImplementedError(
        f"a {getattr(cls, '__name__', cls)!r} centred on a moving track has no kernel form; "
        "use PlummerPotential or HernquistPotential")


def _positions_of(track, times):
    """Positions (len(times), 3) of a centre given as an interpolant with ``evaluate(t)`` (a
    Dopri5 DenseSolution, or any object returning rows (..., >= 3) for an array of times) or as
    a function of time (vectorised over times, or called once per time)."""
    def _np(v):
        return v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else np.asarray(v)

    times = np.asarray(times, dtype=np.float64).reshape(-1)
    if hasattr(track, "evaluate"):
        v = _np(track.evaluate(_lib.f64(times)))
        return np.asarray(v, dtype=np.float64).reshape(times.size, -1)[:, :3]
    try:
        v = np.asarray(_np(track(times)), dtype=np.float64)
        if v.shape == (times.size, 3):
            return v
    except Exception:  # noqa: BLE001 -- a scalar-only function: fall back to one call per time
        pass
    return np.stack([np.asarray(_np(track(float(t))), dtype=np.float64).reshape(-1)[:3]
                     for t in times])


class ProgenitorPotential(Potential):
    """A Plummer or Hernquist sphere (``prog_pot`` = PlummerPotential or HernquistPotential,
    with m and r_s) centred on a moving track: Phi(x - c(t)), c(t) = interp_func.evaluate(t)[:3]
    (reference streamsculptor/potential.py:140-153, where interp_func is a dense diffrax
    solution of the progenitor orbit; here a DenseSolution of ``integrate_orbit(dense=True)``,
    or any object with evaluate(t)).  The centre is sampled on the integration grid (fixed-step
    leapfrog, evaluation, release); adaptive solvers raise."""

    def __init__(self, m, r_s, interp_func, prog_pot, units=None):
        super().__init__(units, {"m": m, "r_s": r_s, "interp_func": interp_func,
                                 "prog_pot": prog_pot})
        self._kind = _track_profile(prog_pot)

    def _kernel_components(self):
        return []

    def _kernel_sources(self):
        gm = np.atleast_1d(self._G * _arr(self.m)).reshape(-1)[:1]
        rs = np.atleast_1d(_arr(self.r_s)).reshape(-1)[:1]
        interp = self.interp_func
        return SampledTrack(self._kind, gm, rs,
                            lambda times: _positions_of(interp, times)[:, None, :])


class TimeDepTranslatingPotential(Potential):
    """A potential translating with a track: Phi_pot(x - c(t)) with c = center_spl(t)
    (reference streamsculptor/potential.py:448-462).  ``pot`` is a PlummerPotential or
    HernquistPotential, or a Potential_Combine of spheres of one of these profiles; other
    profiles have no kernel form and raise.  The centre is sampled on the integration grid
    (fixed-step leapfrog, evaluation, release); adaptive solvers raise."""

    def __init__(self, pot, center_spl, units=None):
        super().__init__(units, {"pot": pot, "center_spl": center_spl})
        parts = pot.potential_list if isinstance(pot, Potential_Combine) else [pot]
        kinds = {_track_profile(type(p)) for p in parts}
        if len(kinds) != 1:
            raise NotImplementedError("the translating spheres must share one profile")
        if any(np.any(_arr(getattr(p, "soft", 0.0)) != 0) for p in parts):
            raise NotImplementedError("softened Hernquist spheres have no kernel form")
        self._kind = kinds.pop()
        self._parts = parts

    def _kernel_components(self):
        return []

    def _kernel_sources(self):
        gm = np.array([p._G * float(_arr(p.m)) for p in self._parts])
        rs = np.array([float(_arr(p.r_s)) for p in self._parts])
        n, fn = len(self._parts), self.center_spl
        return SampledTrack(self._kind, gm, rs,
                            lambda times: np.repeat(_positions_of(fn, times)[:, None, :], n, 1))


def _per_source(pot, xyz, t, what):
    """Per-source potential (M,) or gradient (M, 3) at one point, one kernel call per source
    set row (used by the perturbation helpers)."""
    src = pot._kernel_sources()
    tens = src.tensors()
    n = tens["gm"].numel()
    outs = []
    for j in range(n):
        sub = SourceSet(src.kind, gm=tens["gm"][j:j + 1],
                        rs=tens["rs"][j:j + 1],
                        pos=(tens["pos"][:, j:j + 1, :].contiguous()
                             if src.kind in _lib.TRACK_KINDS else tens["pos"][j:j + 1]),
                        vel=None if tens["vel"] is None else tens["vel"][j:j + 1],
                        t0=None if tens["t0"] is None else tens["t0"][j:j + 1],
                        t_window=None if tens["t_window"] is None else tens["t_window"][j:j + 1],
                        grid=src.grid)
        single = _SourceOnly(sub, pot.units)
        outs.append(single.potential(xyz, t) if what == "phi" else single.gradient(xyz, t))
    return torch.stack(outs, dim=-1 if what == "phi" else -2)


class _SourceOnly(Potential):
    def __init__(self, src, units):
        super().__init__(units, {})
        self._src = src

    def _kernel_components(self):
        return []

    def _kernel_sources(self):
        return self._src


__all__ = ["MiyamotoNagaiDisk", "NFWPotential", "PlummerPotential", "HernquistPotential",
           "Potential_Combine", "GalaMilkyWayPotential", "CustomPotential",
           "SubhaloLinePotential", "SubhaloLinePotential_Custom",
           "SubhaloLinePotentialCustom_fromFunc",
           "SubhaloLinePotential_dRadius", "SubhaloLinePotentialCustom_dRadius_fromFunc",
           "PlummerClumps", "ProgenitorPotential", "TimeDepTranslatingPotential",
           "LiveSubhaloPotential", "usys"]
