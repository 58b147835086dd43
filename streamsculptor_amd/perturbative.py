"""First-order perturbation theory for streams (reference streamsculptor/perturbative.py).

``BaseStreamModel`` holds the zeroth-order quantities (release conditions, forward progenitor
orbit and the Jacobian of the release model with respect to the progenitor's phase-space
point); ``GenerateMassRadiusPerturbation`` integrates the progenitor's responses to each
subhalo backwards from the observation time, maps them through the release Jacobian to the
particles' response initial conditions, and ``compute_perturbation_OTF`` integrates every
particle and its (mass, radius) responses from its release time to the observation time with
the fused gfx950 response kernels: adaptive Dormand-Prince 5(4) (``solver=Dopri5()``, one step
controller over each particle's whole state, as the reference's Runge-Kutta path) or the
fixed-step leapfrog (``solver='leapfrog'``, ``dt0=``).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import solvers as _solvers
from .fields import (MassRadiusPerturbation_OTF, integrate_field, perturb_dopri5,
                     perturb_leapfrog)
from .main import usys


def _method(solver, dt0):
    """'dopri5' or 'leapfrog' (which needs dt0) for the response kernels."""
    method = _solvers.resolve(solver, default="dopri5")
    if method == "leapfrog" and dt0 is None:
        raise ValueError("the fixed-step response kernel needs a step size: pass dt0=")
    return method


class BaseStreamModel:
    """Zeroth-order stream model (reference perturbative.py:243-296).

    prog_w0: progenitor phase-space point at ts[0], integrated forwards; ts: stripping times with
    ts[-1] the observation time.  Attributes: streamICs (pos_close, pos_far, vel_close, vel_far),
    IDs, prog_loc_fwd (len(ts), 6) and dRel_dIC (len(ts), 2, 6, 6)."""

    def __init__(self, potential_base, prog_w0, ts, Msat, seednum, solver=None, units=None,
                 dense=False, cpu=True, dt0=None, kval_arr=1.0, **kwargs):
        if dense:
            raise NotImplementedError("dense=True (diffrax interpolants) is not available")
        self.potential_base = potential_base
        self.units = units or usys
        self.prog_w0 = _lib.f64(prog_w0).reshape(6)
        self.ts = np.asarray(_lib.f64(ts).cpu().numpy(), dtype=np.float64).reshape(-1)
        self.Msat = Msat
        self.seednum = int(seednum)
        self.solver = _solvers.Dopri5() if solver is None else solver
        self.dt0 = dt0
        self.kval_arr = kval_arr
        self.dense = False
        self.stream_interp = None
        self.streamICs = potential_base.gen_stream_ics(ts=self.ts, prog_w0=self.prog_w0,
                                                       Msat=Msat, seed_num=self.seednum,
                                                       solver=self.solver, kval_arr=kval_arr,
                                                       dt0=dt0, **kwargs)
        self.IDs = torch.arange(len(self.ts), device=self.prog_w0.device)
        self.prog_loc_fwd = potential_base.integrate_orbit(
            w0=self.prog_w0, ts=self.ts, t0=self.ts.min(), t1=self.ts.max(), solver=self.solver,
            dt0=dt0, **kwargs).ys
        self.dRel_dIC = self.release_func_jacobian()

    def release_func_jacobian(self):
        """d(release rows)/d(progenitor row): (len(ts), 2, 6, 6), arm 0 = close (leading), arm 1
        = far (trailing), rows = [pos, vel] of the released particle (reference
        perturbative.py:282-296).  Twelve vector-Jacobian products of the release kernel."""
        from .adjoint import release_vjp
        prog = self.prog_loc_fwd.contiguous()
        n = prog.shape[0]
        dev = prog.device
        gms = self.potential_base._G * _lib.f64(
            np.broadcast_to(np.asarray(self.Msat, dtype=np.float64), (n,)).copy())
        idx = torch.arange(n, dtype=torch.int32, device=dev)
        J = torch.empty(n, 2, 6, 6, dtype=torch.float64, device=dev)
        zeros = torch.zeros(n, 3, dtype=torch.float64, device=dev)
        for arm in range(2):
            for row in range(6):
                cot = [zeros, zeros, zeros, zeros]  # pos_close, pos_far, vel_close, vel_far
                slot = (0 if arm == 0 else 1) + (0 if row < 3 else 2)
                e = torch.zeros(n, 3, dtype=torch.float64, device=dev)
                e[:, row % 3] = 1.0
                cot[slot] = e
                g_prog, _ = release_vjp(self.potential_base, prog, idx, gms, self.seednum, *cot,
                                        kval_arr=self.kval_arr)
                J[:, arm, row, :] = g_prog
        return J


class GenerateMassRadiusPerturbation:
    """Perturbation generator (reference perturbative.py:24-135).

    potential_base: smooth potential; potential_perturbation: the subhalos (Plummer or
    Hernquist sources: SubhaloLinePotential, SubhaloLinePotentialCustom_fromFunc, ...);
    potential_structural: their d Phi / d r_s (SubhaloLinePotential_dRadius,
    SubhaloLinePotentialCustom_dRadius_fromFunc) or None for the mass block alone."""

    def __init__(self, potential_base, potential_perturbation, potential_structural,
                 BaseStreamModel=None, units=None, solver=None, dt0=None, **kwargs):
        self.units = units or usys
        self.potential_base = potential_base
        self.potential_perturbation = potential_perturbation
        self.potential_structural = potential_structural
        self.gradientPotentialBase = potential_base.gradient
        self.gradientPotentialPerturbation = potential_perturbation.gradient
        src = potential_perturbation._kernel_sources()
        if src is None:
            raise NotImplementedError("the perturbing potential must consist of subhalo sources")
        self.num_pert = int(src.tensors()["gm"].numel())
        self.jump_ts = None
        if BaseStreamModel is None:
            return
        bs = BaseStreamModel
        self.base_stream = bs
        dt0 = dt0 if dt0 is not None else bs.dt0
        solver = solver if solver is not None else bs.solver  # the base model's method
        _method(solver, dt0)
        self.solver, self.dt0 = solver, dt0
        dev = bs.prog_loc_fwd.device
        self.field_wobs = [bs.prog_loc_fwd[-1],
                           torch.zeros(self.num_pert, 12, dtype=torch.float64, device=dev)]
        flipped = np.flip(bs.ts).copy()
        sol = integrate_field(w0=self.field_wobs, ts=flipped,
                              field=MassRadiusPerturbation_OTF(self), backwards_int=True,
                              solver=solver, dt0=dt0, **kwargs)
        self.prog_base = sol
        self.prog_fieldICs = torch.flip(sol.ys[1], dims=[0]).contiguous()  # (len(ts), Nsh, 12)
        self.perturbation_ICs_lead, self.perturbation_ICs_trail = self.compute_perturbation_ICs()
        pc, pf, vc, vf = bs.streamICs
        self.base_realspace_ICs_lead = torch.cat([pc, vc], dim=1)
        self.base_realspace_ICs_trail = torch.cat([pf, vf], dim=1)

    def compute_perturbation_ICs(self):
        """Particle response ICs = release Jacobian x progenitor responses, mass and radius
        blocks (reference perturbative.py:77-91).  Each (len(ts), Nsh, 12)."""
        J = self.base_stream.dRel_dIC
        F = self.prog_fieldICs
        out = []
        for arm in range(2):
            Ja = J[:, arm]  # (T, 6, 6)
            d_eps = torch.einsum("ijk,ilk->ilj", Ja, F[:, :, :6])
            d_rad = torch.einsum("ijk,ilk->ilj", Ja, F[:, :, 6:])
            out.append(torch.cat([d_eps, d_rad], dim=2).contiguous())
        return out[0], out[1]

    def compute_perturbation_OTF(self, cpu=True, solver=None, rtol=1e-6, atol=1e-6, dtmin=0.05,
                                 dtmax=None, max_steps=10_000, dt0=None):
        """Integrate every released particle and its responses from its release time to the
        observation time (reference perturbative.py:95-135).  Returns
        ([lead (T, 6), dlead (T, Nsh, 12)], [trail (T, 6), dtrail (T, Nsh, 12)]) with
        T = len(ts) - 1."""
        solver = self.solver if solver is None else solver
        dt0 = self.dt0 if dt0 is None else dt0
        method = _method(solver, dt0)
        bs = self.base_stream
        T = len(bs.ts) - 1
        base = self.potential_base
        src = self.potential_perturbation._kernel_sources()
        radius = self.potential_structural is not None
        dd = 12 if radius else 6
        if method == "leapfrog":
            from ._grid import release_grid
            t0, dt, nsteps, kidx = release_grid(bs.ts, dt0)
            start = kidx[:T].astype(np.int32)
        outs = []
        for w0, d0 in ((self.base_realspace_ICs_lead, self.perturbation_ICs_lead),
                       (self.base_realspace_ICs_trail, self.perturbation_ICs_trail)):
            if method == "leapfrog":
                w, d = perturb_leapfrog(base, src, w0[:T].contiguous(),
                                        d0[:T, :, :dd].contiguous(), start, t0, dt, nsteps,
                                        radius)
            else:  # each particle from its release time ts[i] to ts[-1]
                w, d, _, _, _ = perturb_dopri5(base, src, w0[:T].contiguous(),
                                               d0[:T, :, :dd].contiguous(), float(bs.ts[-1]),
                                               t_start=bs.ts[:T], with_radius=radius, rtol=rtol,
                                               atol=atol, dtmin=dtmin, dtmax=dtmax,
                                               max_steps=max_steps)
            if not radius:
                full = torch.zeros(T, self.num_pert, 12, dtype=torch.float64, device=d.device)
                full[:, :, :6] = d
                d = full
            outs.append([w, d])
        return outs[0], outs[1]
