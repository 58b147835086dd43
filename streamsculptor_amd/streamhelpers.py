"""Perturbed-stream helpers (reference streamsculptor/streamhelpers.py:57-146).

Each stream runs on the gfx950 kernels through ``streams.gen_stream``: the progenitor orbit,
the release conditions and the particle orbits each use the potential the reference uses for
them.  ``pot_pert`` is typically a ``SubhaloLinePotential`` (LINE sources) or
``PlummerClumps`` (STATIC sources); the sum ``pot_base + pot_pert`` is formed with
``Potential_Combine`` exactly as in the reference.  Pass ``solver='leapfrog', dt0=...`` for the
fixed-step form.
"""
from __future__ import annotations

from . import solvers as _solvers
from .main import usys
from .potential import Potential_Combine
from .streams import gen_stream, stream_ics


def _total(pot_base, pot_pert):
    return Potential_Combine(potential_list=[pot_base, pot_pert], units=usys)


def gen_stream_ics_pert(pot_base=None, pot_pert=None, ts=None, prog_w0=None, Msat=None,
                        seed_num=None, solver=None, kval_arr=1.0, max_steps=10_000, rtol=1e-7,
                        atol=1e-7, dtmin=0.1, dt0=None):
    """Release conditions for a progenitor integrated in pot_base + pot_pert, with the release
    model evaluated in the smooth pot_base (reference streamhelpers.py:57-85).  Returns
    (pos_close, pos_far, vel_close, vel_far), each (len(ts), 3)."""
    solver = _solvers.Dopri5() if solver is None else solver
    ics, _, _ = stream_ics(_total(pot_base, pot_pert), pot_base, ts, prog_w0, Msat, seed_num,
                           solver=solver, kval_arr=kval_arr, dt0=dt0, rtol=rtol, atol=atol,
                           dtmin=dtmin, max_steps=max_steps)
    return tuple(ics)


def gen_stream_vmapped_with_pert(pot_base=None, pot_pert=None, ts=None, prog_w0=None, Msat=None,
                                 seed_num=None, solver=None, kval_arr=1.0, max_steps=10_000,
                                 rtol=1e-7, atol=1e-7, dtmin=0.1, dt0=None, throw=False):
    """Perturbed stream (reference streamhelpers.py:90-117): progenitor and particles in
    pot_base + pot_pert, release model in pot_base.  Returns (lead, trail), each (T, 6)."""
    solver = _solvers.Dopri5() if solver is None else solver
    total = _total(pot_base, pot_pert)
    return gen_stream(total, pot_base, ts, prog_w0, Msat, seed_num, solver=solver,
                      kval_arr=kval_arr, rtol=rtol, atol=atol, dtmin=dtmin, max_steps=max_steps,
                      throw=throw, dt0=dt0, particle_pot=total)


def gen_stream_vmapped_with_pert_fixed_prog(pot_base=None, pot_pert=None, ts=None, prog_w0=None,
                                            Msat=None, seed_num=None, solver=None, kval_arr=1.0,
                                            max_steps=10_000, rtol=1e-7, atol=1e-7, dtmin=0.1,
                                            dt0=None, throw=False):
    """Perturbed stream with an unperturbed progenitor (reference streamhelpers.py:120-146):
    progenitor and release in pot_base, particles in pot_base + pot_pert."""
    solver = _solvers.Dopri5() if solver is None else solver
    total = _total(pot_base, pot_pert)
    return gen_stream(pot_base, pot_base, ts, prog_w0, Msat, seed_num, solver=solver,
                      kval_arr=kval_arr, rtol=rtol, atol=atol, dtmin=dtmin, max_steps=max_steps,
                      throw=throw, dt0=dt0, particle_pot=total)


def eval_dense_stream(t_eval=None, dense_stream=None):
    """Leading and trailing arm of a dense stream at time t_eval (reference
    streamhelpers.py:24-30): (lead (T, 6), trail (T, 6))."""
    w = dense_stream.evaluate(t_eval)
    return w[0::2], w[1::2]


def eval_dense_stream_id(time=None, interp_func=None, idx=None, lead=True):
    """Trajectory of particle ``idx`` of one arm of a dense stream at ``time`` (reference
    streamhelpers.py:33-53)."""
    w = interp_func.evaluate(time)
    i = 2 * int(idx) + (0 if lead else 1)
    return w[i]

