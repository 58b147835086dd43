"""Solver selectors accepted by ``solver=`` (drop-in for the diffrax solver objects the
reference passes, e.g. ``diffrax.Dopri5(scan_kind='bounded')`` at reference
streamsculptor/main.py:288,344).

* ``Leapfrog()`` / ``"leapfrog"``: fixed-step kick-drift-kick leapfrog; needs ``dt0``.
* ``Dopri5()`` / ``"dopri5"``: adaptive Dormand-Prince 5(4) with an integral step-size controller.
* ``Dopri8()``: the reference's default for ``integrate_orbit``; not built on this path, so it
  raises ``NotImplementedError`` instead of silently running another method.

Objects of the real diffrax classes are recognised by class name.
"""
from __future__ import annotations


class Solver:
    name = "solver"

    def __init__(self, *args, **kwargs):  # scan_kind= etc. are accepted and ignored
        self.kwargs = kwargs

    def __repr__(self):
        return f"{type(self).__name__}()"

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self).__name__)


class Leapfrog(Solver):
    name = "leapfrog"


class Dopri5(Solver):
    name = "dopri5"


class Dopri8(Solver):
    name = "dopri8"


def resolve(solver, default="dopri5") -> str:
    """Map a solver argument (None, str, our classes, diffrax instances) to a method name."""
    if solver is None:
        name = default
    elif isinstance(solver, str):
        name = solver.lower()
    elif isinstance(solver, Solver):
        name = solver.name
    else:
        name = type(solver).__name__.lower()
    if name in ("leapfrog", "kdk", "verlet"):
        return "leapfrog"
    if name in ("dopri5", "tsit5_dopri5"):
        return "dopri5"
    if name == "dopri8":
        raise NotImplementedError(
            "Dopri8 is not built on the MI355X path; pass solver=Dopri5() (adaptive) or "
            "solver='leapfrog' with dt0= (fixed step).")
    raise NotImplementedError(f"solver {solver!r} is not available on the MI355X path")
