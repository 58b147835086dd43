"""Mapping of save/release times onto the fixed-step grid t_k = t0 + k*dt."""
from __future__ import annotations

import numpy as np


def _round_checked(x: np.ndarray, what: str) -> np.ndarray:
    k = np.rint(x)
    if np.any(np.abs(x - k) > 1e-9 * np.maximum(1.0, np.abs(x))):
        raise ValueError(f"{what} must lie on the dt0 grid of the fixed-step solver")
    return k.astype(np.int64)


def leapfrog_grid(ts, t0: float, t1: float, dt0, steps: bool = False):
    """Return (dt, nsteps, save_idx) for integrating from t0 to t1 with step |dt0| in the
    direction of t1 - t0 and saving at ts (which must run in that direction)."""
    if dt0 is None:
        raise ValueError("the leapfrog solver needs a step size: pass dt0=")
    span = t1 - t0
    sgn = 1.0 if span >= 0 else -1.0
    dt = abs(float(dt0)) * sgn
    if dt == 0:
        raise ValueError("dt0 must be non-zero")
    nsteps = int(_round_checked(np.asarray([span / dt]), "t1 - t0")[0])
    if steps:
        return dt, nsteps, np.arange(nsteps + 1, dtype=np.int32)
    ts = np.asarray(ts, dtype=np.float64).reshape(-1)
    kidx = _round_checked((ts - t0) / dt, "save times ts")
    if np.any(kidx < 0) or np.any(kidx > nsteps):
        raise ValueError("save times ts must lie within [t0, t1]")
    if np.any(np.diff(kidx) < 0):
        raise ValueError("save times ts must be ordered along the direction of integration")
    return dt, nsteps, kidx.astype(np.int32)


def release_grid(ts, dt0):
    """For a spray stream on ts: (t0, dt, nsteps, grid index of every ts)."""
    ts = np.asarray(ts, dtype=np.float64).reshape(-1)
    t0, t1 = float(ts[0]), float(ts[-1])
    dt, nsteps, kidx = leapfrog_grid(ts, t0, t1, dt0)
    return t0, dt, nsteps, kidx
