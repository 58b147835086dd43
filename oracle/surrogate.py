"""ORACLE -- test infrastructure only: numpy restatement of the leapfrog with the learned
clump force (analytic host + clump force from the MLP, the exact clump sum in flagged cells and
outside the residual grid) and of its discrete adjoint.

The reference has no surrogate and differentiates forward-mode only (streamsculptor/main.py:160);
this follows the surrogate stream workload's specification.  The integrator is the
kick-drift-kick map of ``oracle.leapfrog`` and the adjoint recursion is the one of
``oracle.adjoint_leapfrog`` (ssc_oracle.cpp, orc_adjoint_leapfrog); only the clump part of the
force and of the Hessian-vector product changes:

  clump gradient    g_scale * MLP(u)[1:4],  u = float32((x - center) * (1 / half))
                    (float64 u with the smooth MLP, emulate=False)
  clump H W         g_scale * (1 / half) * J_u^T (0, W)   (the MLP input VJP, oracle/mlp.py)
  fallback rows     exact Plummer sum and its Hessian (oracle.evaluate with the clump table)

Pinned by tests/test_oracle_surrogate.py: with every cell flagged the adjoint equals
oracle.adjoint_leapfrog over the exact clump sum; with the smooth float64 MLP (emulate=False)
it matches central finite differences of the forward map in the initial rows and the halo
parameters; the MLP VJP itself is pinned against torch autograd and finite differences in
tests/test_oracle_mlp.py.
"""
from __future__ import annotations

import numpy as np

from . import evaluate, host as make_host, host_dgrad
from . import mlp as OM


class Model:
    """Host descriptor ``h`` (oracle.host), clump table ``clumps`` (oracle.Sources, STATIC),
    flat MLP parameters, input scaling, residual grid (lo, cell size, cells per axis) and its
    flags (uint8 [nx, ny, nz], 1 = exact sum)."""

    def __init__(self, h, clumps, params, center, half, g_scale, grid_lo, grid_cell, grid_n,
                 flags):
        self.h = h
        self.clumps = clumps
        self.p = np.asarray(params, dtype=np.float32)
        self.center = np.asarray(center, dtype=np.float64)
        self.inv_half = 1.0 / np.asarray(half, dtype=np.float64)
        self.g_scale = float(g_scale)
        self.lo = np.asarray(grid_lo, dtype=np.float64)
        self.inv_cell = 1.0 / np.asarray(grid_cell, dtype=np.float64)
        self.n = np.asarray(grid_n, dtype=np.int64)
        self.flags = np.asarray(flags, dtype=np.uint8).reshape(tuple(self.n))
        self.h0 = make_host([])

    def fallback(self, x):
        c = (x - self.lo) * self.inv_cell
        inside = ((c >= 0.0) & (c < self.n)).all(1)
        ijk = np.clip(np.floor(np.where(inside[:, None], c, 0.0)).astype(np.int64), 0, self.n - 1)
        flagged = self.flags[ijk[:, 0], ijk[:, 1], ijk[:, 2]] != 0
        return (~inside) | flagged

    def inputs(self, x, emulate=True):
        u = (x - self.center) * self.inv_half
        return u.astype(np.float32) if emulate else u

    def accel(self, x, emulate=True):
        _, gh, _ = evaluate(self.h, x)
        g = np.zeros_like(x)
        fb = self.fallback(x)
        if (~fb).any():
            y, _ = OM.forward(self.p, self.inputs(x[~fb], emulate), emulate)
            g[~fb] = self.g_scale * y[:, 1:]
        if fb.any():
            _, gc, _ = evaluate(self.h0, x[fb], src=self.clumps)
            g[fb] = gc
        return -(gh + g), fb

    def hess_w(self, x, W, emulate=True):
        """(H_host + H_clump) W at each row."""
        _, _, Hh = evaluate(self.h, x)
        out = np.einsum("nij,nj->ni", Hh, W)
        fb = self.fallback(x)
        if (~fb).any():
            dy = np.zeros((int((~fb).sum()), 4))
            dy[:, 1:] = W[~fb]
            dx = OM.input_vjp(self.p, self.inputs(x[~fb], emulate), dy, emulate)
            out[~fb] += self.g_scale * self.inv_half * dx
        if fb.any():
            _, _, Hc = evaluate(self.h0, x[fb], src=self.clumps)
            out[fb] += np.einsum("nij,nj->ni", Hc, W[fb])
        return out


def leapfrog(model, w0, dt, nsteps, start=None, emulate=True, traj=False):
    """Rows w0 (n, 6) from their start steps to step nsteps.  Returns (w_final, fallback
    particle-steps) or, with traj=True, also the positions (nsteps + 1, n, 3) (rows before a
    particle's start hold its initial position)."""
    w0 = np.atleast_2d(np.asarray(w0, dtype=np.float64))
    n = w0.shape[0]
    s = np.zeros(n, dtype=np.int64) if start is None else np.clip(np.asarray(start), 0, nsteps)
    x, v = w0[:, :3].copy(), w0[:, 3:].copy()
    X = np.zeros((nsteps + 1, n, 3)) if traj else None
    if traj:
        X[:] = x[None]
    kmin = int(s.min()) if n else nsteps
    nfb = 0
    h = 0.5 * dt
    a = np.zeros_like(x)
    if kmin < nsteps:
        a, _ = model.accel(x, emulate)
    for k in range(kmin, nsteps):
        act = (k >= s)[:, None]
        v = np.where(act, v + h * a, v)
        x = np.where(act, x + dt * v, x)
        a, fb = model.accel(x, emulate)
        nfb += int((fb & act[:, 0]).sum())
        v = np.where(act, v + h * a, v)
        if traj:
            X[k + 1] = x
    w = np.hstack([x, v])
    return (w, nfb, X) if traj else (w, nfb)


def adjoint(model, w0, dt, nsteps, lam_T, start=None, forcing=None, emulate=True):
    """Discrete adjoint of ``leapfrog``: dict(lam0 (n, 6), g_host (n_comp, 2) = dL/d(G m_c),
    dL/d(p1_c)) for dL/dw_final = lam_T and optional per-row forcing (nsteps + 1, n, 6)."""
    w0 = np.atleast_2d(np.asarray(w0, dtype=np.float64))
    n = w0.shape[0]
    s = np.zeros(n, dtype=np.int64) if start is None else np.clip(np.asarray(start), 0, nsteps)
    _, _, X = leapfrog(model, w0, dt, nsteps, start=s, emulate=emulate, traj=True)
    nc = model.h.n_components
    gh = np.zeros((nc, 2))
    h = 0.5 * dt
    lam_T = np.asarray(lam_T, dtype=np.float64).reshape(n, 6)

    def F(k):
        return np.zeros((n, 6)) if forcing is None else np.asarray(forcing[k])

    def apply(k, W, lx, rows):
        if not rows.any():
            return
        xk = X[k][rows]
        Wr = W[rows]
        lx[rows] -= model.hess_w(xk, Wr, emulate)
        dg = host_dgrad(model.h, xk)
        gh[:, 0] -= np.einsum("ncj,nj->c", dg[:, :, 0], Wr)
        gh[:, 1] -= np.einsum("ncj,nj->c", dg[:, :, 1], Wr)

    f = F(nsteps)
    lx = lam_T[:, :3] + f[:, :3]
    lv = lam_T[:, 3:] + f[:, 3:]
    apply(nsteps, h * lv, lx, s < nsteps)
    for k in range(nsteps - 1, int(s.min()) - 1 if n else nsteps, -1):
        rows = k >= s
        f = F(k)
        lh = lv + dt * lx
        lv_new = lh + f[:, 3:]
        lx_new = lx + f[:, :3]
        W = h * lh + np.where((k > s)[:, None], h * lv_new, 0.0)
        lv = np.where(rows[:, None], lv_new, lv)
        lx = np.where(rows[:, None], lx_new, lx)
        apply(k, W, lx, rows)
    return dict(lam0=np.hstack([lx, lv]), g_host=gh)
