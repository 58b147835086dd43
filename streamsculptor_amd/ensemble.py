"""Differentiable ensembles of spray streams in the clumpy halo with the learned clump force.

Many progenitors (each with its own initial row and stream age) are integrated in one pass:
every progenitor orbit starts at its own grid row of a common leapfrog grid t0 = -T dt .. 0 and
releases a fixed number of spray pairs at evenly spaced steps of its life (reference release
model, streamsculptor/main.py:209-306, one seed for the whole ensemble as under ``jax.vmap``
over progenitors); the particles run from their release steps to t = 0 in the host plus the
surrogate clump force (exact sum where the residual map flags it).

``ensemble_loss`` is the squared distance of the final particle positions to a target and is
differentiable, through the discrete adjoints, with respect to the halo mass and scale radius
and every progenitor's initial row.  The particles use the surrogate adjoint (network input VJP
for the clump Hessian); the release conditions the release-model VJP; the progenitor orbits,
which feel the exact clump sum, the plain adjoint, forced at the release rows.  FP64 throughout
the orbit and adjoint state.
"""
from __future__ import annotations

import copy

import numpy as np
import torch

from . import _lib
from .adjoint import LeapfrogAdjoint, _with_sources, release_vjp

G = _lib.G_USYS


class SurrogateStreamEnsemble:
    """Fixed inputs of the ensemble: the trained ``ClumpSurrogate`` (its host is replaced by
    the halo parameters of each call), per-progenitor stream lengths in grid steps,
    ``n_release`` release steps per progenitor (two particles each), Msat and the release
    seed."""

    def __init__(self, sur, lengths, n_release=1000, msat=1e4, seed=0, dt=1.0, halo_index=2,
                 base_components=None, checkpoint_every=100, kval_arr=1.0):
        from . import workloads as W
        self.sur = sur
        L = np.asarray(lengths, dtype=np.int64)
        if (L < n_release).any():
            raise ValueError("every stream needs at least n_release steps")
        self.lengths = L
        self.S = len(L)
        self.R = int(n_release)
        self.T = int(L.max())
        self.dt = float(dt)
        self.t0 = -self.T * self.dt
        self.msat = float(msat)
        self.seed = int(seed)
        self.E = int(checkpoint_every)
        self.kval_arr = kval_arr
        self.base = base_components or W.host_components()
        self.halo_index = halo_index
        self.prog_start = (self.T - L).astype(np.int32)                     # (S,)
        i = np.arange(self.R)
        self.rel_step = (self.prog_start[:, None] + (i[None, :] * L[:, None]) // self.R)  # (S, R)
        rel = self.rel_step.reshape(-1)
        start = np.concatenate([rel, rel]).astype(np.int32)                 # lead block, trail block
        self.order = np.argsort(start, kind="stable")                       # particle launch order
        self.start_sorted = start[self.order]
        self.n = start.size
        self.particle_steps = int((self.T - start).sum())

    # ------------------------------------------------------------------ model pieces
    def host(self, m_halo, rs_halo):
        from .main import Potential
        comps = [list(c) for c in self.base]
        comps[self.halo_index] = (self.base[self.halo_index][0], [G * float(m_halo), float(rs_halo)])

        class _Host(Potential):
            def __init__(self, cs):
                super().__init__(None, {})
                self._G = G
                self._cs = cs

            def _kernel_components(self):
                return self._cs

        return _Host(comps)

    def _surrogate_for(self, host):
        s = copy.copy(self.sur)
        s.host = host
        return s

    def forward(self, m_halo, rs_halo, prog_w0):
        """Final particle rows (in ensemble order: lead particles of every (stream, release),
        then the trail particles) and everything the backward pass needs."""
        host = self.host(m_halo, rs_halo)
        src = self.sur.clumps._kernel_sources()
        total = _with_sources(host, src)
        prog_w0 = _lib.f64(prog_w0).reshape(self.S, 6).contiguous()
        dev = prog_w0.device
        T = self.T
        _, ys = total._leapfrog(prog_w0, self.t0, self.dt, T, start=self.prog_start,
                                save_idx=np.arange(T + 1, dtype=np.int32))
        rows = torch.as_tensor(self.rel_step.reshape(-1), device=dev)
        sidx = torch.arange(self.S, device=dev).repeat_interleave(self.R)
        prog_rows = ys[rows, sidx].contiguous()                              # (S R, 6)
        del ys
        rel_idx = torch.arange(self.R, dtype=torch.int32, device=dev).repeat(self.S)
        gms = torch.full((self.S * self.R,), G * self.msat, dtype=torch.float64, device=dev)
        pc, pf, vc, vf = host._release(prog_rows, rel_idx, gms, self.seed, self.kval_arr)
        w0 = torch.cat([torch.cat([pc, vc], 1), torch.cat([pf, vf], 1)], 0)
        order = torch.as_tensor(self.order, device=dev)
        sadj = self._surrogate_for(host).adjoint(w0[order], self.t0, self.dt, T,
                                                 start=self.start_sorted,
                                                 checkpoint_every=self.E)
        wT_sorted = sadj.forward()
        wT = torch.empty_like(wT_sorted)
        wT[order] = wT_sorted
        return dict(host=host, src=src, total=total, prog_rows=prog_rows, rel_idx=rel_idx,
                    gms=gms, sadj=sadj, order=order, wT=wT, prog_w0=prog_w0, rows=rows,
                    sidx=sidx)

    def backward(self, st, lam_T):
        """(dL/dm_halo, dL/drs_halo, dL/dprog_w0 (S, 6)) for dL/d(final rows) = lam_T."""
        T, S, R = self.T, self.S, self.R
        hi = self.halo_index
        order = st["order"]
        A = st["sadj"].backward(lam_T[order])
        lam0 = torch.empty_like(A["lam0"])
        lam0[order] = A["lam0"]
        g_host = A["g_host"].clone()
        nr = S * R
        gp, gh = release_vjp(st["host"], st["prog_rows"], st["rel_idx"], st["gms"], self.seed,
                             lam0[:nr, :3], lam0[nr:, :3], lam0[:nr, 3:], lam0[nr:, 3:],
                             self.kval_arr)
        g_host += gh
        dev = lam_T.device
        F = torch.zeros(T + 1, S, 6, dtype=torch.float64, device=dev)
        F.index_put_((st["rows"], st["sidx"]), gp, accumulate=True)
        pa = LeapfrogAdjoint(st["total"], st["prog_w0"], self.t0, self.dt, T,
                             start=self.prog_start, checkpoint_every=self.E, src=st["src"])
        pa.forward()
        C = pa.backward(torch.zeros(S, 6, dtype=torch.float64, device=dev), forcing=F,
                        want_sources=False)
        g_host += C["g_host"]
        return G * g_host[hi, 0], g_host[hi, 1], C["lam0"]


class _EnsembleLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, halo, prog_w0, model, target):
        st = model.forward(halo[0], halo[1], prog_w0)
        d = st["wT"][:, :3] - target
        loss = (d * d).sum()
        ctx.model, ctx.st, ctx.d = model, st, d
        return loss

    @staticmethod
    def backward(ctx, gout):
        lam_T = torch.zeros_like(ctx.st["wT"])
        lam_T[:, :3] = 2.0 * ctx.d
        gm_h, grs_h, g_w0 = ctx.model.backward(ctx.st, lam_T)
        go = gout.to(torch.float64)
        return torch.stack([gm_h, grs_h]) * go, g_w0 * go, None, None


def ensemble_loss(halo, prog_w0, model, target):
    """L = sum_i |x_i(0) - target_i|^2 over the ensemble's particles, differentiable with respect
    to halo = (m_halo, r_s_halo) and the progenitors' initial rows (S, 6) (torch tensors on the
    GPU)."""
    return _EnsembleLoss.apply(halo, prog_w0, model, _lib.f64(target))
