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
    seed.  The progenitor rows are given at each progenitor's own first grid row
    (t = -length dt)."""

    def __init__(self, sur, lengths, n_release=1000, msat=1e4, seed=0, dt=1.0, halo_index=2,
                 base_components=None, checkpoint_every=100, kval_arr=1.0, clumps=None):
        from . import workloads as W
        self.sur = sur            # may be set after construction (e.g. trained on this ensemble)
        self.clumps = clumps      # clump Potential when sur is None
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

    def particles(self, host, prog_w0, want_positions=False):
        """Progenitor orbits (host + exact clump sum, every grid row kept) and the released
        particle rows in ensemble order: the lead particle of every (stream, release), then
        the trail particles.  want_positions adds the progenitors' positions at every row
        after their starts."""
        src = self.sur_clumps()._kernel_sources()
        total = _with_sources(host, src)
        prog_w0 = _lib.f64(prog_w0).reshape(self.S, 6).contiguous()
        dev = prog_w0.device
        T = self.T
        _, ys = total._leapfrog(prog_w0, self.t0, self.dt, T, start=self.prog_start,
                                save_idx=np.arange(T + 1, dtype=np.int32))
        rows = torch.as_tensor(self.rel_step.reshape(-1), device=dev)
        sidx = torch.arange(self.S, device=dev).repeat_interleave(self.R)
        prog_rows = ys[rows, sidx].contiguous()                              # (S R, 6)
        prog_pos = None
        if want_positions:
            keep = (torch.arange(T + 1, device=dev)[:, None]
                    >= torch.as_tensor(self.prog_start, device=dev)[None, :])
            prog_pos = ys[..., :3][keep]
        del ys
        rel_idx = torch.arange(self.R, dtype=torch.int32, device=dev).repeat(self.S)
        gms = torch.full((self.S * self.R,), G * self.msat, dtype=torch.float64, device=dev)
        pc, pf, vc, vf = host._release(prog_rows, rel_idx, gms, self.seed, self.kval_arr)
        w0 = torch.cat([torch.cat([pc, vc], 1), torch.cat([pf, vf], 1)], 0)
        return dict(src=src, total=total, prog_w0=prog_w0, prog_rows=prog_rows, rows=rows,
                    sidx=sidx, rel_idx=rel_idx, gms=gms, w0=w0, prog_pos=prog_pos)

    def sur_clumps(self):
        return self.clumps if self.sur is None else self.sur.clumps

    def training_positions(self, host, prog_w0, n_sample=1 << 16, stride=16, seed=0):
        """Positions along plain-method (exact clump sum) orbits of the ensemble: every row of
        every progenitor orbit and every ``stride``-th row of ``n_sample`` random particles."""
        from .surrogate import orbit_positions
        P = self.particles(host, prog_w0, want_positions=True)
        rng = np.random.default_rng(seed)
        sel = np.sort(rng.choice(self.n, size=min(n_sample, self.n), replace=False))
        start = np.concatenate([self.rel_step.reshape(-1)] * 2).astype(np.int32)
        pos = orbit_positions(P["total"], P["w0"][torch.as_tensor(sel, device=P["w0"].device)],
                              start[sel], self.t0, self.dt, self.T, stride=stride)
        return torch.cat([P["prog_pos"], pos], 0)

    def forward(self, m_halo, rs_halo, prog_w0, exact=False):
        """Final particle rows (ensemble order) and everything the backward pass needs.
        exact=True integrates the particles with the exact clump sum (plain adjoint) instead
        of the surrogate."""
        host = self.host(m_halo, rs_halo)
        st = self.particles(host, prog_w0)
        st.pop("prog_pos")
        order = torch.as_tensor(self.order, device=st["w0"].device)
        w0s = st.pop("w0")[order]
        if exact:
            adj = LeapfrogAdjoint(st["total"], w0s, self.t0, self.dt, self.T,
                                  start=self.start_sorted, checkpoint_every=self.E, src=st["src"])
        else:
            adj = self._surrogate_for(host).adjoint(w0s, self.t0, self.dt, self.T,
                                                    start=self.start_sorted,
                                                    checkpoint_every=self.E)
        del w0s
        wT_sorted = adj.forward()
        wT = torch.empty_like(wT_sorted)
        wT[order] = wT_sorted
        st.update(host=host, padj=adj, exact=exact, order=order, wT=wT)
        return st

    def backward(self, st, lam_T, events=None):
        """(dL/dm_halo, dL/drs_halo, dL/dprog_w0 (S, 6)) for dL/d(final rows) = lam_T.
        ``events`` (start, end) are recorded around the particles' adjoint."""
        T, S, R = self.T, self.S, self.R
        hi = self.halo_index
        order = st["order"]
        if events:
            events[0].record()
        if st["exact"]:
            A = st["padj"].backward(lam_T[order], want_sources=False)
        else:
            A = st["padj"].backward(lam_T[order])
        if events:
            events[1].record()
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
    def forward(ctx, halo, prog_w0, model, target, exact, events):
        st = model.forward(halo[0], halo[1], prog_w0, exact=exact)
        d = st["wT"][:, :3] - target
        loss = (d * d).sum()
        ctx.model, ctx.st, ctx.d, ctx.events = model, st, d, events
        return loss

    @staticmethod
    def backward(ctx, gout):
        lam_T = torch.zeros_like(ctx.st["wT"])
        lam_T[:, :3] = 2.0 * ctx.d
        gm_h, grs_h, g_w0 = ctx.model.backward(ctx.st, lam_T, events=ctx.events)
        ctx.st = None
        go = gout.to(torch.float64)
        return torch.stack([gm_h, grs_h]) * go, g_w0 * go, None, None, None, None


def ensemble_loss(halo, prog_w0, model, target, exact=False, events=None):
    """L = sum_i |x_i(0) - target_i|^2 over the ensemble's particles, differentiable with respect
    to halo = (m_halo, r_s_halo) and the progenitors' initial rows (S, 6) (torch tensors on the
    GPU).  exact=True: particles in the exact clump sum instead of the surrogate -- on the
    2000-progenitor clumpy-halo ensemble the surrogate's gradient differs from the exact one by
    12-17 % (halo) and 83 % (progenitor rows, normwise) and takes 2.8x as long, so use
    exact=True where the gradient matters.  ``events`` (start, end) are recorded around the
    particles' adjoint in the backward pass."""
    return _EnsembleLoss.apply(halo, prog_w0, model, _lib.f64(target), exact, events)
