"""Reverse-mode gradients of fixed-step spray streams (discrete leapfrog adjoint).

The reference differentiates forward-mode only (``jax.jacfwd`` through diffrax ForwardMode,
reference streamsculptor/main.py:160); this module provides the discrete adjoint of the
leapfrog map with checkpoints every ``checkpoint_every`` steps, FP64 throughout.

``stream_loss`` differentiates the squared distance of a perturbed spray stream's final
positions to a target stream with respect to the halo mass and scale radius, the masses of
live subhalos and the progenitor's initial phase-space row.  All dependencies are
differentiated: particle forces, the release conditions (through the tidal radius, i.e. third
derivatives of the host potential), the progenitor orbit (which feels the subhalos) and the
subhalo orbits (which feel the host).  Subhalo scale radii are held fixed.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .main import SourceSet

G = _lib.G_USYS


class LeapfrogAdjoint:
    """Checkpointed forward/backward of ssc_leapfrog for a set of orbits in ``pot``."""

    def __init__(self, pot, w0, t0, dt, nsteps, start=None, checkpoint_every=100, src=None):
        self.h, psrc = pot._descriptor()
        self.src = src if src is not None else psrc
        self.sd = self.src.desc() if self.src is not None else None
        self.w0 = _lib.f64(w0).reshape(-1, 6).contiguous()
        self.n = self.w0.shape[0]
        self.start = None if start is None else _lib.i32(start)
        self.t0, self.dt, self.nsteps, self.E = float(t0), float(dt), int(nsteps), int(checkpoint_every)
        m = 0 if self.src is None else self.src.tensors()["gm"].numel()
        self.m = m
        nbytes = _lib.load().ssc_adjoint_workspace_bytes(self.n, self.nsteps, self.E, m)
        self.ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=self.w0.device)
        self.w_out = None

    def _srcp(self):
        return None if self.sd is None else _lib.ctypes.byref(self.sd)

    def forward(self):
        self.w_out = torch.empty_like(self.w0)
        _lib.call("ssc_adjoint_forward", _lib.ctypes.byref(self.h), self._srcp(), _lib.ptr(self.w0),
                  _lib.ptr(self.start), self.n, self.t0, self.dt, self.nsteps, self.E,
                  _lib.ptr(self.w_out), _lib.ptr(self.ws), self.ws.numel())
        return self.w_out

    def forward_from_rows(self, rows):
        """Take the checkpoints from orbits already integrated with every grid row saved,
        rows (nsteps + 1, n, 6) (a leapfrog with save_idx = 0..nsteps: the same arithmetic as
        ``forward``), instead of integrating again."""
        rows = _lib.f64(rows).reshape(self.nsteps + 1, self.n, 6).contiguous()
        _lib.call("ssc_adjoint_checkpoints_from_rows", _lib.ptr(rows), self.n, self.nsteps, self.E,
                  self.m, _lib.ptr(self.ws), self.ws.numel())
        self.w_out = rows[-1].clone()
        return self.w_out

    def backward(self, lam_T, forcing=None, want_sources=True, rows=None):
        """Adjoint sweep; rows = (k_lo, k_hi) sweeps only those grid rows (called from the last
        range down, see ssc_adjoint_backward_range), returning that range's share."""
        lam_T = _lib.f64(lam_T).reshape(self.n, 6).contiguous()
        F = None if forcing is None else _lib.f64(forcing).contiguous()
        dev = lam_T.device
        lam0 = torch.empty_like(lam_T)
        nc = self.h.n_components
        gh = torch.zeros(max(nc, 1), 2, dtype=torch.float64, device=dev)
        gg = gp = None
        if want_sources and self.m:
            gg = torch.zeros(self.m, dtype=torch.float64, device=dev)
            prow = self.nsteps + 1 if self.src.kind == _lib.SRC_TRACK else 1
            gp = torch.zeros(prow, self.m, 3, dtype=torch.float64, device=dev)
        k_lo, k_hi = (0, self.nsteps) if rows is None else rows
        _lib.call("ssc_adjoint_backward_range", _lib.ctypes.byref(self.h), self._srcp(),
                  _lib.ptr(self.start), self.n, self.t0, self.dt, self.nsteps, self.E, int(k_lo),
                  int(k_hi), _lib.ptr(lam_T), _lib.ptr(F), _lib.ptr(lam0), _lib.ptr(gh),
                  _lib.ptr(gg), _lib.ptr(gp), _lib.ptr(self.ws), self.ws.numel())
        return dict(lam0=lam0, g_host=gh[:nc], g_gm=gg, g_pos=gp)


def _release_vjp_into(pot, prog_w, rel_idx, g_msat, seed, kval_arr, g_pc, g_pf, g_vc, g_vf,
                      g_prog, g_host, scratch):
    """ssc_release_vjp into caller-provided outputs (contiguous device tensors; scratch of
    at least (nrel, 2 n_components) doubles)."""
    h, _ = pot._descriptor()
    kv = np.asarray(kval_arr, dtype=np.float64)
    kvp = None
    if kv.ndim > 0:
        kv = np.ascontiguousarray(np.ones(8) * kv)
        kvp = kv.ctypes.data_as(_lib.ctypes.c_void_p)
    _lib.call("ssc_release_vjp", _lib.ctypes.byref(h), _lib.ptr(prog_w), _lib.ptr(rel_idx),
              _lib.ptr(g_msat), prog_w.shape[0], int(seed), kvp, _lib.ptr(g_pc), _lib.ptr(g_pf),
              _lib.ptr(g_vc), _lib.ptr(g_vf), _lib.ptr(g_prog), _lib.ptr(g_host),
              _lib.ptr(scratch))


def release_vjp(pot, prog_w, rel_idx, g_msat, seed, g_pc, g_pf, g_vc, g_vf, kval_arr=1.0):
    """dL/d(progenitor rows) (nrel, 6) and dL/d(host G m, p1) (n_comp, 2) of the release model
    in the analytic potential ``pot`` (the release of the perturbed streams, reference
    streamhelpers.py:57-66, which these gradients follow)."""
    h, src = pot._descriptor()
    if src is not None:
        raise NotImplementedError("the release-model VJP covers analytic potentials only: "
                                  "release in the smooth base potential")
    prog_w = _lib.f64(prog_w).contiguous()
    nrel = prog_w.shape[0]
    dev = prog_w.device
    kv = np.asarray(kval_arr, dtype=np.float64)
    kvp = None
    if kv.ndim > 0:
        kv = np.ascontiguousarray(np.ones(8) * kv)
        kvp = kv.ctypes.data_as(_lib.ctypes.c_void_p)
    g_prog = torch.empty(nrel, 6, dtype=torch.float64, device=dev)
    gh = torch.zeros(max(h.n_components, 1), 2, dtype=torch.float64, device=dev)
    scratch = torch.empty(nrel, 2 * max(h.n_components, 1), dtype=torch.float64, device=dev)
    _lib.call("ssc_release_vjp", _lib.ctypes.byref(h), _lib.ptr(prog_w), _lib.ptr(_lib.i32(rel_idx)),
              _lib.ptr(_lib.f64(g_msat)), nrel, int(seed), kvp,
              *[_lib.ptr(_lib.f64(a).contiguous()) for a in (g_pc, g_pf, g_vc, g_vf)],
              _lib.ptr(g_prog), _lib.ptr(gh), _lib.ptr(scratch))
    return g_prog, gh[:h.n_components]


# ------------------------------------------------------------------------------------------
# the perturbed-stream loss (live subhalos, direct N x M sums)
# ------------------------------------------------------------------------------------------
class StreamModel:
    """Fixed inputs of a perturbed spray stream: host components other than the halo, the
    subhalo initial rows and radii, the release grid and the spray layout.

    ``checkpoint_every`` (default 100) keeps a checkpoint every that many steps and re-runs each
    segment in the backward pass; 1 keeps every particle row of the forward pass instead (48 B
    per particle per step, 19.2 GB for 1e5 particles over 4000 steps) and re-runs nothing.
    ``pipeline`` row ranges of the backward pass overlap the particles' sweep with the release
    model and the progenitor and subhalo adjoints of the ranges already swept."""

    def __init__(self, ts, sub_w0, sub_rs, msat, per_step, seed=0, dt=1.0, halo_index=2,
                 base_components=None, checkpoint_every=100, kval_arr=1.0, pipeline=8):
        from . import workloads as W
        self.ts = np.asarray(ts, dtype=np.float64)
        self.t0 = float(self.ts[0])
        self.dt = float(dt)
        self.nsteps = int(round((self.ts[-1] - self.ts[0]) / dt))
        self.sub_w0 = _lib.f64(sub_w0).reshape(-1, 6)
        self.sub_rs = _lib.f64(sub_rs).reshape(-1)
        self.msat = float(msat)
        self.per_step = int(per_step)
        self.seed = int(seed)
        self.E = int(checkpoint_every)
        self.pipeline = max(1, int(pipeline))
        self.kval_arr = kval_arr
        self.base = base_components or W.host_components()
        self.halo_index = halo_index

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

    def track_sources(self, host, sub_masses):
        _, ys = host._leapfrog(self.sub_w0, self.t0, self.dt, self.nsteps,
                               save_idx=np.arange(self.nsteps + 1, dtype=np.int32))
        src = SourceSet(_lib.SRC_TRACK, gm=G * _lib.f64(sub_masses), rs=self.sub_rs,
                        pos=ys[:, :, :3].contiguous(), grid=(self.t0, self.dt))
        return src, ys

    def layout(self):
        from .workloads import spray_layout
        even, odd = spray_layout(self.per_step)
        T = self.nsteps
        lay = np.array([even, odd], dtype=np.int64)      # (parity, per_step, [seed, arm])
        par = np.arange(T) % 2
        rel = np.repeat(np.arange(T), self.per_step)
        return lay[par, :, 0].reshape(-1), lay[par, :, 1].reshape(-1), rel

    def _device_layout(self, dev):
        """The spray layout's index tensors and the release masses on the device, made once
        (every loss evaluation uses the same)."""
        key = str(dev)
        if getattr(self, "_dl_key", None) != key:
            T = self.nsteps
            sel_s, sel_a, rel = self.layout()
            si, ai, ri = (torch.as_tensor(a, device=dev) for a in (sel_s, sel_a, rel))
            # rows of (seed, arm, release) in the backward pass's adjoint table
            # [seed][arm][x|v][row][3] viewed as rows of 3: positions, and velocities T + 1 on
            flat = (si * 4 + ai * 2) * (T + 1) + ri
            self._dl = dict(si=si, ai=ai, ri=ri, flat_x=flat, flat_v=flat + (T + 1),
                            start=_lib.i32(rel.astype(np.int32)),
                            idx=torch.arange(T + 1, dtype=torch.int32, device=dev),
                            gms=torch.full((T + 1,), G * self.msat, dtype=torch.float64,
                                           device=dev))
            self._dl_key = key
        return self._dl

    def forward(self, m_halo, rs_halo, sub_masses, prog_w0):
        """Final particle rows and everything the backward pass needs."""
        host = self.host(m_halo, rs_halo)
        src, sub_traj = self.track_sources(host, sub_masses)
        total = _with_sources(host, src)
        T = self.nsteps
        _, pys = total._leapfrog(_lib.f64(prog_w0).reshape(1, 6), self.t0, self.dt, T,
                                 save_idx=np.arange(T + 1, dtype=np.int32))
        prog_rows = pys[:, 0, :].contiguous()
        nseed = (self.per_step + 1) // 2
        dev = prog_rows.device
        dl = self._device_layout(dev)
        table = host._release_seeds(prog_rows, dl["idx"], dl["gms"], self.seed, nseed,
                                    self.kval_arr)
        si, ai, ri = dl["si"], dl["ai"], dl["ri"]
        w0 = table[si, ai, ri].contiguous()
        start = dl["start"]
        padj = LeapfrogAdjoint(total, w0, self.t0, self.dt, T, start=start,
                               checkpoint_every=self.E, src=src)
        wT = padj.forward()
        # the progenitor's and the subhalos' rows are all saved already: they are the
        # checkpoints (every row stored)
        pa = LeapfrogAdjoint(total, _lib.f64(prog_w0).reshape(1, 6), self.t0, self.dt, T,
                             checkpoint_every=1, src=src)
        pa.forward_from_rows(pys)
        sa = LeapfrogAdjoint(host, self.sub_w0, self.t0, self.dt, T, checkpoint_every=1)
        sa.forward_from_rows(sub_traj)
        return dict(host=host, src=src, total=total, prog_rows=prog_rows, wT=wT, padj=padj,
                    pa=pa, sa=sa, idx=(si, ai, ri), nseed=nseed,
                    prog_w0=_lib.f64(prog_w0).reshape(1, 6))

    def _ranges(self):
        """Row ranges (k_lo, k_hi) of the pipelined backward pass, from the last."""
        T = self.nsteps
        J = min(self.pipeline, T)
        q = 1 if self.E == 1 else self.E
        cuts = sorted({0, T} | {min(T, (T * j // J) // q * q) for j in range(1, J)})
        return [(cuts[i], cuts[i + 1]) for i in range(len(cuts) - 1)][::-1]

    def backward(self, st, lam_T):
        """Gradients of L with dL/d(final rows) = lam_T: (dL/dm_halo, dL/drs_halo, dL/dm_j,
        dL/dprog_w0).

        The particles' sweep runs on the current stream one row range at a time; after each
        range, a second stream takes the release-model VJP of the releases in that range (their
        adjoints are final once the sweep has passed their release rows) and the progenitor and
        subhalo adjoints over the same rows, while the current stream sweeps the next range."""
        T = self.nsteps
        host, src = st["host"], st["src"]
        hi = self.halo_index
        dev = lam_T.device
        M = self.sub_w0.shape[0]
        nc = max(host._descriptor()[0].n_components, 1)
        dl = self._device_layout(dev)
        ps = self.per_step
        # accumulators (current stream)
        g_host = torch.zeros(nc, 2, dtype=torch.float64, device=dev)
        g_gm = torch.zeros(M, dtype=torch.float64, device=dev)
        g_table = torch.zeros(st["nseed"], 2, 2, T + 1, 3, dtype=torch.float64, device=dev)
        g_prog_rows = torch.zeros(T + 1, 1, 6, dtype=torch.float64, device=dev)
        F = torch.zeros(T + 1, M, 6, dtype=torch.float64, device=dev)
        side_host = torch.zeros_like(g_host)
        side_gm = torch.zeros_like(g_gm)
        gms, rel_all = dl["gms"], dl["idx"]
        z1 = torch.zeros(1, 6, dtype=torch.float64, device=dev)
        zM = torch.zeros(M, 6, dtype=torch.float64, device=dev)
        main = torch.cuda.current_stream(dev)
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(main)
        keep = []
        C = None
        for lo, hi_ in self._ranges():
            A = st["padj"].backward(lam_T, rows=(lo, hi_))
            g_host += A["g_host"]
            g_gm += A["g_gm"]
            done = torch.cuda.Event()
            done.record(main)
            keep.append(A)
            top = hi_ + 1 if hi_ == T else hi_
            with torch.cuda.stream(side):
                side.wait_event(done)
                # releases at rows [lo, hi): the particles in that slice (sorted by release row);
                # the adjoint table keeps positions and velocities apart, [seed][arm][x|v][row],
                # so each seed's four VJP inputs are contiguous slices
                a, b = lo * ps, min(hi_, T) * ps
                lam0 = A["lam0"][a:b]
                g_table.view(-1, 3).index_add_(0, dl["flat_x"][a:b], lam0[:, :3])
                g_table.view(-1, 3).index_add_(0, dl["flat_v"][a:b], lam0[:, 3:])
                nr = hi_ - lo
                gp_all = torch.empty(st["nseed"], nr, 6, dtype=torch.float64, device=dev)
                gh_all = torch.empty(st["nseed"], nc, 2, dtype=torch.float64, device=dev)
                scratch = torch.empty(nr, 2 * nc, dtype=torch.float64, device=dev)
                pw = st["prog_rows"][lo:hi_]
                for s in range(st["nseed"]):
                    _release_vjp_into(host, pw, rel_all[lo:hi_], gms[lo:hi_], self.seed + s,
                                      self.kval_arr, g_table[s, 0, 0, lo:hi_],
                                      g_table[s, 1, 0, lo:hi_], g_table[s, 0, 1, lo:hi_],
                                      g_table[s, 1, 1, lo:hi_], gp_all[s], gh_all[s], scratch)
                g_prog_rows[lo:hi_, 0] += gp_all.sum(0)
                side_host += gh_all.sum(0)
                keep.append((gp_all, gh_all, scratch))
                # progenitor orbit (in host + subhalos) over the same rows
                C = st["pa"].backward(z1, forcing=g_prog_rows, rows=(lo, hi_))
                side_host += C["g_host"]
                side_gm += C["g_gm"]
                # subhalo orbits (in the host), forced by the source-position adjoints
                F[lo:top, :, :3] = A["g_pos"][lo:top] + C["g_pos"][lo:top]
                D = st["sa"].backward(zM, forcing=F, rows=(lo, hi_))
                side_host += D["g_host"]
                keep.append((C, D))
        main.wait_stream(side)
        g_host += side_host
        g_gm += side_gm
        return (G * g_host[hi, 0], g_host[hi, 1], G * g_gm, C["lam0"].reshape(6))


def _with_sources(host, src):
    from .main import Potential

    class _Total(Potential):
        def __init__(self):
            super().__init__(None, {})
            self._G = G

        def _kernel_components(self):
            return host._kernel_components()

        def _kernel_sources(self):
            return src

    return _Total()


class _StreamLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, halo, sub_masses, prog_w0, model, target):
        st = model.forward(halo[0], halo[1], sub_masses, prog_w0)
        d = st["wT"][:, :3] - target
        loss = (d * d).sum()
        ctx.model, ctx.st, ctx.d = model, st, d
        return loss

    @staticmethod
    def backward(ctx, gout):
        lam_T = torch.zeros_like(ctx.st["wT"])
        lam_T[:, :3] = 2.0 * ctx.d
        gm_h, grs_h, g_m, g_w0 = ctx.model.backward(ctx.st, lam_T)
        go = gout.to(torch.float64)
        ghalo = torch.stack([gm_h, grs_h]) * go
        return ghalo, g_m * go, g_w0 * go, None, None


def stream_loss(halo, sub_masses, prog_w0, model, target):
    """L = sum_i |x_i(T) - target_i|^2 of the perturbed spray stream, differentiable with respect
    to halo = (m_halo, r_s_halo), the subhalo masses and the progenitor row (torch tensors on the
    GPU; use .backward() or torch.autograd.grad)."""
    return _StreamLoss.apply(halo, sub_masses, prog_w0, model, _lib.f64(target))
