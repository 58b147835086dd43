"""Learned clump-force surrogate.

A small MLP stands in for the direct sum over a clumpy halo's Plummer clumps: it maps scaled
(x, y, z) to the scaled (Phi, grad Phi) of the clump sum.  It is trained on the GPU by the
project's own BF16 MFMA kernels on positions sampled along plain-method (exact-sum) orbits,
with exact labels from the potential-evaluation kernel.  After training, a residual map over
a grid of cells records where the surrogate's force error exceeds ``threshold`` (relative to the
total host + clump acceleration) or where no validation sample fell; positions there, and
positions outside the training box, fall back to the exact sum during integration.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .mlp import MLP


@dataclass
class Scaling:
    center: tuple
    half: tuple
    phi_scale: float
    g_scale: float


class ClumpSurrogate:
    def __init__(self, host, clumps, mlp, scaling, flags, grid_lo, grid_cell, grid_n,
                 threshold=1e-3, report=None):
        self.host = host          # analytic host Potential (no sources)
        self.clumps = clumps      # Potential whose kernel sources are the static clumps
        self.mlp = mlp
        self.scaling = scaling
        self.flags = flags        # uint8 [nx, ny, nz]: 1 = use the exact clump sum
        self.grid_lo = grid_lo
        self.grid_cell = grid_cell
        self.grid_n = grid_n
        self.threshold = threshold
        self.report = report or {}

    # ------------------------------------------------------------------ evaluation
    def scale_inputs(self, xyz):
        s = self.scaling
        x = _lib.f64(xyz).reshape(-1, 3)
        c = torch.as_tensor(s.center, dtype=torch.float64, device=x.device)
        h = torch.as_tensor(s.half, dtype=torch.float64, device=x.device)
        return ((x - c) / h).to(torch.float32)

    def predict(self, xyz):
        """(phi, grad) of the clump sum predicted by the network (float64 tensors)."""
        y = self.mlp.infer(self.scale_inputs(xyz)).double()
        return y[:, 0] * self.scaling.phi_scale, y[:, 1:] * self.scaling.g_scale

    def exact(self, xyz):
        phi, g = self.clumps._eval(xyz, 0.0, want_phi=True, want_grad=True)
        return phi.reshape(-1), g.reshape(-1, 3)

    def cell_index(self, xyz):
        x = _lib.f64(xyz).reshape(-1, 3)
        lo = torch.as_tensor(self.grid_lo, dtype=torch.float64, device=x.device)
        cell = torch.as_tensor(self.grid_cell, dtype=torch.float64, device=x.device)
        n = torch.as_tensor(self.grid_n, device=x.device)
        ijk = torch.floor((x - lo) / cell).long()
        inside = ((ijk >= 0) & (ijk < n)).all(1)
        ijk = torch.minimum(torch.clamp(ijk, min=0), n - 1)
        flat = (ijk[:, 0] * self.grid_n[1] + ijk[:, 1]) * self.grid_n[2] + ijk[:, 2]
        return flat, inside

    def fallback_mask(self, xyz):
        flat, inside = self.cell_index(xyz)
        return (~inside) | (self.flags.reshape(-1)[flat] != 0)

    def scaling_vector(self):
        s = self.scaling
        return np.array(list(s.center) + list(s.half) + [s.phi_scale, s.g_scale] +
                        list(self.grid_lo) + list(self.grid_cell) + list(self.grid_n),
                        dtype=np.float64)

    def leapfrog(self, w0, t0, dt, nsteps, start=None, flags=None):
        """Leapfrog of rows w0 (n, 6) in host + clumps with the surrogate clump force and the
        exact-sum fallback (ssc_surrogate_leapfrog).  Returns (w_final, fallback particle-steps).
        ``flags`` overrides the residual map (uint8 of the same shape)."""
        h = self.host._host_only()
        src = self.clumps._kernel_sources()
        sd = src.desc()
        w0 = _lib.f64(w0).reshape(-1, 6).contiguous()
        n = w0.shape[0]
        out = torch.empty_like(w0)
        st = None if start is None else _lib.i32(start)
        fl = (self.flags if flags is None else flags).to(device=w0.device, dtype=torch.uint8).contiguous()
        cnt = torch.zeros(1, dtype=torch.int64, device=w0.device)
        sv = np.ascontiguousarray(self.scaling_vector())
        _lib.call("ssc_surrogate_leapfrog", _lib.ctypes.byref(h), _lib.ctypes.byref(sd),
                  sv.ctypes.data_as(_lib.ctypes.c_void_p), _lib.ptr(self.mlp.pack()),
                  _lib.ptr(fl), _lib.ptr(w0), _lib.ptr(st), n, float(t0), float(dt), int(nsteps),
                  _lib.ptr(out), _lib.ptr(cnt))
        return out, cnt

    def adjoint(self, w0, t0, dt, nsteps, start=None, checkpoint_every=100, flags=None):
        """Checkpointed discrete adjoint of ``leapfrog`` (SurrogateAdjoint)."""
        return SurrogateAdjoint(self, w0, t0, dt, nsteps, start, checkpoint_every, flags)

    # ------------------------------------------------------------------ training
    @classmethod
    def train(cls, host, clumps, positions, n_train=1 << 22, n_val=1 << 18, batch=8192,
              iters=96000, lr=2e-3, lr_min=1e-5, seed=0, threshold=1e-3, grid_cells=64,
              weight_cut=30.0, log=None):
        """Train on ``positions`` (sampled along plain-method orbits, (N, 3) on the GPU).
        Loss weights 1 / (1 + (|g| / g_ref)^2), g_ref = weight_cut x the median clump force:
        the spikes near clumps (which the residual map sends to the exact sum anyway) do not
        set the fit.  The defaults (96,000 iterations of 8,192 points, cosine learning rate
        2e-3 -> 1e-5, weight_cut 30) came from a sweep of the final stream-position error
        against the plain method on the clumpy-halo workload."""
        t_start = time.perf_counter()
        dev = _lib.device()
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        pos = _lib.f64(positions).reshape(-1, 3)
        perm = torch.randperm(pos.shape[0], generator=g, device=dev)
        n_train = min(n_train, pos.shape[0] - n_val)
        xtr, xva = pos[perm[:n_train]], pos[perm[n_train:n_train + n_val]]
        # exact labels (potential-evaluation kernel on the clump table)
        phi_tr, g_tr = _eval_chunked(clumps, xtr)
        lo = xtr.min(0).values
        hi = xtr.max(0).values
        center = 0.5 * (lo + hi)
        half = 0.5 * (hi - lo) * 1.02 + 1e-3
        phi_scale = float(phi_tr.std().item()) or 1.0
        g_scale = float(g_tr.norm(dim=1).pow(2).mean().sqrt().item()) or 1.0
        scaling = Scaling(tuple(center.tolist()), tuple(half.tolist()), phi_scale, g_scale)
        xs = ((xtr - center) / half).float()
        tgt = torch.cat([(phi_tr / phi_scale)[:, None], g_tr / g_scale], 1).float()
        # down-weight the near-field spikes so the smooth far field sets the fit
        gmag = g_tr.norm(dim=1)
        gref = weight_cut * float(gmag.median().item())
        wts = (1.0 / (1.0 + (gmag / gref) ** 2)).float()
        net = MLP(seed=seed)
        losses = []
        for it in range(iters):
            idx = torch.randint(0, n_train, (batch,), generator=g, device=dev)
            loss, _ = net.loss_backward(xs[idx], tgt[idx], wts[idx])
            frac = it / max(1, iters - 1)
            cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * frac))
            net.adam(cur)
            if it % 500 == 0 or it == iters - 1:
                losses.append((it, float(loss.item())))
                if log:
                    log(f"surrogate training iter {it}: weighted loss {losses[-1][1]:.3e}")
        t_train = time.perf_counter() - t_start
        obj = cls(host, clumps, net, scaling, None, None, None, None, threshold)
        obj._residual_map(xva, grid_cells, lo, hi)
        obj.report.update(train_seconds=t_train, losses=losses, n_train=n_train, n_val=n_val,
                          iters=iters, batch=batch)
        return obj

    def _residual_map(self, xva, cells, lo, hi):
        """Per-cell maximum of |a_surrogate - a_exact| / |a_total| over validation positions;
        cells above ``threshold`` or without samples fall back to the exact sum."""
        phi_e, g_e = _eval_chunked(self.clumps, xva)
        _, g_s = self.predict(xva)
        g_h = self.host._eval(xva, 0.0, want_grad=True).reshape(-1, 3)
        res = (g_s - g_e).norm(dim=1) / (g_h + g_e).norm(dim=1)
        n = int(cells)
        lo_ = lo - 1e-6
        cell = (hi - lo + 2e-6) / n
        self.grid_lo = tuple(lo_.tolist())
        self.grid_cell = tuple(cell.tolist())
        self.grid_n = (n, n, n)
        flat, _ = self.cell_index(xva)
        worst = torch.full((n ** 3,), -1.0, dtype=torch.float64, device=xva.device)
        self.cell_worst = worst.scatter_reduce(0, flat, res, reduce="amax")
        self._xva = xva
        self.report.update(val_residual_median=float(res.median().item()),
                           val_residual_p99=float(torch.quantile(res.float(), 0.99).item()),
                           val_residual_max=float(res.max().item()))
        self.set_threshold(self.threshold)

    def set_threshold(self, threshold):
        """Re-flag the cells for another residual threshold (cells without validation samples
        stay flagged)."""
        self.threshold = float(threshold)
        worst = self.cell_worst
        flags = ((worst < 0) | (worst > self.threshold)).to(torch.uint8)
        self.flags = flags.reshape(self.grid_n).contiguous()
        fb = self.fallback_mask(self._xva)
        self.report.update(flagged_cells=int(flags.sum().item()), cells=int(worst.numel()),
                           threshold=self.threshold,
                           val_fallback_fraction=float(fb.double().mean().item()))


class SurrogateAdjoint:
    """Forward (with checkpoints every ``checkpoint_every`` steps) and backward of the surrogate
    leapfrog for rows w0 (ssc_surrogate_adjoint_forward / _backward).  The backward returns
    dL/dw0 and dL/d(host parameters) for dL/dw_final = lam_T and optional per-row forcing."""

    def __init__(self, sur, w0, t0, dt, nsteps, start=None, checkpoint_every=100, flags=None):
        self.sur = sur
        self.h = sur.host._host_only()
        self.src = sur.clumps._kernel_sources()   # keeps the descriptor's tensors alive
        self.sd = self.src.desc()
        self.w0 = _lib.f64(w0).reshape(-1, 6).contiguous()
        self.n = self.w0.shape[0]
        self.start = None if start is None else _lib.i32(start)
        self.t0, self.dt, self.nsteps, self.E = float(t0), float(dt), int(nsteps), int(checkpoint_every)
        fl = sur.flags if flags is None else flags
        self.flags = fl.to(device=self.w0.device, dtype=torch.uint8).contiguous()
        self.sv = np.ascontiguousarray(sur.scaling_vector())
        nbytes = _lib.load().ssc_surrogate_adjoint_workspace_bytes(self.n, self.nsteps, self.E)
        self.ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=self.w0.device)
        self.w_out = None

    def _svp(self):
        return self.sv.ctypes.data_as(_lib.ctypes.c_void_p)

    def forward(self):
        self.w_out = torch.empty_like(self.w0)
        cnt = torch.zeros(1, dtype=torch.int64, device=self.w0.device)
        _lib.call("ssc_surrogate_adjoint_forward", _lib.ctypes.byref(self.h),
                  _lib.ctypes.byref(self.sd), self._svp(), _lib.ptr(self.sur.mlp.pack()),
                  _lib.ptr(self.flags), _lib.ptr(self.w0), _lib.ptr(self.start), self.n, self.t0,
                  self.dt, self.nsteps, self.E, _lib.ptr(self.w_out), _lib.ptr(cnt),
                  _lib.ptr(self.ws), self.ws.numel())
        self.fallback_forward = cnt
        return self.w_out

    def backward(self, lam_T, forcing=None):
        lam_T = _lib.f64(lam_T).reshape(self.n, 6).contiguous()
        F = None if forcing is None else _lib.f64(forcing).contiguous()
        lam0 = torch.empty_like(lam_T)
        nc = self.h.n_components
        gh = torch.zeros(max(nc, 1), 2, dtype=torch.float64, device=lam_T.device)
        cnt = torch.zeros(1, dtype=torch.int64, device=lam_T.device)
        mlp = self.sur.mlp
        _lib.call("ssc_surrogate_adjoint_backward", _lib.ctypes.byref(self.h),
                  _lib.ctypes.byref(self.sd), self._svp(), _lib.ptr(mlp.pack()),
                  _lib.ptr(mlp.pack_vjp()), _lib.ptr(self.flags), _lib.ptr(self.start), self.n,
                  self.dt, self.nsteps, self.E, _lib.ptr(lam_T), _lib.ptr(F), _lib.ptr(lam0),
                  _lib.ptr(gh), _lib.ptr(cnt), _lib.ptr(self.ws), self.ws.numel())
        return dict(lam0=lam0, g_host=gh[:nc], fallback=cnt)


def _eval_chunked(pot, xyz, chunk=1 << 20):
    phis, gs = [], []
    for i in range(0, xyz.shape[0], chunk):
        p, gg = pot._eval(xyz[i:i + chunk], 0.0, want_phi=True, want_grad=True)
        phis.append(p.reshape(-1))
        gs.append(gg.reshape(-1, 3))
    return torch.cat(phis), torch.cat(gs)


def orbit_positions(pot, w0, start, t0, dt, nsteps, stride=1):
    """Positions along plain-method orbits: integrate rows w0 from their start steps and keep
    every ``stride``-th grid row at or after each start.  Returns (N, 3) float64."""
    kidx = np.arange(0, nsteps + 1, stride, dtype=np.int32)
    _, ys = pot._leapfrog(w0, t0, dt, nsteps, start=start, save_idx=kidx)
    st = torch.as_tensor(np.asarray(start), device=ys.device)
    keep = torch.as_tensor(kidx, device=ys.device)[:, None] >= st[None, :]
    return ys[..., :3][keep]
