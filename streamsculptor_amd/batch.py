"""Batches of independent stream simulations in single kernel launches.

The reference runs many perturbed-stream simulations at once by ``jax.vmap`` over the
simulation parameters (examples/GD1_and_SubhaloFitting_GPU.ipynb, cells 18-27: 700 streams,
each with its own straight-line subhalo).  Here a batch of B simulations shares one host
potential and differs in its subhalos: the subhalos of all simulations form one grouped source
table (group b = simulation b) and every orbit carries its simulation index, so the progenitor
orbits, the release model and all particle orbits of the batch each take one launch.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import adaptive
from .main import SourceSet, _as_float
from .potential import SubhaloLinePotential


def _arr(x, dtype=np.float64):
    return np.asarray(_as_float(x), dtype=dtype)


class SubhaloLineBatch:
    """B simulations with nsh straight-line subhalos each (one SubhaloLinePotential per
    simulation, reference streamsculptor/potential.py:802-830; profile="hernquist" gives
    Hernquist subhalos, as SubhaloLinePotentialCustom_fromFunc with HernquistPotential,
    :1161-1213).  Arrays: m, a, subhalo_t0, t_window (B, nsh); subhalo_x0, subhalo_v
    (B, nsh, 3)."""

    def __init__(self, m, a, subhalo_x0, subhalo_v, subhalo_t0, t_window, units=None,
                 profile="plummer"):
        if profile not in ("plummer", "hernquist"):
            raise NotImplementedError("subhalo profiles with a kernel form: plummer, hernquist")
        self.profile = profile
        m = _arr(m)
        if m.ndim != 2:
            raise ValueError("m must have shape (B, nsh)")
        self.B, self.nsh = m.shape
        shp = (self.B, self.nsh)
        self.m = m
        self.a = np.broadcast_to(_arr(a), shp).copy()
        self.subhalo_x0 = _arr(subhalo_x0).reshape(*shp, 3)
        self.subhalo_v = _arr(subhalo_v).reshape(*shp, 3)
        self.subhalo_t0 = np.broadcast_to(_arr(subhalo_t0), shp).copy()
        self.t_window = np.broadcast_to(_arr(t_window), shp).copy()
        from .main import usys
        self.units = units or usys
        self._G = self.units.G if getattr(self.units, "G", None) is not None else _lib.G_USYS
        self._src = None

    def __len__(self):
        return self.B

    def __getitem__(self, b):
        """The subhalo potential of simulation b."""
        if self.profile == "hernquist":
            from .potential import HernquistPotential, SubhaloLinePotentialCustom_fromFunc
            return SubhaloLinePotentialCustom_fromFunc(
                HernquistPotential, self.m[b], self.a[b], self.subhalo_x0[b], self.subhalo_v[b],
                self.subhalo_t0[b], self.t_window[b], units=self.units)
        return SubhaloLinePotential(self.m[b], self.a[b], self.subhalo_x0[b], self.subhalo_v[b],
                                    self.subhalo_t0[b], self.t_window[b], units=self.units)

    def sources(self):
        if self._src is None:
            kind = _lib.SRC_LINE_HERNQUIST if self.profile == "hernquist" else _lib.SRC_LINE
            self._src = SourceSet(kind, gm=self._G * self.m.reshape(-1),
                                  rs=self.a.reshape(-1), pos=self.subhalo_x0.reshape(-1, 3),
                                  vel=self.subhalo_v.reshape(-1, 3),
                                  t0=self.subhalo_t0.reshape(-1),
                                  t_window=self.t_window.reshape(-1))
        return self._src

    def desc(self, group):
        """Grouped sources descriptor; ``group`` (device int32) holds each orbit's simulation."""
        sd = self.sources().desc_grouped(self.nsh, group)
        return sd


def _host(pot_base):
    h, src = pot_base._descriptor()
    if src is not None:
        raise NotImplementedError("the base potential of a batch must be analytic")
    return h


def integrate_orbit_batch(pot_base, batch, w0, ts, t0=None, t1=None, rtol=1e-7, atol=1e-7,
                          dtmin=0.3, dtmax=None, max_steps=10_000, throw=False):
    """Orbit b (w0[b] or a shared w0) integrated in pot_base + batch[b] with Dopri5 and saved at
    ts: (B, len(ts), 6) (reference integrate_orbit per simulation, main.py:125-163)."""
    h = _host(pot_base)
    tsn = _arr(ts).reshape(-1)
    t0v = float(tsn.min()) if t0 is None else float(t0)
    t1v = float(tsn.max()) if t1 is None else float(t1)
    direction = 1.0 if t1v >= t0v else -1.0
    order = np.argsort(direction * tsn, kind="stable")
    w = _lib.f64(w0).reshape(-1, 6)
    if w.shape[0] == 1:
        w = w.expand(batch.B, 6).contiguous()
    group = torch.arange(batch.B, dtype=torch.int32, device=w.device)
    sd = batch.desc(group)
    _, ys, _ = adaptive.run(h, sd, w, t1v, t0=t0v, ts=tsn[order], rtol=rtol, atol=atol,
                            dtmin=dtmin, dtmax=dtmax, max_steps=max_steps, throw=throw)
    if not np.array_equal(order, np.arange(order.size)):
        ys = ys[:, torch.as_tensor(np.argsort(order), device=ys.device)]
    return ys


def gen_stream_batch_with_pert(pot_base, batch, ts, prog_w0, Msat, seed_num, kval_arr=1.0,
                               rtol=1e-7, atol=1e-7, dtmin=0.1, dtmax=None, max_steps=10_000,
                               throw=False, fixed_prog=False, events=None, return_stats=False):
    """B perturbed streams (reference streamhelpers.gen_stream_vmapped_with_pert, and
    gen_stream_vmapped_with_pert_fixed_prog when fixed_prog=True, streamhelpers.py:90-146) with
    Dopri5: progenitor b in pot_base + batch[b] (pot_base alone when fixed_prog), release model
    in pot_base, particles in pot_base + batch[b].  prog_w0: (6,) or (B, 6) at ts[0].
    Returns (lead (B, T, 6), trail (B, T, 6)), T = len(ts) - 1, and with return_stats the
    particles' (B, T, 2, 3) {accepted, rejected, status}.  ``events``: a pair of
    torch.cuda.Event recorded around the particle launch."""
    h = _host(pot_base)
    tsn = _arr(ts).reshape(-1)
    T = tsn.size - 1
    B = batch.B
    kw = dict(rtol=rtol, atol=atol, dtmin=dtmin, dtmax=dtmax, max_steps=max_steps, throw=throw)
    w = _lib.f64(prog_w0).reshape(-1, 6)
    dev = w.device
    if fixed_prog:
        if w.shape[0] != 1:
            raise ValueError("fixed_prog=True takes one progenitor point")
        _, ys, _ = adaptive.run(h, None, w, float(tsn.max()), t0=float(tsn.min()), ts=tsn, **kw)
        prog = ys.expand(B, T + 1, 6)
    else:
        prog = integrate_orbit_batch(pot_base, batch, w, tsn, **kw)
    rows = prog.reshape(B * (T + 1), 6).contiguous()
    rel_idx = torch.arange(T + 1, dtype=torch.int32, device=dev).repeat(B)
    gms = pot_base._G * _lib.f64(np.broadcast_to(_arr(Msat), (T + 1,)).copy()).repeat(B)
    pc, pf, vc, vf = pot_base._release(rows, rel_idx, gms.contiguous(), seed_num, kval_arr)
    pc, pf, vc, vf = (a.reshape(B, T + 1, 3)[:, :T] for a in (pc, pf, vc, vf))
    w0 = torch.empty(B, T, 2, 6, dtype=torch.float64, device=dev)
    w0[:, :, 0, :3], w0[:, :, 0, 3:] = pc, vc
    w0[:, :, 1, :3], w0[:, :, 1, 3:] = pf, vf
    t_start = torch.as_tensor(np.repeat(tsn[:T], 2), device=dev).repeat(B)
    group = torch.arange(B, dtype=torch.int32, device=dev).repeat_interleave(2 * T)
    sd = batch.desc(group)
    if events:
        events[0].record()
    out, _, stats = adaptive.run(h, sd, w0.reshape(-1, 6), float(tsn[-1]), t_start=t_start, **kw)
    if events:
        events[1].record()
    out = out.reshape(B, T, 2, 6)
    if return_stats:
        return out[:, :, 0], out[:, :, 1], stats.reshape(B, T, 2, 3)
    return out[:, :, 0], out[:, :, 1]
