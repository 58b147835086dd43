"""Spray-stream driver shared by ``Potential.gen_stream_vmapped`` (reference
streamsculptor/main.py:343-368) and the perturbed-stream helpers (reference
streamsculptor/streamhelpers.py:57-146).

Three potentials take part, as in the reference: the progenitor orbit is integrated in
``prog_pot``, release conditions come from ``release_pot`` (the smooth base potential in the
perturbed variants, streamhelpers.py:58-66) and particles are integrated in ``particle_pot``.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from . import solvers as _solvers
from ._grid import release_grid


def stream_ics(prog_pot, release_pot, ts, prog_w0, Msat, seed_num, solver=None, kval_arr=1.0,
               dt0=None, **adaptive_kw):
    """(pos_close, pos_far, vel_close, vel_far), each (len(ts), 3), plus the grid used."""
    method = _solvers.resolve(solver)
    ts_np = np.asarray(ts.detach().cpu().numpy() if isinstance(ts, torch.Tensor) else ts,
                       dtype=np.float64).reshape(-1)
    if method == "leapfrog":
        t0, dt, nsteps, kidx = release_grid(ts_np, dt0)
        _, ys = prog_pot._leapfrog(_lib.f64(prog_w0).reshape(1, 6), t0, dt, nsteps, save_idx=kidx)
        ws = ys[:, 0, :].contiguous()
        grid = (t0, dt, nsteps, kidx)
    else:
        ws = prog_pot.integrate_orbit(w0=prog_w0, ts=ts_np, solver=solver, dt0=dt0,
                                      **adaptive_kw).ys
        grid = None
    ics = release_pot._release_along(ws, ts_np, Msat, seed_num, kval_arr)
    return ics, grid, ts_np


def _static_in_time(pot):
    """True when the potential does not depend on time (no LINE or TRACK sources): a fixed-step
    orbit in it can be restarted at any grid row with the same arithmetic."""
    _, src = pot._descriptor()
    return src is None or src.kind == _lib.SRC_STATIC


def _restartable(pot):
    """True when a fixed-step orbit in the potential can restart at a grid row with the sources
    it would meet in one pass: no sources, STATIC sources, or TRACK sources given as a table on
    the grid (a window reads the table's own rows).  Straight-line sources and tracks sampled
    from a function form their positions from the window's times."""
    from .main import SampledTrack
    _, src = pot._descriptor()
    return (src is None or src.kind == _lib.SRC_STATIC or
            (src.kind in _lib.TRACK_KINDS and not isinstance(src, SampledTrack)))


_SIDE = {}


def _side_stream(dev):
    """One second stream per device, kept for the process: a stream of its own keeps its
    hardware queue, and creating one per call costs host time ahead of the first launch."""
    key = torch.device(dev).index
    if key not in _SIDE:
        _SIDE[key] = torch.cuda.Stream(device=dev)
    return _SIDE[key]


def _leapfrog_stream_pipelined(prog_pot, release_pot, particle_pot, ts_np, prog_w0, Msat,
                               seed_num, kval_arr, dt0, chunks):
    """The fixed-step spray stream in ``chunks`` windows of grid rows on two streams: the
    current stream advances the progenitor orbit through a window and releases the particles
    of that window; a second stream then advances every particle released so far through the
    same window while the progenitor moves on through the next.  The single-orbit chain (one
    lane, latency-bound) thus overlaps the particles' integration instead of preceding all of
    it, and two streams keep clear of hardware-queue sharing, which serialises streams mapped
    to one queue.  Every orbit takes the steps of the one-pass path (a restart at a grid row
    recomputes the acceleration there from the same state; a particle starts at its own row):
    in the host alone the rows are the same bit for bit; with sources the acceleration at a
    window's first row comes from the kernel's prologue, whose FMA contraction can differ from
    the step loop's in the last bit.  Returns the particle rows (2T, 6), lead/trail
    interleaved, in release order."""
    from .main import _as_float
    t0, dt, nsteps, kidx = release_grid(ts_np, dt0)
    T = len(ts_np) - 1
    dev = _lib.device()
    w = _lib.f64(prog_w0).reshape(1, 6)
    # windows of release rows [cuts[c], cuts[c + 1]): grid rows [kidx[cuts[c]], kidx[cuts[c + 1]]]
    cuts = sorted({0, T} | {T * j // chunks for j in range(1, chunks)})
    msat = np.broadcast_to(np.asarray(_as_float(Msat), dtype=np.float64), (T + 1,))
    gms_all = release_pot._G * _lib.f64(np.ascontiguousarray(msat[:T]))
    # every window's index arrays in one host-to-device copy, up front
    rows_all = np.repeat(kidx[:T], 2).astype(np.int64)
    parts, spans = [], []
    for c in range(len(cuts) - 1):
        a, b = cuts[c], cuts[c + 1]
        g0 = int(kidx[a])
        for arr in ((kidx[a:b] - g0), np.maximum(rows_all[:2 * b] - g0, 0)):
            spans.append((sum(p.size for p in parts), arr.size))
            parts.append(arr.astype(np.int32))
    flat = _lib.i32(np.concatenate(parts))
    save_all = [flat[o:o + k] for o, k in spans[0::2]]
    start_all = [flat[o:o + k] for o, k in spans[1::2]]
    idx_all = torch.arange(T, dtype=torch.int32, device=dev)
    main = torch.cuda.current_stream(dev)
    side = _side_stream(dev)
    state = None
    keep = []
    for c in range(len(cuts) - 1):
        a, b = cuts[c], cuts[c + 1]
        g0 = int(kidx[a])
        g1 = int(kidx[b]) if b < T else nsteps
        # progenitor rows kidx[a..b-1], and its state at the next window's first row g1
        pn = (g1 if b < T else int(kidx[T - 1])) - g0
        w_next, ys = prog_pot._leapfrog(w, t0 + g0 * dt, dt, pn, save_idx=save_all[c])
        ws = ys[:, 0, :].contiguous()
        pc, pf, vc, vf = release_pot._release(ws, idx_all[a:b].contiguous(),
                                              gms_all[a:b].contiguous(), seed_num, kval_arr,
                                              t_rel=ts_np[a:b])
        new = torch.empty(2 * (b - a), 6, dtype=torch.float64, device=dev)
        new[0::2, :3], new[0::2, 3:] = pc, vc
        new[1::2, :3], new[1::2, 3:] = pf, vf
        side.wait_stream(main)
        with torch.cuda.stream(side):
            w_in = new if state is None else torch.cat([state, new], 0)
            state, _ = particle_pot._leapfrog(w_in, t0 + g0 * dt, dt, g1 - g0,
                                              start=start_all[c])
        keep.append((new, w_in))
        w = w_next
    main.wait_stream(side)
    return state


def gen_stream(prog_pot, release_pot, ts, prog_w0, Msat, seed_num, solver=None, kval_arr=1.0,
               rtol=1e-7, atol=1e-7, dtmin=0.3, dtmax=None, max_steps=10_000, throw=False,
               dt0=None, particle_pot=None, chunks=8):
    """Returns (lead (T, 6), trail (T, 6)) with T = len(ts) - 1.  With the fixed-step solver in
    time-independent potentials the progenitor orbit and the particles are pipelined over
    ``chunks`` windows of grid rows (the one-pass rows, to the last bit with sources; chunks=1:
    one pass after the other)."""
    particle_pot = particle_pot or prog_pot
    method = _solvers.resolve(solver)
    if method == "leapfrog" and chunks > 1:
        ts_np = np.asarray(ts.detach().cpu().numpy() if isinstance(ts, torch.Tensor) else ts,
                           dtype=np.float64).reshape(-1)
        T = len(ts_np) - 1
        if (T >= 4 * chunks and _static_in_time(prog_pot) and _static_in_time(release_pot)
                and _static_in_time(particle_pot)):
            out = _leapfrog_stream_pipelined(prog_pot, release_pot, particle_pot, ts_np,
                                             prog_w0, Msat, seed_num, kval_arr, dt0, chunks)
            return out[0::2], out[1::2]
    adaptive_kw = dict(rtol=rtol, atol=atol, dtmin=dtmin, dtmax=dtmax, max_steps=max_steps)
    (pc, pf, vc, vf), grid, ts_np = stream_ics(prog_pot, release_pot, ts, prog_w0, Msat, seed_num,
                                               solver=solver, kval_arr=kval_arr, dt0=dt0,
                                               **adaptive_kw)
    T = len(ts_np) - 1
    w0 = torch.empty(2 * T, 6, dtype=torch.float64, device=pc.device)
    w0[0::2, :3], w0[0::2, 3:] = pc[:T], vc[:T]
    w0[1::2, :3], w0[1::2, 3:] = pf[:T], vf[:T]
    if method == "leapfrog":
        t0, dt, nsteps, kidx = grid
        start = np.repeat(kidx[:T], 2).astype(np.int32)
        out, _ = particle_pot._leapfrog(w0, t0, dt, nsteps, start=start)
    else:
        from . import adaptive
        t_start = np.repeat(ts_np[:T], 2)
        out = adaptive.integrate_to(particle_pot, w0, t_start, float(ts_np[-1]), throw=throw,
                                    **adaptive_kw)
    return out[0::2], out[1::2]


def gen_stream_dense(prog_pot, release_pot, ts, prog_w0, Msat, seed_num, solver=None,
                     kval_arr=1.0, rtol=1e-7, atol=1e-7, dtmin=0.3, dtmax=None, max_steps=10_000):
    """Dense spray stream: a batched DenseSolution over 2 (len(ts) - 1) particles, particle
    2i + arm released at ts[i] (arm 0 = lead / close, 1 = trail / far) and integrated to ts[-1]
    with Dopri5 (reference gen_stream_vmapped_dense, main.py:409-430)."""
    if _solvers.resolve(solver) == "leapfrog":
        raise NotImplementedError("dense streams need an adaptive solver (Dopri5)")
    from . import adaptive
    adaptive_kw = dict(rtol=rtol, atol=atol, dtmin=dtmin, dtmax=dtmax, max_steps=max_steps)
    (pc, pf, vc, vf), _, ts_np = stream_ics(prog_pot, release_pot, ts, prog_w0, Msat, seed_num,
                                            solver=solver, kval_arr=kval_arr, **adaptive_kw)
    T = len(ts_np) - 1
    w0 = torch.empty(2 * T, 6, dtype=torch.float64, device=pc.device)
    w0[0::2, :3], w0[0::2, 3:] = pc[:T], vc[:T]
    w0[1::2, :3], w0[1::2, 3:] = pf[:T], vf[:T]
    t_start = np.repeat(ts_np[:T], 2)
    h, src = adaptive._sources(prog_pot)
    sd = src.desc() if src is not None else None
    return adaptive.run_dense(h, sd, w0, float(ts_np[-1]), t0=float(ts_np[0]), t_start=t_start,
                              batched=True, **adaptive_kw)


def spray_stream(prog_pot, release_pot, ts, prog_w0, Msat, seed_num, per_step, dt0,
                 kval_arr=1.0, particle_pot=None, integrate=None, chunks=8):
    """Final rows of the many-particle spray stream -- the particles of ``spray_particles``, in
    its order, advanced with the fixed-step leapfrog to the last grid row -- with the progenitor
    orbit pipelined with the particles over ``chunks`` windows of grid rows on two streams, as
    ``gen_stream`` does for two particles per step: the current stream advances the progenitor
    through a window and releases that window's particles (every seed in one launch); the
    second stream then advances every particle released so far through the window while the
    progenitor moves on.  ``integrate(w, t0, dt, nsteps, start)`` advances rows (default: the
    leapfrog in ``particle_pot``, else ``prog_pot``); TRACK tables are cut to each window's
    rows.  The rows equal the one-pass order (``spray_particles``, then one integration) up to
    the last bits of the accelerations re-formed at the windows' first rows.  Fewer than
    4 * chunks release steps, or sources that form their positions from the time (straight
    lines, sampled tracks): one pass."""
    from .main import _as_float
    from .workloads import spray_layout
    particle_pot = particle_pot or prog_pot
    if integrate is None:
        def integrate(w, t_start, step, n, start):
            return particle_pot._leapfrog(w, t_start, step, n, start=start)[0]
    ts_np = np.asarray(ts, dtype=np.float64).reshape(-1)
    T = len(ts_np) - 1
    if (chunks <= 1 or T < 4 * chunks or not _restartable(prog_pot)
            or not _restartable(particle_pot)):
        w0, start, _, (t0, dt, nsteps, _) = spray_particles(
            prog_pot, release_pot, ts_np, prog_w0, Msat, seed_num, per_step, dt0, kval_arr)
        return integrate(w0, t0, dt, nsteps, start)
    t0, dt, nsteps, kidx = release_grid(ts_np, dt0)
    dev = _lib.device()
    nseed = (per_step + 1) // 2
    lay = np.array(spray_layout(per_step), dtype=np.int64)  # (parity, per_step, [seed, arm])
    msat = np.broadcast_to(np.asarray(_as_float(Msat), dtype=np.float64), (T + 1,))
    gms_all = release_pot._G * _lib.f64(np.ascontiguousarray(msat[:T]))
    idx_all = torch.arange(T, dtype=torch.int32, device=dev)
    cuts = sorted({0, T} | {T * j // chunks for j in range(1, chunks)})
    # every window's index arrays in one host-to-device copy: the progenitor rows it saves,
    # (seed, arm, release) of its new particles, and the start rows of all particles so far
    parts, spans = [], []
    for c in range(len(cuts) - 1):
        a, b = cuts[c], cuts[c + 1]
        g0 = int(kidx[a])
        sel = lay[np.arange(a, b) % 2]                            # (b - a, per_step, 2)
        rel_all = np.repeat(np.arange(b), per_step)
        for arr in (kidx[a:b] - g0, sel[..., 0].reshape(-1), sel[..., 1].reshape(-1),
                    np.repeat(np.arange(b - a), per_step), np.maximum(kidx[rel_all] - g0, 0)):
            spans.append((sum(p.size for p in parts), arr.size))
            parts.append(np.asarray(arr).astype(np.int32))
    flat = _lib.i32(np.concatenate(parts))
    views = [flat[o:o + k] for o, k in spans]
    main = torch.cuda.current_stream(dev)
    side = _side_stream(dev)
    w = _lib.f64(prog_w0).reshape(1, 6)
    state, keep = None, []
    for c in range(len(cuts) - 1):
        a, b = cuts[c], cuts[c + 1]
        g0 = int(kidx[a])
        g1 = int(kidx[b]) if b < T else nsteps
        save, s_i, a_i, r_i, start = views[5 * c:5 * c + 5]
        pn = (g1 if b < T else int(kidx[T - 1])) - g0
        w_next, ys = prog_pot._leapfrog(w, t0 + g0 * dt, dt, pn, save_idx=save)
        table = release_pot._release_seeds(ys[:, 0, :].contiguous(), idx_all[a:b].contiguous(),
                                           gms_all[a:b].contiguous(), seed_num, nseed,
                                           kval_arr, t_rel=ts_np[a:b])
        new = table[s_i.long(), a_i.long(), r_i.long()]
        side.wait_stream(main)
        with torch.cuda.stream(side):
            w_in = new if state is None else torch.cat([state, new], 0)
            state = integrate(w_in, t0 + g0 * dt, dt, g1 - g0, start)
        keep.append((new, w_in))
        w = w_next
    main.wait_stream(side)
    return state


def spray_particles(prog_pot, release_pot, ts, prog_w0, Msat, seed_num, per_step, dt0,
                    kval_arr=1.0):
    """Initial rows of a many-particle spray stream with ``per_step`` particles released at
    every ts[i], i < len(ts)-1: ceil(per_step/2) release-model draws (seeds seed_num,
    seed_num+1, ...) per step, lead/trail alternating (workloads.spray_layout).  Returns
    (w0 (N, 6), start_step (N,), arm (N,), grid) ordered by release step."""
    from .workloads import spray_layout
    ts_np = np.asarray(ts, dtype=np.float64).reshape(-1)
    t0, dt, nsteps, kidx = release_grid(ts_np, dt0)
    _, ys = prog_pot._leapfrog(_lib.f64(prog_w0).reshape(1, 6), t0, dt, nsteps, save_idx=kidx)
    ws = ys[:, 0, :].contiguous()
    T = len(ts_np) - 1
    even, odd = spray_layout(per_step)
    nseed = (per_step + 1) // 2
    dev = ws.device
    from .main import _as_float
    # table[s, arm, i, 6] of every seed in one launch
    msat = np.broadcast_to(np.asarray(_as_float(Msat), dtype=np.float64), (T + 1,))
    gms = release_pot._G * _lib.f64(np.ascontiguousarray(msat))
    idx = torch.arange(T + 1, dtype=torch.int32, device=dev)
    table = release_pot._release_seeds(ws, idx, gms, seed_num, nseed, kval_arr, t_rel=ts_np)
    lay = np.array([even, odd], dtype=np.int64)            # (parity, per_step, [seed, arm])
    par = np.arange(T) % 2
    sel_s, sel_a = lay[par, :, 0], lay[par, :, 1]
    rel = np.repeat(np.arange(T), per_step)
    sar = torch.as_tensor(np.stack([sel_s.reshape(-1), sel_a.reshape(-1), rel]), device=dev)
    w0 = table[sar[0], sar[1], sar[2]].contiguous()
    start = kidx[rel].astype(np.int32)
    return w0, start, sel_a.reshape(-1), (t0, dt, nsteps, kidx)
