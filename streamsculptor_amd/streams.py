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


def _leapfrog_stream_pipelined(prog_pot, release_pot, particle_pot, ts_np, prog_w0, Msat,
                               seed_num, kval_arr, dt0, chunks):
    """The fixed-step spray stream in ``chunks`` ranges of release rows: the progenitor orbit
    advances one range at a time on the current stream, each range's releases follow it, and
    the particles released there start integrating on a stream of their own while the
    progenitor advances through the next range -- the single-orbit chain (one lane, latency
    bound) overlaps the particles' integration instead of preceding all of it.  Every orbit
    sees the arithmetic of the unchunked path (a restart at a grid row recomputes the
    acceleration there from the same state, and each particle starts at its own row), so the
    rows are the same bit for bit.  Returns the particle rows (2T, 6), lead/trail interleaved."""
    t0, dt, nsteps, kidx = release_grid(ts_np, dt0)
    T = len(ts_np) - 1
    dev = _lib.device()
    w = _lib.f64(prog_w0).reshape(1, 6)
    cuts = sorted({0, T} | {T * j // chunks for j in range(1, chunks)})
    from .main import _as_float
    msat = np.broadcast_to(np.asarray(_as_float(Msat), dtype=np.float64), (T + 1,))
    gms_all = release_pot._G * _lib.f64(np.ascontiguousarray(msat[:T]))
    # index arrays on the device up front: no host-to-device copy inside the pipeline
    start_all = _lib.i32(np.repeat(kidx[:T], 2).astype(np.int32))
    save_all = [_lib.i32((kidx[cuts[c]:cuts[c + 1]] - kidx[cuts[c]]).astype(np.int32))
                for c in range(len(cuts) - 1)]
    main = torch.cuda.current_stream(dev)
    sides = [torch.cuda.Stream(device=dev) for _ in range(min(3, len(cuts) - 2))]
    out = torch.empty(2 * T, 6, dtype=torch.float64, device=dev)
    keep = []
    for c in range(len(cuts) - 1):
        a, b = cuts[c], cuts[c + 1]
        # progenitor rows kidx[a..b-1], and its state at the next range's first row
        r0 = int(kidx[a])
        r1 = int(kidx[b]) if b < T else int(kidx[T - 1])
        w_next, ys = prog_pot._leapfrog(w, t0 + r0 * dt, dt, r1 - r0, save_idx=save_all[c])
        ws = ys[:, 0, :].contiguous()
        idx = torch.arange(a, b, dtype=torch.int32, device=dev)
        pc, pf, vc, vf = release_pot._release(ws, idx, gms_all[a:b].contiguous(), seed_num,
                                              kval_arr, t_rel=ts_np[a:b])
        w0 = torch.empty(2 * (b - a), 6, dtype=torch.float64, device=dev)
        w0[0::2, :3], w0[0::2, 3:] = pc, vc
        w0[1::2, :3], w0[1::2, 3:] = pf, vf
        start = start_all[2 * a:2 * b]
        last = c == len(cuts) - 2
        st = main if (last or not sides) else sides[c % len(sides)]
        if st is not main:
            st.wait_stream(main)
        with torch.cuda.stream(st):
            o, _ = particle_pot._leapfrog(w0, t0, dt, nsteps, start=start)
            out[2 * a:2 * b] = o
        keep.append((w0, o))
        w = w_next
    for s in sides:
        main.wait_stream(s)
    return out


def gen_stream(prog_pot, release_pot, ts, prog_w0, Msat, seed_num, solver=None, kval_arr=1.0,
               rtol=1e-7, atol=1e-7, dtmin=0.3, dtmax=None, max_steps=10_000, throw=False,
               dt0=None, particle_pot=None, chunks=4):
    """Returns (lead (T, 6), trail (T, 6)) with T = len(ts) - 1.  With the fixed-step solver in
    time-independent potentials the progenitor orbit and the particles are pipelined over
    ``chunks`` ranges of release rows (same rows bit for bit; chunks=1: one pass after the
    other)."""
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
    # table[s, arm, i, 6]
    table = torch.empty(nseed, 2, T + 1, 6, dtype=torch.float64, device=dev)
    for s in range(nseed):
        pc, pf, vc, vf = release_pot._release_along(ws, ts_np, Msat, seed_num + s, kval_arr)
        table[s, 0, :, :3], table[s, 0, :, 3:] = pc, vc
        table[s, 1, :, :3], table[s, 1, :, 3:] = pf, vf
    lay = np.array([even, odd], dtype=np.int64)            # (parity, per_step, [seed, arm])
    par = np.arange(T) % 2
    sel_s, sel_a = lay[par, :, 0], lay[par, :, 1]
    rel = np.repeat(np.arange(T), per_step)
    s_idx = torch.as_tensor(sel_s.reshape(-1), device=dev)
    a_idx = torch.as_tensor(sel_a.reshape(-1), device=dev)
    r_idx = torch.as_tensor(rel, device=dev)
    w0 = table[s_idx, a_idx, r_idx].contiguous()
    start = kidx[rel].astype(np.int32)
    return w0, start, sel_a.reshape(-1), (t0, dt, nsteps, kidx)
