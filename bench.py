#!/usr/bin/env python3
"""Throughput of the stream-particle integration path on one MI355X.

Metric: particle-steps/s (one spray particle advanced one 1-Myr leapfrog step).  Default
workload (clumpy halo): the 3-component host plus 512 static Plummer clumps; 1e6 spray
particles (250 released per step) integrated to the present over 4000 x 1 Myr, i.e.
2,000,500,000 particle-steps per bench step.  A step is one full stream -- progenitor orbit,
release conditions, particle integration -- with inputs resident in HBM.  The clump force is
computed two ways: the plain FP64 direct sum and the learned BF16-MFMA surrogate (trained on
the GPU before the timed region); ``value`` is the rate of the faster form that meets the
10 pc bar on the median final stream-position error against the plain run (the exact plain
form always does).

Other workloads: --workload plain (8000 particles, host only), subhalos (1e5 particles, 64
live subhalos, direct sums), responses (1e5 particles + 64 first-order response vectors each),
adjoint (the subhalos stream's loss and its gradient with respect to the halo mass and radius,
the 64 subhalo masses and the progenitor row: forward with checkpoints every 100 steps plus
the full reverse chain), gd1 (a batch of 700 perturbed GD-1-like streams with adaptive Dopri5,
one straight-line subhalo each: the workload of the reference's GPU example, reported as
simulations/s and particle-steps/s), ensemble (the clumps potential and surrogate with 2000
progenitors of 2-6 Gyr streams, 4e6 particles: loss + gradient with respect to the halo and
every progenitor's initial row through the surrogate adjoint, checkpoints every 100 steps),
derivs (one batch of the reference's generate_derivs loop with an adaptive solver: 10,000
spray particles of a 4 Gyr stream and first-order mass and radius responses to 128
straight-line Hernquist subhalos, rtol = atol = 1e-11, the progenitor's backward field solve
and every particle's Dopri5 response solve under one controller per particle).

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12      # MI355X FP64 vector, FLOP/s
PEAK_BF16 = 2.5e15       # MI355X dense BF16 MFMA
ERR_LIMIT_PC = 10.0
# HBM bytes per launch of the plain clumps leapfrog kernel at the default 1e6 particles, from
# rocprofv3 FETCH_SIZE (x 2.0, calibrated on a 1 GiB copy) + WRITE_SIZE
TRAFFIC_CLUMPS_PLAIN = 100.4e6
CPU_SECONDS = 10.0       # target CPU time of the baseline sample


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name, particles=None):
    import torch
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.potential import (LiveSubhaloPotential, PlummerClumps,
                                              Potential_Combine)
    from streamsculptor_amd.main import usys
    host = W.host_potential()
    ts = W.release_times()
    extra = {}
    if name == "clumps":
        pos, m, rs = W.clump_table()
        clumps = PlummerClumps(m, rs, pos, units=usys)
        pot = Potential_Combine([host, clumps], units=usys)
        total, nsrc = 1_000_000, 512
        extra["clumps"] = clumps
    elif name in ("subhalos", "responses", "adjoint"):
        sw0, m, rs = W.subhalo_ics()
        live = LiveSubhaloPotential.from_orbits(host, sw0, m, rs, float(ts[0]), W.DT, W.T_STEPS)
        pot = Potential_Combine([host, live], units=usys) if name == "subhalos" else host
        extra["live"] = live
        total, nsrc = 100_000, 64
    else:
        pot, total, nsrc = host, 8000, 0
    if particles:
        total = particles
    per_step = W.particles_per_step(total)
    back = pot.integrate_orbit(w0=W.PROG_TODAY, ts=[0.0, ts[0]], t0=0.0, t1=float(ts[0]),
                               solver="leapfrog", dt0=W.DT).ys[-1]
    torch.cuda.synchronize()
    return dict(name=name, pot=pot, host=host, ts=ts, prog_w0=back, per_step=per_step,
                total=total, nsrc=nsrc, psteps=W.particle_steps(per_step), **extra)


def spray(wl):
    """Spray rows of the workload.  The perturbed streams (live subhalos, responses) release in
    the smooth host as the reference's perturbed-stream helpers do (streamhelpers.py:57-66,
    perturbative.py:253-278); the clumpy halo releases in the whole potential, clumps included
    (gen_stream_vmapped, main.py:287-306)."""
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.streams import spray_particles
    rel = wl["host"] if wl["name"] in ("subhalos", "responses") else wl["pot"]
    return spray_particles(wl["pot"], rel, wl["ts"], wl["prog_w0"], W.MSAT, 0,
                           wl["per_step"], W.DT)


def spray_rows(wl):
    """Initial rows and start rows of the workload's spray in integration order, computed once
    outside the timed region for the sample check."""
    if "spray_rows" not in wl:
        w0, start, _, _ = spray(wl)
        wl["spray_rows"] = (w0, start)
    return wl["spray_rows"]


def plain_stream_rows(wl):
    """Initial rows and start rows of the plain 8000-particle stream in the order gen_stream
    integrates them (per release: lead, then trail), computed once outside the timed region for
    the sample check."""
    import numpy as np
    import torch
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.streams import stream_ics
    if "rows" not in wl:
        (pc, pf, vc, vf), (t0, dt, nsteps, kidx), ts_np = stream_ics(
            wl["pot"], wl["pot"], wl["ts"], wl["prog_w0"], W.MSAT, 0, solver="leapfrog",
            dt0=W.DT)
        T = len(ts_np) - 1
        w0 = torch.empty(2 * T, 6, dtype=torch.float64, device=pc.device)
        w0[0::2, :3], w0[0::2, 3:] = pc[:T], vc[:T]
        w0[1::2, :3], w0[1::2, 3:] = pf[:T], vf[:T]
        wl["rows"] = (w0, np.repeat(kidx[:T], 2).astype(np.int32))
    return wl["rows"]


def run_plain(wl, ev=None):
    if wl["name"] == "plain":
        # the plain 8000-particle stream through gen_stream: progenitor orbit, releases and particles,
        # pipelined over ranges of release rows
        import torch
        from streamsculptor_amd import workloads as W
        from streamsculptor_amd.streams import gen_stream
        if ev:
            ev[0].record()
        lead, trail = gen_stream(wl["pot"], wl["pot"], wl["ts"], wl["prog_w0"], W.MSAT, 0,
                                 solver="leapfrog", dt0=W.DT)
        if ev:
            ev[1].record()
        out = torch.empty(2 * lead.shape[0], 6, dtype=torch.float64, device=lead.device)
        out[0::2], out[1::2] = lead, trail
        w0, start = plain_stream_rows(wl)
        return w0, start, out
    if wl["name"] == "subhalos":
        # the live-subhalo stream through spray_stream: the progenitor orbit in the live
        # subhalos pipelined with the particles over windows of grid rows
        from streamsculptor_amd import workloads as W
        from streamsculptor_amd.streams import spray_stream
        if ev:
            ev[0].record()
        out = spray_stream(wl["pot"], wl["host"], wl["ts"], wl["prog_w0"], W.MSAT, 0,
                           wl["per_step"], W.DT)
        if ev:
            ev[1].record()
        w0, start = spray_rows(wl)
        return w0, start, out
    w0, start, arm, (t0, dt, nsteps, _) = spray(wl)
    if ev:
        ev[0].record()
    if wl["name"] == "responses":
        from streamsculptor_amd.fields import perturb_leapfrog
        out, _ = perturb_leapfrog(wl["host"], wl["live"]._kernel_sources(), w0, None, start, t0,
                                  dt, nsteps, False)
    else:
        out, _ = wl["pot"]._leapfrog(w0, t0, dt, nsteps, start=start)
    if ev:
        ev[1].record()
    return w0, start, out


def adjoint_setup(wl, checkpoint_every=100):
    import numpy as np
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.adjoint import StreamModel
    sw0, m, rs = W.subhalo_ics()
    model = StreamModel(wl["ts"], sw0, rs, W.MSAT, wl["per_step"], seed=0,
                        checkpoint_every=checkpoint_every)
    prog_w0 = wl["prog_w0"].cpu().numpy()
    halo = np.array([W.HOST["m_halo"], W.HOST["rs_halo"]])
    st = model.forward(halo[0] * 1.05, halo[1] * 0.97, m, prog_w0)
    target = st["wT"][:, :3].clone()
    return dict(model=model, halo=halo, m=m, prog_w0=prog_w0, target=target)


def run_adjoint(wl, ad, ev=None):
    import torch
    from streamsculptor_amd.adjoint import stream_loss
    halo = torch.tensor(ad["halo"], dtype=torch.float64, device="cuda", requires_grad=True)
    ms = torch.tensor(ad["m"], dtype=torch.float64, device="cuda", requires_grad=True)
    pw = torch.tensor(ad["prog_w0"], dtype=torch.float64, device="cuda", requires_grad=True)
    if ev:
        ev[0].record()
    L = stream_loss(halo, ms, pw, ad["model"], ad["target"])
    L.backward()
    if ev:
        ev[1].record()
    return L, halo.grad, ms.grad, pw.grad


def run_surrogate(wl, sur, ev=None):
    w0, start, arm, (t0, dt, nsteps, _) = spray(wl)
    if ev:
        ev[0].record()
    out, cnt = sur.leapfrog(w0, t0, dt, nsteps, start=start)
    if ev:
        ev[1].record()
    return out, cnt


def timed(fn, steps, warmup, label):
    import torch
    for i in range(warmup):
        fn(None)
        torch.cuda.synchronize()
        log(f"{label}: warmup {i + 1}/{warmup} done")
    evs, res = [], None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        res = fn(ev)
        evs.append(ev)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern = sum(a.elapsed_time(b) for a, b in evs) / len(evs) / 1e3
    log(f"{label}: {steps} steps in {wall:.3f} s; dominant kernel {kern * 1e3:.1f} ms avg")
    return wall, kern, res


def train_surrogate(wl, iters):
    import numpy as np
    from streamsculptor_amd.surrogate import ClumpSurrogate, orbit_positions
    w0, start, arm, (t0, dt, nsteps, _) = spray(wl)
    sel = np.linspace(0, w0.shape[0] - 1, 2100).astype(np.int64)
    P = orbit_positions(wl["pot"], w0[sel], start[sel], t0, dt, nsteps)
    return ClumpSurrogate.train(wl["host"], wl["clumps"], P, iters=iters, log=log)


def host_cpu():
    """(threads, description) of the host cores the CPU baseline runs on: every CPU this
    process may run on, capped at 16 (the host share of one GPU on a shared node), with nproc
    and the CPU model reported beside it."""
    avail = len(os.sched_getaffinity(0))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = min(avail, 16)
    return threads, {"nproc": os.cpu_count(), "affinity": avail, "cpu_model": model}


def cpu_baseline(wl, w0, start, out, sample_every):
    """The oracle (-O3 x86-64-v3 + OpenMP build of the same algorithm) on a strided sample of
    the same particles, repeated until about CPU_SECONDS of CPU work; also returns the sample's
    max row error against the GPU result."""
    import numpy as np
    import oracle as O
    from streamsculptor_amd import workloads as W
    cores, cpu = host_cpu()
    h = O.host(W.host_components())
    src = None
    if wl["name"] == "clumps":
        pos, m, rs = W.clump_table()
        src = O.Sources(O.SRC_STATIC, gm=W.G * m, rs=rs, pos=pos)
    elif wl["name"] in ("subhalos", "responses"):
        tens = wl["live"]._kernel_sources().tensors()
        src = O.Sources(O.SRC_TRACK, gm=tens["gm"].cpu().numpy(), rs=tens["rs"].cpu().numpy(),
                        pos=tens["pos"].cpu().numpy())
    w0s = w0[::sample_every].cpu().numpy()
    sts = start[::sample_every]
    nsteps = len(wl["ts"]) - 1
    reps, el = 0, 0.0
    while reps == 0 or (el < CPU_SECONDS and reps < 1000):
        t0 = time.perf_counter()
        if wl["name"] == "responses":
            ref, _ = O.perturb_leapfrog(h, src, w0s, float(wl["ts"][0]), W.DT, nsteps, start=sts,
                                        nthreads=cores)
        else:
            ref = O.leapfrog(h, w0s, float(wl["ts"][0]), W.DT, nsteps, start=sts, src=src,
                             nthreads=cores, fast=True)
        el += time.perf_counter() - t0
        reps += 1
    psteps = int((nsteps - sts).sum())
    g = out[::sample_every].cpu().numpy()
    err = float((np.linalg.norm(g - ref, axis=1) / np.linalg.norm(ref, axis=1)).max())
    every = f"every {sample_every}th particle" if sample_every > 1 else "every particle"
    return dict(value=reps * psteps / el, unit="particle-steps/s", cores=cores, **cpu, kind="port",
                sample=f"{every} of the same stream ({len(w0s)} particles, {psteps} "
                       f"particle-steps) x {reps} passes, {el:.1f} s, oracle OpenMP"), err


def gd1_setup(B):
    import numpy as np
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.batch import SubhaloLineBatch
    from streamsculptor_amd.main import usys
    from streamsculptor_amd.potential import GalaMilkyWayPotential
    mw = GalaMilkyWayPotential(units=usys)
    grid = np.linspace(0.0, -W.GD1_AGE, int(W.GD1_AGE) + 1)
    back = mw.integrate_orbit(w0=W.GD1_TODAY, ts=grid, t0=0.0, t1=-W.GD1_AGE,
                              solver="dopri5").ys.cpu().numpy()
    pr = W.gd1_subhalo_prior(B)
    k = np.clip(np.rint(pr["t_impact"]).astype(np.int64), 1, len(grid) - 1)
    x0 = back[k, :3] + pr["b"][:, None] * pr["dir_b"]
    v0 = back[k, 3:] + pr["v"][:, None] * pr["dir_v"]
    batch = SubhaloLineBatch(pr["m"][:, None], pr["a"][:, None], x0[:, None], v0[:, None],
                             grid[k][:, None], np.full((B, 1), 100.0), units=usys)
    return dict(mw=mw, batch=batch, ts=W.gd1_ts(), B=B)


def run_gd1(g, ev=None):
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.batch import gen_stream_batch_with_pert, integrate_orbit_batch
    ic = integrate_orbit_batch(g["mw"], g["batch"], W.GD1_TODAY, ts=[-W.GD1_AGE], t0=0.0,
                               t1=-W.GD1_AGE)[:, 0]
    lead, trail, st = gen_stream_batch_with_pert(
        g["mw"], g["batch"], g["ts"], ic, 1e4, 532, rtol=1e-6, atol=1e-6, dtmin=0.05,
        max_steps=10_000, events=ev, return_stats=True)
    return ic, lead, trail, st


def gd1_main(args):
    import numpy as np
    import torch
    from streamsculptor_amd import flops
    from streamsculptor_amd import workloads as W
    B = args.sims
    g = gd1_setup(B)
    T = len(g["ts"]) - 1
    log(f"gd1: {B} simulations x {2 * T} particles, Dopri5 rtol = atol = 1e-6")
    wall, kern, (ic, lead, trail, st) = timed(lambda ev: run_gd1(g, ev), args.steps,
                                              args.warmup, "gd1")
    st = st.cpu().numpy().reshape(-1, 3)
    acc, rej, failed = int(st[:, 0].sum()), int(st[:, 1].sum()), int((st[:, 2] != 0).sum())
    fl = flops.dopri5_step(flops.HOST_GRAD + flops.HERNQUIST_GRAD, 0)
    achieved = fl * (acc + rej) / kern
    sims_s = B * args.steps / wall
    res = {"metric": "perturbed GD-1-like stream simulations/s (adaptive Dopri5, batched)",
           "value": sims_s, "unit": "simulations/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "gd1 batch", "simulations": B, "particles_per_sim": 2 * T,
                      "rtol": 1e-6, "atol": 1e-6, "dtmin": 0.05, "subhalos_per_sim": 1},
           "particle_steps_per_s": acc / (wall / args.steps),
           "accepted_steps": acc, "rejected_steps": rej, "failed_orbits": failed,
           "reference_a100": {"simulations_per_s": 700 / 11.30,
                              "note": "700 sims per 11.30 s batch including the KDE "
                                      "likelihood (GD1_and_SubhaloFitting_GPU.ipynb:754)"},
           "roofline": {"bound": "fp64-valu", "achieved": achieved / 1e12,
                        "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP64, "traffic": None,
                        "kernel": "dopri5_kernel<LINE, grouped> (particles)",
                        "flop_per_step": fl}}
    log(f"gd1: {sims_s:.1f} simulations/s, {acc} accepted + {rej} rejected steps per batch, "
        f"particle kernel {kern * 1e3:.1f} ms, {failed} failed orbits")
    if not args.no_cpu:
        import oracle as O
        cores, cpu = host_cpu()
        h = O.host(W.gala_mw_components())
        bt = g["batch"]
        ic_np = ic.cpu().numpy()
        sample = list(range(0, B, max(1, B // args.cpu_sims)))[:args.cpu_sims]
        st_g = st.reshape(B, T, 2, 3)
        t0 = time.perf_counter()
        errs, flips, psteps = [], 0, 0
        for b in sample:
            src = O.Sources(O.SRC_LINE, gm=W.G * bt.m[b], rs=bt.a[b], pos=bt.subhalo_x0[b],
                            vel=bt.subhalo_v[b], t0=bt.subhalo_t0[b], t_window=bt.t_window[b])
            lo, tr, sto = O.stream_dopri5(h, ic_np[b], g["ts"], 1e4, 532, src=src,
                                          release_host=h, nthreads=cores, fast=True,
                                          rtol=1e-6, atol=1e-6,
                                          dtmin=0.05, max_steps=10_000)
            psteps += int(sto[:, 0].sum())
            sto = sto.reshape(T, 2, 3)
            flips += int(np.sum(np.any(sto[:, :, :2] != st_g[b, :, :, :2], axis=2)))
            for a_, b_ in ((lead[b], lo), (trail[b], tr)):
                a_ = a_.cpu().numpy()
                errs.append(np.linalg.norm(a_ - b_, axis=1) / np.linalg.norm(b_, axis=1))
        el = time.perf_counter() - t0
        errs = np.concatenate(errs)
        res["cpu_baseline"] = {"value": len(sample) / el, "unit": "simulations/s",
                               "cores": cores, **cpu, "kind": "port",
                               "particle_steps_per_s": psteps / el,
                               "sample": f"{len(sample)} of the {B} simulations, oracle Dopri5 "
                                         f"(OpenMP), {el:.1f} s"}
        # Rounding differences between two FP64 implementations grow along these orbits (a
        # 1e-14 change of the initial row moves final rows by up to ~1e-6), and a step whose
        # error estimate lands within rounding of the acceptance threshold takes a different
        # (equally valid) step sequence: the distribution and the count of such orbits say
        # more than the maximum alone.
        res["parity_sample"] = {"orbits": int(errs.size), "median_rel_err": float(np.median(errs)),
                                "p999_rel_err": float(np.quantile(errs, 0.999)),
                                "max_rel_err": float(errs.max()),
                                "orbits_with_other_step_sequence": flips}
        res["parity_sample_max_rel_err"] = float(errs.max())
        err = float(np.quantile(errs, 0.999))
        log(f"cpu baseline {len(sample) / el:.2f} simulations/s on {cores} cores; sample "
            f"parity p99.9 {err:.2e}, {flips} orbits with another step sequence")
    print(json.dumps(res), flush=True)


def derivs_setup(n_arm, m):
    import numpy as np
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.main import usys
    from streamsculptor_amd.perturbative import BaseStreamModel
    from streamsculptor_amd.potential import (HernquistPotential,
                                              SubhaloLinePotentialCustom_dRadius_fromFunc,
                                              SubhaloLinePotentialCustom_fromFunc)
    from streamsculptor_amd.solvers import Dopri5
    rng = np.random.default_rng(11)
    base = W.host_potential()
    ts = np.linspace(-4000.0, 0.0, n_arm + 1)
    bs = BaseStreamModel(base, W.PROG_TODAY, ts, 1e4, 0, solver=Dopri5(), rtol=1e-10,
                         atol=1e-10, dtmin=0.01)
    prog = bs.prog_loc_fwd.cpu().numpy()
    # impacts on the stream's path: a point of the progenitor orbit at a random time after the
    # first quarter, offset by up to 5 scale radii, with a 150 km/s-like relative velocity
    k = rng.integers(n_arm // 4, n_arm, m)
    mass = 10 ** rng.uniform(6.0, 8.0, m)
    rs = 1.05 * np.sqrt(mass / 1e8)
    u = rng.normal(size=(m, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    x0 = prog[k, :3] + rng.uniform(0.0, 5.0, m)[:, None] * rs[:, None] * u
    v = prog[k, 3:] + rng.normal(scale=0.15, size=(m, 3))
    tw = np.full(m, 150.0)
    pert = SubhaloLinePotentialCustom_fromFunc(HernquistPotential, mass, rs, x0, v, ts[k], tw,
                                               units=usys)
    struct = SubhaloLinePotentialCustom_dRadius_fromFunc(HernquistPotential, mass, rs, x0, v,
                                                         ts[k], tw, units=usys)
    return dict(base=base, bs=bs, pert=pert, struct=struct, ts=ts, m=m,
                mass=mass, rs=rs, x0=x0, v=v, tc=ts[k], tw=tw)


DERIVS_TOL = dict(rtol=1e-11, atol=1e-11, dtmin=0.01, max_steps=100_000)


def run_derivs(d, ev=None):
    """One batch: the progenitor's backward field solve, response ICs, then every particle's
    adaptive response solve (both arms; ``ev`` brackets the particle launches)."""
    import torch
    from streamsculptor_amd.fields import perturb_dopri5
    from streamsculptor_amd.perturbative import GenerateMassRadiusPerturbation
    gen = GenerateMassRadiusPerturbation(d["base"], d["pert"], d["struct"],
                                         BaseStreamModel=d["bs"], **DERIVS_TOL)
    T = len(d["ts"]) - 1
    src = d["pert"]._kernel_sources()
    w0 = torch.cat([gen.base_realspace_ICs_lead[:T], gen.base_realspace_ICs_trail[:T]])
    d0 = torch.cat([gen.perturbation_ICs_lead[:T], gen.perturbation_ICs_trail[:T]])
    tst = torch.cat([torch.as_tensor(d["ts"][:T])] * 2)
    if ev:
        ev[0].record()
    out = perturb_dopri5(d["base"], src, w0, d0, float(d["ts"][-1]), t_start=tst,
                         with_radius=True, **DERIVS_TOL)
    if ev:
        ev[1].record()
    return w0, d0, tst, out


def derivs_main(args):
    import numpy as np
    from streamsculptor_amd import flops
    from streamsculptor_amd import workloads as W
    n_arm, m = 5000, 128
    d = derivs_setup(n_arm, m)
    log(f"derivs: {2 * n_arm} particles x {m} Hernquist subhalos (mass + radius blocks), "
        f"Dopri5 rtol = atol = 1e-11")
    wall, kern, (w0, d0, tst, out) = timed(lambda ev: run_derivs(d, ev), args.steps,
                                           args.warmup, "derivs")
    st = out[4].cpu().numpy()
    acc, rej, failed = int(st[:, 0].sum()), int(st[:, 1].sum()), int((st[:, 2] != 0).sum())
    fl = flops.perturb_dopri5_step(2 * m)
    achieved = fl * (acc + rej) / kern
    res = {"metric": "particle-steps/s of adaptive first-order responses (Dopri5, one "
                     "controller per particle over base row + responses)",
           "value": acc / kern, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": "generate_derivs batch", "particles": 2 * n_arm,
                      "subhalos": m, "profile": "hernquist", "radius_block": True,
                      "response_vectors": 2 * m, **DERIVS_TOL},
           "batches_per_s": args.steps / wall, "accepted_steps": acc, "rejected_steps": rej,
           "failed": failed,
           "roofline": {"bound": "fp64-valu", "achieved": achieved / 1e12,
                        "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP64, "traffic": None,
                        "kernel": "perturb_dopri5_kernel<LINE_HERNQUIST, 1, 1>",
                        "kernel_ms": kern * 1e3, "flop_per_step": fl}}
    log(f"derivs: {acc / kern:.3e} particle-steps/s ({acc} accepted + {rej} rejected), "
        f"kernel {kern * 1e3:.1f} ms, {achieved / 1e12:.2f} TFLOP/s, {failed} failed")
    if not args.no_cpu:
        import oracle as O
        cores, cpu = host_cpu()
        h = O.host(W.host_components())
        so = O.Sources(O.SRC_LINE_HERNQUIST, gm=W.G * d["mass"], rs=d["rs"], pos=d["x0"],
                       vel=d["v"], t0=d["tc"], t_window=d["tw"])
        n = w0.shape[0]
        every = args.cpu_sample_every or 10
        sel = np.arange(0, n, every)
        w_np, d_np, t_np = w0.cpu().numpy()[sel], d0.cpu().numpy()[sel], tst.cpu().numpy()[sel]
        tol = dict(DERIVS_TOL)
        t0 = time.perf_counter()
        wr, dwr, _, _, sto = O.perturb_dopri5(h, so, w_np, float(d["ts"][-1]), t_start=t_np,
                                              dw0=d_np, with_radius=True, nthreads=cores,
                                              **tol)
        el = time.perf_counter() - t0
        dw = out[1].cpu().numpy()[sel]
        rel = np.linalg.norm((dw - dwr).reshape(len(sel), -1), axis=1) / np.maximum(
            np.linalg.norm(dwr.reshape(len(sel), -1), axis=1), 1e-300)
        res["cpu_baseline"] = {"value": float(sto[:, 0].sum()) / el, "unit": "particle-steps/s",
                               "cores": cores, **cpu, "kind": "port",
                               "sample": f"every {every}th particle ({len(sel)} of {n}), "
                                         f"oracle Dopri5 responses (OpenMP), {el:.1f} s"}
        res["parity_sample_max_rel_err"] = float(rel.max())
        res["parity_sample_same_steps"] = float((sto == st[sel]).all(axis=1).mean())
        log(f"cpu baseline {float(sto[:, 0].sum()) / el:.3e} particle-steps/s on {cores} "
            f"cores; sample parity {rel.max():.2e}")
    print(json.dumps(res), flush=True)


GRAD_BAR = 0.05


def ensemble_main(args):
    """Config: the clumps potential and surrogate with a differentiable ensemble of streams
    from many progenitors of different ages (loss + gradient through the surrogate adjoint)."""
    import numpy as np
    import torch
    from streamsculptor_amd import flops
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.ensemble import SurrogateStreamEnsemble, ensemble_loss
    from streamsculptor_amd.main import usys
    from streamsculptor_amd.potential import PlummerClumps
    from streamsculptor_amd.surrogate import ClumpSurrogate
    S = args.streams
    lengths, prog = W.ensemble_setup(S)
    host = W.host_potential()
    cpos, cm, crs = W.clump_table()
    clumps = PlummerClumps(cm, crs, cpos, units=usys)
    ens = SurrogateStreamEnsemble(None, lengths, n_release=args.releases, msat=W.MSAT, seed=0,
                                  checkpoint_every=100, clumps=clumps)
    mh, rsh = W.HOST["m_halo"], W.HOST["rs_halo"]
    ps = ens.particle_steps
    log(f"ensemble: {S} progenitors (lengths {lengths.min()}-{lengths.max()} steps), "
        f"{ens.n} particles, {ps} particle-steps per loss")
    t0 = time.perf_counter()
    P = ens.training_positions(ens.host(mh, rsh), prog)
    sur = ClumpSurrogate.train(host, clumps, P, iters=args.train_iters, log=log)
    ens.sur = sur
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    rep = {k: v for k, v in sur.report.items() if k != "losses"}
    log(f"surrogate trained on {P.shape[0]} ensemble positions in {t_train:.1f} s: {rep}")
    del P
    target = ens.forward(mh * 1.05, rsh * 0.97, prog)["wT"][:, :3].clone()
    torch.cuda.synchronize()

    def step(ev, exact=False):
        halo = torch.tensor([mh, rsh], dtype=torch.float64, device="cuda", requires_grad=True)
        pw = torch.tensor(prog, dtype=torch.float64, device="cuda", requires_grad=True)
        L = ensemble_loss(halo, pw, ens, target, exact=exact, events=ev)
        L.backward()
        return L.detach(), halo.grad, pw.grad

    wall, kern, (L, gh, gw) = timed(step, args.steps, args.warmup, "ensemble loss + gradient")
    fl = flops.mlp_input_vjp()
    achieved = fl * ps / kern
    res = {"metric": "particle-steps/s of loss + gradient through the surrogate (ensemble)",
           "value": ps * args.steps / wall, "unit": "particle-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16-mfma/f64", "data": "synthetic",
           "config": {"workload": "surrogate ensemble adjoint", "progenitors": S,
                      "stream_lengths": [int(lengths.min()), int(lengths.max())],
                      "releases_per_stream": args.releases, "particles": ens.n,
                      "particle_steps": ps, "checkpoint_every": 100, "clumps": 512,
                      "gradients": "halo (m, r_s) and every progenitor's initial row"},
           "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_BF16 / 1e12,
                        "unit": "TFLOP/s", "frac": achieved / PEAK_BF16, "traffic": None,
                        "kernel": "surrogate_adjoint_segment_kernel (particles' reverse sweep)",
                        "kernel_ms": kern * 1e3, "flop_per_particle_step": fl},
           "train_seconds": t_train, "train": rep, "loss": float(L.item()),
           "grad_halo": gh.tolist()}
    log(f"ensemble: {res['value']:.3e} particle-steps/s (loss + gradient), reverse sweep "
        f"{kern * 1e3:.0f} ms = {achieved / 1e12:.0f} TFLOP/s")
    if not args.no_exact:
        # the same loss and gradient with the exact clump sum for the particles (plain adjoint)
        t1 = time.perf_counter()
        Le, ghe, gwe = step(None, exact=True)
        torch.cuda.synchronize()
        t_exact = time.perf_counter() - t1
        st_s = ens.forward(mh, rsh, prog)["wT"]
        st_e = ens.forward(mh, rsh, prog, exact=True)["wT"]
        med_pc = float(torch.linalg.norm(st_s[:, :3] - st_e[:, :3], dim=1).median().item()) * 1e3
        rel_h = (torch.abs(gh - ghe) / torch.abs(ghe)).tolist()
        rel_w = float((torch.linalg.norm(gw - gwe) / torch.linalg.norm(gwe)).item())
        res["exact"] = {"seconds_loss_and_gradient": t_exact,
                        "particle_steps_per_s": ps / t_exact, "loss": float(Le.item()),
                        "grad_halo": ghe.tolist(), "median_final_position_err_pc": med_pc,
                        "grad_halo_rel_diff": rel_h, "grad_progenitors_rel_diff": rel_w,
                        # the surrogate gradient's bar: within GRAD_BAR of the exact-sum
                        # gradient for each halo parameter and normwise over the progenitors
                        "gradient_bar": GRAD_BAR,
                        "gradient_bar_met": bool(max(rel_h) <= GRAD_BAR and rel_w <= GRAD_BAR)}
        log(f"exact clump sum: {t_exact:.1f} s per loss + gradient; surrogate median position "
            f"error {med_pc:.2f} pc, halo gradient rel. diff {rel_h}, progenitor rows {rel_w:.2e}")
    if not args.no_cpu:
        import threadpoolctl
        import oracle as O
        import oracle.surrogate as OS
        sc = sur.scaling
        h = O.host(W.host_components())
        osrc = O.Sources(O.SRC_STATIC, gm=W.G * cm, rs=crs, pos=cpos)
        om = OS.Model(h, osrc, sur.mlp.params.cpu().numpy(), sc.center, sc.half, sc.g_scale,
                      sur.grid_lo, sur.grid_cell, sur.grid_n, sur.flags.cpu().numpy())
        parts = ens.particles(ens.host(mh, rsh), prog)
        sel = np.arange(0, ens.n, ens.n // args.cpu_particles)[:args.cpu_particles]
        start = np.concatenate([ens.rel_step.reshape(-1)] * 2)[sel]
        w0s = parts["w0"][torch.as_tensor(sel, device="cuda")].cpu().numpy()
        lam = np.random.default_rng(0).normal(size=(len(sel), 6))
        with threadpoolctl.threadpool_limits(1):
            tc = time.perf_counter()
            OS.adjoint(om, w0s, 1.0, ens.T, lam, start=start)
            el = time.perf_counter() - tc
        psteps = int((ens.T - start).sum())
        res["cpu_baseline"] = {"value": psteps / el, "unit": "particle-steps/s", "cores": 1,
                               **host_cpu()[1], "kind": "port",
                               "sample": f"{len(sel)} particles of the ensemble (every "
                                         f"{ens.n // args.cpu_particles}th), forward + adjoint "
                                         f"through the oracle surrogate (numpy, BF16 emulation), "
                                         f"{psteps} particle-steps, {el:.1f} s"}
        log(f"cpu baseline {psteps / el:.3e} particle-steps/s (1 core)")
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="clumps",
                    choices=["clumps", "subhalos", "responses", "adjoint", "plain", "gd1",
                             "ensemble", "derivs"])
    ap.add_argument("--streams", type=int, default=2000, help="ensemble: progenitors")
    ap.add_argument("--releases", type=int, default=1000, help="ensemble: release steps each")
    ap.add_argument("--cpu-particles", type=int, default=16, help="ensemble: CPU sample size")
    ap.add_argument("--no-exact", action="store_true", help="ensemble: skip the exact run")
    ap.add_argument("--no-stored", action="store_true",
                    help="adjoint: skip the run with every particle row stored")
    ap.add_argument("--sims", type=int, default=700, help="gd1: simulations per batch")
    ap.add_argument("--cpu-sims", type=int, default=600, help="gd1: simulations in the CPU sample")
    ap.add_argument("--particles", type=int, default=None)
    ap.add_argument("--train-iters", type=int, default=96000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--cpu-sample-every", type=int, default=None,
                    help="CPU baseline stride (default: 16 for clumps, 1 otherwise)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: particles are independent, so a multi-GPU run is "
              "N copies of this one on disjoint particle ranges; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import streamsculptor_amd  # noqa: F401
    from streamsculptor_amd import _lib, flops
    _lib.require_gpu()
    log(f"device {torch.cuda.get_device_name(0)}; building workload {args.workload}")
    if args.workload == "gd1":
        gd1_main(args)
        return
    if args.workload == "ensemble":
        ensemble_main(args)
        return
    if args.workload == "derivs":
        derivs_main(args)
        return
    wl = build_workload(args.workload, args.particles)
    ps = wl["psteps"]
    log(f"{wl['total']} particles, {ps} particle-steps per step, {wl['nsrc']} sources")

    if args.workload == "adjoint":
        ad = adjoint_setup(wl)
        wall_a, kern_a, (L, gh, gm, gw) = timed(lambda ev: run_adjoint(wl, ad, ev), args.steps,
                                                args.warmup, "adjoint")
        fl_a = 5025
        res = {"metric": "particle-steps/s of loss + gradient (discrete leapfrog adjoint)",
               "value": ps * args.steps / wall_a, "unit": "particle-steps/s", "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall_a / args.steps,
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
               "data": "synthetic",
               "config": {"workload": "adjoint of the subhalos stream", "particles": wl["total"],
                          "sources": wl["nsrc"], "steps_per_orbit": 4000, "dt_myr": 1.0,
                          "particle_steps": ps, "checkpoint_every": 100,
                          "gradients": "halo (m, r_s), 64 subhalo masses, progenitor row"},
               "roofline": {"bound": "fp64-valu", "achieved": fl_a * ps / kern_a / 1e12,
                            "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                            "frac": fl_a * ps / kern_a / PEAK_FP64, "traffic": None,
                            "kernel": "loss + backward (all kernels)",
                            "flop_per_particle_step": fl_a},
               "loss": float(L.item()), "grad_halo": gh.tolist(),
               "grad_prog_w0": gw.tolist()}
        if args.no_stored:
            print(json.dumps(res), flush=True)
            return
        # the same loss and gradient with every particle row kept (nothing re-run backwards)
        del ad
        ad1 = adjoint_setup(wl, checkpoint_every=1)
        wall_1, _, (L1, gh1, _, _) = timed(lambda ev: run_adjoint(wl, ad1, ev), args.steps,
                                           args.warmup, "adjoint, every row stored")
        res["stored_rows"] = {"checkpoint_every": 1, "ms_per_step": 1e3 * wall_1 / args.steps,
                              "value": ps * args.steps / wall_1, "loss": float(L1.item()),
                              "grad_halo": gh1.tolist()}
        print(json.dumps(res), flush=True)
        return

    wall_p, kern_p, (w0, start, out_p) = timed(lambda ev: run_plain(wl, ev), args.steps,
                                               args.warmup, "plain")
    if args.workload == "responses":
        fl_p = flops.responses(wl["nsrc"])
    else:
        fl_p = flops.plain(wl["nsrc"])
    plain = {"value": ps * args.steps / wall_p, "ms_per_step": 1e3 * wall_p / args.steps,
             "kernel_ms": kern_p * 1e3,
             "roofline": {"bound": "fp64-valu", "achieved": fl_p * ps / kern_p / 1e12,
                          "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                          "frac": fl_p * ps / kern_p / PEAK_FP64,
                          "traffic": (TRAFFIC_CLUMPS_PLAIN if args.workload == "clumps"
                                      and wl["total"] == 1_000_000 else None),
                          "kernel": ("perturb_base_kernel + perturb_rows_kernel (one launch chain)"
                                     if args.workload == "responses" else
                                     "the whole stream: progenitor orbit, releases and particle "
                                     "leapfrog, pipelined" if args.workload in ("plain", "subhalos")
                                     else "leapfrog_kernel"),
                          "flop_per_particle_step": fl_p}}
    result = {
        "metric": "particle-steps/s, surrogate vs plain FP64 force, at <=10 pc median stream error",
        "value": plain["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": plain["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.workload} stream", "particles": wl["total"],
                   "sources": wl["nsrc"], "steps_per_orbit": 4000, "dt_myr": 1.0,
                   "particle_steps": ps, "form": "plain"},
        "roofline": plain["roofline"],
        "plain": plain,
    }

    if args.workload == "clumps" and not args.no_surrogate:
        t0 = time.perf_counter()
        sur = train_surrogate(wl, args.train_iters)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        rep = {k: v for k, v in sur.report.items() if k != "losses"}
        log(f"surrogate trained in {t_train:.1f} s: {rep}")
        wall_s, kern_s, (out_s, cnt) = timed(lambda ev: run_surrogate(wl, sur, ev), args.steps,
                                             args.warmup, "surrogate")
        dpos = torch.linalg.norm(out_s[:, :3] - out_p[:, :3], dim=1)
        med_pc = float(dpos.median().item()) * 1e3
        from streamsculptor_amd.workloads import MSAT  # noqa: F401
        host_only, _ = wl["host"]._leapfrog(w0, float(wl["ts"][0]), 1.0, len(wl["ts"]) - 1,
                                            start=start)
        med_host_pc = float(torch.linalg.norm(host_only[:, :3] - out_p[:, :3], dim=1).median().item()) * 1e3
        fl_mlp = flops.mlp()
        fb_frac = int(cnt.item()) / ps
        surrogate = {"value": ps * args.steps / wall_s, "ms_per_step": 1e3 * wall_s / args.steps,
                     "kernel_ms": kern_s * 1e3, "median_err_pc": med_pc,
                     "median_err_pc_host_only": med_host_pc, "fallback_fraction": fb_frac,
                     "train_seconds": t_train, "train": rep,
                     "roofline": {"bound": "mfma", "achieved": fl_mlp * ps / kern_s / 1e12,
                                  "peak": PEAK_BF16 / 1e12, "unit": "TFLOP/s",
                                  "frac": fl_mlp * ps / kern_s / PEAK_BF16, "traffic": None,
                                  "kernel": "surrogate_leapfrog_kernel",
                                  "flop_per_particle_step": fl_mlp}}
        result["surrogate"] = surrogate
        result["surrogate_vs_plain"] = surrogate["value"] / plain["value"]
        log(f"surrogate {surrogate['value']:.3e} vs plain {plain['value']:.3e} particle-steps/s; "
            f"median error {med_pc:.2f} pc (host only {med_host_pc:.1f} pc); "
            f"fallback {fb_frac:.3f}")
        # value: the faster form that meets the 10 pc bar (the exact plain form always does)
        if med_pc <= ERR_LIMIT_PC and surrogate["value"] > plain["value"]:
            result.update(value=surrogate["value"], ms_per_step=surrogate["ms_per_step"],
                          roofline=surrogate["roofline"], dtype="bf16-mfma/f64")
            result["config"]["form"] = "surrogate"
        elif med_pc <= ERR_LIMIT_PC:
            result["config"]["form"] = "plain (surrogate within 10 pc but slower)"
        else:
            result["config"]["form"] = "plain (surrogate error above 10 pc)"

    if not args.no_cpu:
        log("timing the CPU baseline (oracle on a sample)")
        every = args.cpu_sample_every or (16 if args.workload == "clumps" else 1)
        cb, err = cpu_baseline(wl, w0, start, out_p, every)
        result["cpu_baseline"] = cb
        result["parity_sample_max_rel_err"] = err
        log(f"cpu baseline {cb['value']:.3e} particle-steps/s on {cb['cores']} cores; "
            f"sample parity {err:.2e}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
