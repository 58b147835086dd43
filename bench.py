#!/usr/bin/env python3
"""Throughput of the stream-particle integration path on one MI355X.

Metric: particle-steps/s (one spray particle advanced one 1-Myr leapfrog step).  The default
workload is the clumpy halo: the 3-component host plus 512 static Plummer clumps, 1e6 spray
particles (250 released per step) integrated to the present over 4000 x 1 Myr, i.e.
2,000,500,000 particle-steps per bench step.  A step is one full stream: progenitor orbit,
release conditions and the particle integration, with inputs resident in HBM.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload clumps|subhalos|plain]
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
PEAK_HBM = 8.0e12        # B/s


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name, particles=None):
    import numpy as np
    import torch
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.potential import (LiveSubhaloPotential, PlummerClumps,
                                              Potential_Combine)
    from streamsculptor_amd.main import usys
    host = W.host_potential()
    ts = W.release_times()
    if name == "clumps":
        pos, m, rs = W.clump_table()
        pot = Potential_Combine([host, PlummerClumps(m, rs, pos, units=usys)], units=usys)
        total, nsrc = 1_000_000, 512
    elif name == "subhalos":
        sw0, m, rs = W.subhalo_ics()
        live = LiveSubhaloPotential.from_orbits(host, sw0, m, rs, float(ts[0]), W.DT, W.T_STEPS)
        pot = Potential_Combine([host, live], units=usys)
        total, nsrc = 100_000, 64
    else:
        pot, total, nsrc = host, 8000, 0
    if particles:
        total = particles
    per_step = W.particles_per_step(total)
    # the progenitor's present-day point integrated back 4 Gyr in the full potential
    back = pot.integrate_orbit(w0=W.PROG_TODAY, ts=[0.0, ts[0]], t0=0.0, t1=float(ts[0]),
                               solver="leapfrog", dt0=W.DT).ys[-1]
    torch.cuda.synchronize()
    return dict(name=name, pot=pot, host=host, ts=ts, prog_w0=back, per_step=per_step,
                total=total, nsrc=nsrc, psteps=W.particle_steps(per_step))


def run_stream(wl, timer=None):
    import torch
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.streams import spray_particles
    pot = wl["pot"]
    w0, start, arm, (t0, dt, nsteps, _) = spray_particles(
        pot, pot, wl["ts"], wl["prog_w0"], W.MSAT, 0, wl["per_step"], W.DT)
    if timer is not None:
        timer[0].record()
    out, _ = pot._leapfrog(w0, t0, dt, nsteps, start=start)
    if timer is not None:
        timer[1].record()
    return w0, start, out


def cpu_baseline(wl, w0, start, out, sample_every):
    """The oracle (-O3 x86-64-v3 + OpenMP build of the same algorithm) on a strided sample of
    the same particles; also returns the sample's max row error against the GPU result."""
    import numpy as np
    import oracle as O
    from streamsculptor_amd import workloads as W
    cores = min(len(os.sched_getaffinity(0)), 16)
    h = O.host(W.host_components())
    src = None
    if wl["name"] == "clumps":
        pos, m, rs = W.clump_table()
        src = O.Sources(O.SRC_STATIC, gm=W.G * m, rs=rs, pos=pos)
    elif wl["name"] == "subhalos":
        tens = wl["pot"]._kernel_sources().tensors()
        src = O.Sources(O.SRC_TRACK, gm=tens["gm"].cpu().numpy(), rs=tens["rs"].cpu().numpy(),
                        pos=tens["pos"].cpu().numpy())
    w0s = w0[::sample_every].cpu().numpy()
    sts = start[::sample_every]
    nsteps = len(wl["ts"]) - 1
    t0 = time.perf_counter()
    ref = O.leapfrog(h, w0s, float(wl["ts"][0]), W.DT, nsteps, start=sts, src=src,
                     nthreads=cores, fast=True)
    el = time.perf_counter() - t0
    psteps = int((nsteps - sts).sum())
    g = out[::sample_every].cpu().numpy()
    err = float((np.linalg.norm(g - ref, axis=1) / np.linalg.norm(ref, axis=1)).max())
    return dict(value=psteps / el, unit="particle-steps/s", cores=cores, kind="port",
                sample=f"every {sample_every}th particle of the same stream ({len(w0s)} particles, "
                       f"{psteps} particle-steps, {el:.1f} s, oracle -O3 x86-64-v3 OpenMP)"), err


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="clumps", choices=["clumps", "subhalos", "plain"])
    ap.add_argument("--particles", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample-every", type=int, default=100)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: particles are independent, so a multi-GPU run is "
              "N copies of this one (see the sharding notes); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import streamsculptor_amd  # noqa: F401
    from streamsculptor_amd import _lib, flops
    _lib.require_gpu()
    log(f"device {torch.cuda.get_device_name(0)}; building workload {args.workload}")
    wl = build_workload(args.workload, args.particles)
    log(f"{wl['total']} particles, {wl['psteps']} particle-steps per step, {wl['nsrc']} sources")

    for i in range(args.warmup):
        run_stream(wl)
        torch.cuda.synchronize()
        log(f"warmup {i + 1}/{args.warmup} done")

    kern_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        w0, start, out = run_stream(wl, ev)
        kern_ms.append(ev)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern = [a.elapsed_time(b) for a, b in kern_ms]
    log(f"timed {args.steps} steps in {wall:.3f} s; particle kernel {sum(kern) / len(kern):.1f} ms avg")

    ms_per_step = 1e3 * wall / args.steps
    value = wl["psteps"] * args.steps / wall
    fl = flops.plain(wl["nsrc"])
    kavg = sum(kern) / len(kern) / 1e3
    achieved = fl * wl["psteps"] / kavg
    result = {
        "metric": "particle-steps/s, surrogate vs plain FP64 force, at <=10 pc median stream error",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} stream, plain FP64 force",
                   "particles": wl["total"], "sources": wl["nsrc"], "steps_per_orbit": 4000,
                   "dt_myr": 1.0, "particle_steps": wl["psteps"], "form": "plain"},
        "roofline": {"bound": "fp64-valu", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64, "traffic": None,
                     "kernel": "leapfrog_kernel", "kernel_ms": kavg * 1e3,
                     "flop_per_particle_step": fl},
    }
    if not args.no_cpu:
        log("timing the CPU baseline (oracle on a sample)")
        cb, err = cpu_baseline(wl, w0, start, out, args.cpu_sample_every)
        result["cpu_baseline"] = cb
        result["parity_sample_max_rel_err"] = err
        log(f"cpu baseline {cb['value']:.3e} particle-steps/s on {cb['cores']} cores; "
            f"sample parity {err:.2e}")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
