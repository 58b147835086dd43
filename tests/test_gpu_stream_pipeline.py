"""The fixed-step spray stream pipelined over windows of grid rows (the progenitor orbit
advancing window by window while, on a second stream, every particle released so far advances
through the window behind it) against the one-pass path: the plain spray stream (host only,
4000 steps, two particles per step) bit for bit in 4, 7 and the default 8 windows; a clumpy
halo with a mass-loss history (Msat per release) to 1e-11 (a restart forms the acceleration at
a window's first row in the kernel's prologue, whose FMA contraction can differ in the last bit
from the step loop's when the source sum is split over lanes; grown along 1000 steps among the
clumps: 1.5e-13 measured); a time-dependent potential takes the one-pass path."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return ((a - b).norm(dim=1) / b.norm(dim=1)).max().item()


def _both(pot, ts, prog_w0, msat, seed, exact=True):
    from streamsculptor_amd.streams import gen_stream
    a = gen_stream(pot, pot, ts, prog_w0, msat, seed, solver="leapfrog", dt0=1.0, chunks=4)
    b = gen_stream(pot, pot, ts, prog_w0, msat, seed, solver="leapfrog", dt0=1.0, chunks=1)
    c = gen_stream(pot, pot, ts, prog_w0, msat, seed, solver="leapfrog", dt0=1.0, chunks=7)
    e = max(_rel(x, y) for x, y in zip(a + c, b + b))
    print(f"[pipeline] windows vs one pass: {e:.2e}", flush=True)
    if exact:
        assert all(torch.equal(x, y) for x, y in zip(a + c, b + b))
    assert e <= 1e-11
    return a, b


def test_pipelined_stream_equals_one_pass(oracle, ssc):
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.main import usys
    from streamsculptor_amd.potential import PlummerClumps, Potential_Combine
    host = W.host_potential()
    ts = W.release_times()
    back = host.integrate_orbit(w0=W.PROG_TODAY, ts=[0.0, ts[0]], t0=0.0, t1=float(ts[0]),
                                solver="leapfrog", dt0=1.0).ys[-1]
    (l4, t4), (l1, t1) = _both(host, ts, back, W.MSAT, 0)
    # the public entry point takes the pipelined path and still matches the oracle stream
    lead, trail = host.gen_stream_vmapped(ts=ts, prog_w0=back.cpu().numpy(), Msat=W.MSAT,
                                          seed_num=0, solver="leapfrog", dt0=1.0)
    assert torch.equal(lead, l1) and torch.equal(trail, t1)
    h = oracle.host(W.host_components())
    ref_lead, _ = oracle.stream_leapfrog(h, back.cpu().numpy(), ts, W.MSAT, 0, 1.0)
    e = (np.linalg.norm(lead.cpu().numpy() - ref_lead, axis=1) /
         np.linalg.norm(ref_lead, axis=1)).max()
    assert e <= 1e-9
    # static clumps, Msat falling along the stream
    pos, m, rs = W.clump_table(96, seed=3)
    pot = Potential_Combine([host, PlummerClumps(m, rs, pos, units=usys)], units=usys)
    ts2 = np.linspace(-1000.0, 0.0, 1001)
    msat = np.linspace(2e4, 1e3, ts2.size)
    _both(pot, ts2, back, msat, 3, exact=False)


def test_time_dependent_potential_takes_one_pass(ssc):
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.main import usys
    from streamsculptor_amd.potential import (HernquistPotential, Potential_Combine,
                                              TimeDepTranslatingPotential)
    from streamsculptor_amd.streams import _static_in_time
    host = W.host_potential()

    def centre(t):
        t = np.asarray(t, dtype=np.float64)
        return np.stack([40.0 + 0.0 * t, -30.0 - 0.02 * t, 10.0 + 0.0 * t], axis=-1)

    lmc = TimeDepTranslatingPotential(HernquistPotential(m=1e11, r_s=15.0, units=usys), centre,
                                      units=usys)
    pot = Potential_Combine([host, lmc], units=usys)
    assert _static_in_time(host) and not _static_in_time(pot)
    ts = np.linspace(-500.0, 0.0, 501)
    back = np.array([11.0, 1.0, 0.5, 0.02, 0.17, 0.04])
    (l4, t4), (l1, t1) = _both(pot, ts, back, 1e4, 1)
    assert torch.equal(l4, l1) and torch.equal(t4, t1)


def _spray_both(pot, rel_pot, ts, prog_w0, per_step, seed=0):
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.streams import spray_particles, spray_stream
    w0, start, _, (t0, dt, nsteps, _) = spray_particles(pot, rel_pot, ts, prog_w0, W.MSAT, seed,
                                                        per_step, 1.0)
    one, _ = pot._leapfrog(w0, t0, dt, nsteps, start=start)
    piped = spray_stream(pot, rel_pot, ts, prog_w0, W.MSAT, seed, per_step, 1.0)
    e = _rel(piped, one)
    print(f"[spray pipeline] {w0.shape[0]} particles, windows vs one pass: {e:.2e}", flush=True)
    return piped, one, e


def test_spray_stream_equals_one_pass(ssc):
    """The many-particle spray with the progenitor pipelined over windows (spray_stream)
    against spray_particles + one integration: bit for bit in the host alone; to 1e-11 with
    STATIC clumps and with live subhalos on a TRACK table (the acceleration at a window's first
    row comes from the kernel's prologue); a potential with straight-line subhalos takes the
    one-pass path."""
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.main import usys
    from streamsculptor_amd.potential import (LiveSubhaloPotential, PlummerClumps,
                                              Potential_Combine, SubhaloLinePotential)
    from streamsculptor_amd.streams import _restartable
    host = W.host_potential()
    ts = np.linspace(-400.0, 0.0, 401)
    back = host.integrate_orbit(w0=W.PROG_TODAY, ts=[0.0, ts[0]], t0=0.0, t1=float(ts[0]),
                                solver="leapfrog", dt0=1.0).ys[-1]
    piped, one, _ = _spray_both(host, host, ts, back, 25)
    assert tuple(piped.shape) == (25 * 400, 6) and torch.equal(piped, one)
    pos, m, rs = W.clump_table(96, seed=3)
    clumpy = Potential_Combine([host, PlummerClumps(m, rs, pos, units=usys)], units=usys)
    assert _spray_both(clumpy, clumpy, ts, back, 24, seed=2)[2] <= 1e-11
    sw0, sm, srs = W.subhalo_ics(n=16, seed=2)
    live = LiveSubhaloPotential.from_orbits(host, sw0, sm, srs, float(ts[0]), 1.0, 400)
    pot = Potential_Combine([host, live], units=usys)
    assert _restartable(pot)
    assert _spray_both(pot, host, ts, back, 25)[2] <= 1e-11
    rng = np.random.default_rng(5)
    line = SubhaloLinePotential(10 ** rng.uniform(6, 8, 8), rng.uniform(0.1, 0.5, 8),
                                rng.normal(scale=15.0, size=(8, 3)),
                                rng.normal(scale=0.2, size=(8, 3)), rng.uniform(-400, 0, 8),
                                rng.uniform(50, 200, 8), units=usys)
    pot_l = Potential_Combine([host, line], units=usys)
    assert not _restartable(pot_l)
    piped, one, _ = _spray_both(pot_l, host, ts, back, 6)
    assert torch.equal(piped, one)
