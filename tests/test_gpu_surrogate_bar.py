"""The learned clump force against its accuracy bar, at the clumpy-halo workload's shape: the
3-component host plus the 512 Plummer clumps of clump_table(seed=1), 1e5 spray particles
released over the full 4000 x 1 Myr.  For three training seeds of the default training
schedule (positions along 2,100 plain-method orbits, residual threshold 1e-3), the median
distance of the surrogate run's final positions to the plain FP64 run's is at most 10 pc."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def clumpy_workload(ssc):
    """The 1e5-particle clumpy-halo stream, its plain FP64 run and the training positions."""
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.main import usys
    from streamsculptor_amd.potential import PlummerClumps, Potential_Combine
    from streamsculptor_amd.streams import spray_particles
    from streamsculptor_amd.surrogate import orbit_positions
    host = W.host_potential()
    pos, m, rs = W.clump_table()
    clumps = PlummerClumps(m, rs, pos, units=usys)
    pot = Potential_Combine([host, clumps], units=usys)
    ts = W.release_times()
    per_step = W.particles_per_step(100_000)
    prog = pot.integrate_orbit(w0=W.PROG_TODAY, ts=[0.0, ts[0]], t0=0.0, t1=float(ts[0]),
                               solver="leapfrog", dt0=W.DT).ys[-1]
    w0, start, _, (t0, dt, nsteps, _) = spray_particles(pot, pot, ts, prog, W.MSAT, 0, per_step,
                                                        W.DT)
    plain, _ = pot._leapfrog(w0, t0, dt, nsteps, start=start)
    sel = np.linspace(0, w0.shape[0] - 1, 2100).astype(np.int64)
    P = orbit_positions(pot, w0[sel], start[sel], t0, dt, nsteps)
    return dict(host=host, clumps=clumps, w0=w0, start=start, grid=(t0, dt, nsteps),
                plain=plain, P=P)


# one test per seed: each ends (and reports) within about a minute and a half
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_surrogate_median_error(clumpy_workload, seed):
    from streamsculptor_amd.surrogate import ClumpSurrogate
    cw = clumpy_workload
    t0, dt, nsteps = cw["grid"]
    print(f"[surrogate bar] training seed {seed}", flush=True)
    sur = ClumpSurrogate.train(cw["host"], cw["clumps"], cw["P"], seed=seed)
    out, cnt = sur.leapfrog(cw["w0"], t0, dt, nsteps, start=cw["start"])
    d = (out[:, :3] - cw["plain"][:, :3]).norm(dim=1) * 1e3
    fb = float(cnt.item()) / float((nsteps - cw["start"]).sum())
    med = float(d.median())
    print(f"[surrogate bar] seed {seed}: median {med:.2f} pc, 90th percentile "
          f"{float(d.quantile(0.9)):.1f} pc, fallback {fb:.3f} of particle-steps", flush=True)
    assert med <= 10.0


def test_clumpy_stream_responds_linearly(ssc):
    """What the bar measures: the plain clumpy-halo stream is not chaotic over 4000 steps, so the
    surrogate's final-position error is the linear response to its force error.  Relative
    perturbations of every initial row by 1e-10 and 1e-8 move the final positions by medians in
    the ratio 1:100, and clump masses scaled by 1 + 1e-6 and 1 + 1e-5 in the ratio 1:10 (to 2 %)."""
    import torch
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.main import usys
    from streamsculptor_amd.potential import PlummerClumps, Potential_Combine
    from streamsculptor_amd.streams import spray_particles
    host = W.host_potential()
    pos, m, rs = W.clump_table()
    pot = Potential_Combine([host, PlummerClumps(m, rs, pos, units=usys)], units=usys)
    ts = W.release_times()
    prog = pot.integrate_orbit(w0=W.PROG_TODAY, ts=[0.0, ts[0]], t0=0.0, t1=float(ts[0]),
                               solver="leapfrog", dt0=W.DT).ys[-1]
    w0, start, _, (t0, dt, nsteps, _) = spray_particles(pot, pot, ts, prog, W.MSAT, 0,
                                                        W.particles_per_step(100_000), W.DT)
    base, _ = pot._leapfrog(w0, t0, dt, nsteps, start=start)
    g = torch.Generator(device=w0.device)
    g.manual_seed(5)
    noise = torch.randn(w0.shape, generator=g, device=w0.device, dtype=torch.float64)

    def med(out):
        return float(((out[:, :3] - base[:, :3]).norm(dim=1) * 1e3).median())

    ic = [med(pot._leapfrog(w0 * (1 + e * noise), t0, dt, nsteps, start=start)[0])
          for e in (1e-10, 1e-8)]
    cm = []
    for e in (1e-6, 1e-5):
        p2 = Potential_Combine([host, PlummerClumps(m * (1 + e), rs, pos, units=usys)],
                               units=usys)
        cm.append(med(p2._leapfrog(w0, t0, dt, nsteps, start=start)[0]))
    print(f"[linear response] initial rows 1e-10, 1e-8: {ic[0]:.3e}, {ic[1]:.3e} pc; "
          f"clump masses 1e-6, 1e-5: {cm[0]:.3e}, {cm[1]:.3e} pc", flush=True)
    assert abs(ic[1] / ic[0] / 100.0 - 1.0) < 0.02
    assert abs(cm[1] / cm[0] / 10.0 - 1.0) < 0.02
