"""integrate_orbit_batch_vmapped with a per-orbit ts array (N, M) (reference
streamsculptor/main.py:186-202; the scan twin :166-183): orbit i runs from min(ts[i]) to
max(ts[i]) -- or between the t0 and t1 given -- and is saved at ts[i].  Each orbit against the
oracle integrating that orbit alone: fixed-step rows to 1e-9, Dopri5 rows to 1e-6 (the same
controller; both sides take the same steps here), and the adaptive batch equals single-orbit
GPU runs bit for bit (one orbit per lane, the same arithmetic)."""
import numpy as np
import pytest
import torch

from test_gpu_core import bound_orbits

pytestmark = pytest.mark.gpu


def _ts_rows(rng, n, m, lo, hi, grid=True):
    rows = []
    for _ in range(n):
        a = rng.uniform(*lo)
        b = rng.uniform(*hi)
        r = np.sort(rng.uniform(a, b, m))
        r[0], r[-1] = a, b
        rows.append(np.rint(r) if grid else r)
    return np.array(rows)


def test_per_orbit_ts_dopri5(oracle, ssc):
    from streamsculptor_amd import workloads as W
    rng = np.random.default_rng(71)
    host = W.host_potential()
    h = oracle.host(W.host_components())
    n, m = 48, 9
    w0 = bound_orbits(rng, n)
    ts2 = _ts_rows(rng, n, m, (0.0, 300.0), (700.0, 1400.0), grid=False)
    ts2[5] = ts2[5][::-1].copy()  # a row given in descending order: saved rows follow it
    sol = host.integrate_orbit_batch_vmapped(w0=w0, ts=ts2, solver=ssc.Dopri5())
    ys = sol.ys.cpu().numpy()
    assert ys.shape == (n, m, 6)
    worst = 0.0
    for i in range(n):
        order = np.argsort(ts2[i])
        _, yo, st = oracle.dopri5(h, w0[i], ts2[i].max(), t0=ts2[i].min(), ts=ts2[i][order])
        got = ys[i][order]
        e = np.linalg.norm(got - yo[0], axis=1) / np.linalg.norm(yo[0], axis=1)
        worst = max(worst, e.max())
        single = host.integrate_orbit(w0=w0[i], ts=ts2[i], solver=ssc.Dopri5()).ys.cpu().numpy()
        np.testing.assert_array_equal(ys[i], single)
    print(f"[per-orbit ts] dopri5 worst row vs oracle {worst:.2e}")
    assert worst < 1e-6
    # shared t0 / t1 given: every orbit integrates backwards from t0 = 0 to t1 = -900
    tsb = -_ts_rows(rng, n, m, (0.0, 1.0), (899.0, 900.0), grid=False)
    sol = host.integrate_orbit_batch_vmapped(w0=w0, ts=tsb, t0=0.0, t1=-900.0,
                                             solver=ssc.Dopri5())
    for i in (0, 17, 47):
        _, yo, _ = oracle.dopri5(h, w0[i], -900.0, t0=0.0, ts=tsb[i])
        e = np.linalg.norm(sol.ys[i].cpu().numpy() - yo[0], axis=1) / np.linalg.norm(yo[0], axis=1)
        assert e.max() < 1e-6


def test_per_orbit_ts_leapfrog(oracle, ssc):
    from streamsculptor_amd import workloads as W
    rng = np.random.default_rng(72)
    host = W.host_potential()
    h = oracle.host(W.host_components())
    n, m = 300, 6
    w0 = bound_orbits(rng, n)
    ts2 = _ts_rows(rng, n, m, (-1000.0, -400.0), (-300.0, 0.0))
    sol = host.integrate_orbit_batch_vmapped(w0=w0, ts=ts2, solver="leapfrog", dt0=1.0)
    ys = sol.ys.cpu().numpy()
    assert ys.shape == (n, m, 6)
    worst = 0.0
    for i in range(0, n, 7):
        t0 = ts2[i].min()
        k = np.rint(ts2[i] - t0).astype(np.int64)
        _, tr = oracle.leapfrog(h, w0[i], t0, 1.0, int(k.max()), traj=True)
        ref = tr[k, 0]
        e = np.linalg.norm(ys[i] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        worst = max(worst, e.max())
    print(f"[per-orbit ts] leapfrog worst row vs oracle {worst:.2e}")
    assert worst < 1e-9
    off = ts2.copy()
    off[3, 2] += 0.5  # one save time off the dt0 grid of the others
    with pytest.raises(ValueError):
        host.integrate_orbit_batch_vmapped(w0=w0, ts=off, solver="leapfrog", dt0=1.0)
