"""Sizes at and beyond the documented limits: the largest source table a call takes (1024
Plummer sources, STATIC) against the oracle, one source more rejected with a clear message;
the adjoint's row-range checks; empty release and Dopri5 calls."""
import numpy as np
import pytest
import torch

from test_gpu_core import bound_orbits

pytestmark = pytest.mark.gpu

G = 4.498502151469553e-12


def _clumps(rng, m):
    pos = rng.normal(scale=15.0, size=(m, 3))
    mass = 10 ** rng.uniform(5, 7, m)
    return pos, mass, 1.62 * np.sqrt(mass / 1e8)


def test_largest_source_table_and_one_more(oracle, ssc):
    from streamsculptor_amd import _lib
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.main import usys
    from streamsculptor_amd.potential import PlummerClumps, Potential_Combine
    rng = np.random.default_rng(1024)
    host = W.host_potential()
    h = oracle.host(W.host_components())
    pos, mass, rs = _clumps(rng, _lib.MAX_SOURCES)
    pot = Potential_Combine([host, PlummerClumps(mass, rs, pos, units=usys)], units=usys)
    n, nsteps = 300, 200
    w0 = bound_orbits(rng, n)
    start = np.sort(rng.integers(0, nsteps, n)).astype(np.int32)
    out, _ = pot._leapfrog(w0, -200.0, 1.0, nsteps, start=start)
    so = oracle.Sources(oracle.SRC_STATIC, gm=G * mass, rs=rs, pos=pos)
    ref = oracle.leapfrog(h, w0, -200.0, 1.0, nsteps, start=start, src=so)
    got = out.cpu().numpy()
    e = (np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)).max()
    print(f"[limits] 1024 sources: leapfrog vs oracle {e:.2e}")
    assert e < 1e-9
    pos2, mass2, rs2 = _clumps(rng, _lib.MAX_SOURCES + 1)
    pot2 = Potential_Combine([host, PlummerClumps(mass2, rs2, pos2, units=usys)], units=usys)
    with pytest.raises(ValueError, match="at most 1024 sources"):
        pot2._leapfrog(w0, -200.0, 1.0, nsteps)


def test_adjoint_range_checks(ssc):
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.adjoint import LeapfrogAdjoint
    rng = np.random.default_rng(7)
    host = W.host_potential()
    w0 = bound_orbits(rng, 64)
    adj = LeapfrogAdjoint(host, w0, 0.0, 1.0, 100, checkpoint_every=25)
    adj.forward()
    lam = torch.as_tensor(rng.normal(size=(64, 6)), device="cuda")
    for bad in ((10, 100), (0, 60), (50, 50), (50, 101), (-25, 25)):
        with pytest.raises(RuntimeError, match="k_lo|bad argument"):
            adj.backward(lam, rows=bad)
    full = adj.backward(lam)
    adj.backward(lam, rows=(50, 100))
    part = adj.backward(lam, rows=(0, 50))
    assert torch.equal(part["lam0"], full["lam0"])


def test_empty_calls(ssc):
    from streamsculptor_amd import workloads as W
    host = W.host_potential()
    empty = torch.empty(0, 6, dtype=torch.float64, device="cuda")
    outs = host.release_model(x=empty[:, :3], v=empty[:, 3:], Msat=1e4, i=np.arange(0),
                              t=np.zeros(0), seed_num=0)
    assert all(o.shape[0] == 0 for o in outs)
    sol = host.integrate_orbit_batch_vmapped(w0=empty, ts=np.array([0.0, 10.0]),
                                             solver=ssc.Dopri5())
    assert sol.ys.shape[0] == 0
