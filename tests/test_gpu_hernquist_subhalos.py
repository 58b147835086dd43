"""Straight-line Hernquist subhalos (SubhaloLinePotentialCustom_fromFunc with
HernquistPotential, reference streamsculptor/potential.py:1161-1213; source kind
LINE_HERNQUIST) against the oracle, whose Hernquist line source is pinned to the Hernquist
component in test_oracle_potentials.py: evaluation (Phi, gradient, Hessian), fixed-step
leapfrog at 1e-9 with ragged starts on every launch path, adaptive Dopri5 at 1e-6 single and
as a grouped batch of simulations.  Their first-order responses are covered in
test_gpu_perturb.py and test_gpu_perturbative.py."""
import numpy as np
import pytest

from test_gpu_core import bound_orbits, rel_err_rows

pytestmark = pytest.mark.gpu

G = 4.498502151469553e-12


def _subhalos(rng, m, t_lo, t_hi):
    mass = 10 ** rng.uniform(6.5, 8.5, m)
    a = 1.05 * np.sqrt(mass / 1e8)
    pos = rng.normal(scale=9.0, size=(m, 3))
    vel = rng.normal(scale=0.2, size=(m, 3))
    tc = rng.uniform(t_lo, t_hi, m)
    tw = rng.uniform(50, 400, m)
    return mass, a, pos, vel, tc, tw


def _pair(oracle, ssc, rng, m, t_lo, t_hi):
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.potential import (HernquistPotential, Potential_Combine,
                                              SubhaloLinePotentialCustom_fromFunc)
    mass, a, pos, vel, tc, tw = _subhalos(rng, m, t_lo, t_hi)
    sub = SubhaloLinePotentialCustom_fromFunc(HernquistPotential, mass, a, pos, vel, tc, tw,
                                              units=ssc.usys)
    pot = Potential_Combine([W.host_potential(), sub], units=ssc.usys)
    h = oracle.host(W.host_components())
    src = oracle.Sources(oracle.SRC_LINE_HERNQUIST, gm=G * mass, rs=a, pos=pos, vel=vel, t0=tc,
                         t_window=tw)
    return pot, sub, h, src


def test_eval_parity(oracle, ssc):
    rng = np.random.default_rng(31)
    pot, sub, h, src = _pair(oracle, ssc, rng, 48, -100.0, 100.0)
    pts = rng.normal(scale=10.0, size=(800, 3))
    for t in (0.0, 60.0):
        phi_o, g_o, H_o = oracle.evaluate(h, pts, t=t, src=src)
        phi = pot.potential(pts, t).cpu().numpy()
        g = pot.gradient(pts, t).cpu().numpy()
        H = pot.jacobian_force(pts, t).cpu().numpy()
        np.testing.assert_allclose(phi, phi_o, rtol=1e-12)
        assert np.abs(g - g_o).max() <= 1e-12 * np.abs(g_o).max()
        assert np.abs(H - H_o).max() <= 1e-11 * np.abs(H_o).max()
    # per-subhalo potentials sum to the subhalo potential
    per = sub.potential_per_SH(pts[:5], 0.0).cpu().numpy()
    tot = sub.potential(pts[:5], 0.0).cpu().numpy()
    np.testing.assert_allclose(per.sum(-1), tot, rtol=1e-12)


@pytest.mark.parametrize("n", [1, 700, 30000])
def test_leapfrog_parity(oracle, ssc, n):
    """n = 1: one wave per orbit; 700: lanes per orbit split the subhalo sum; 30000: one
    particle per lane."""
    rng = np.random.default_rng(32 + n)
    nsteps, t0, dt = 800, -800.0, 1.0
    pot, _, h, src = _pair(oracle, ssc, rng, 32, t0, 0.0)
    w0 = bound_orbits(rng, n)
    start = np.sort(rng.integers(0, nsteps + 5, n)).astype(np.int32)
    print(f"[hernquist subhalos] leapfrog n={n}", flush=True)
    ref = oracle.leapfrog(h, w0, t0, dt, nsteps, start=start, src=src)
    out, _ = pot._leapfrog(w0, t0, dt, nsteps, start=start)
    assert rel_err_rows(out, ref) <= 1e-9


def test_dopri5_single_and_batch(oracle, ssc):
    from streamsculptor_amd import adaptive
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.batch import SubhaloLineBatch, integrate_orbit_batch
    rng = np.random.default_rng(33)
    pot, _, h, src = _pair(oracle, ssc, rng, 24, 200.0, 1200.0)
    w0 = bound_orbits(rng, 300)
    t_start = np.sort(rng.uniform(0.0, 1000.0, 300))
    out, _, st = adaptive.dopri5(pot, w0, 1500.0, t_start=t_start)
    out_o, _, st_o = oracle.dopri5(h, w0, 1500.0, t_start=t_start, src=src)
    # orbits on the oracle's step sequence agree to 1e-6; a step that straddles a subhalo's
    # window edge (the force jumps there) has an error estimate that can land within rounding
    # of the threshold, and such an orbit continues on another valid sequence
    same = np.all(st.cpu().numpy()[:, :2] == st_o[:, :2], axis=1)
    out = out.cpu().numpy()
    err = np.linalg.norm(out - out_o, axis=1) / np.linalg.norm(out_o, axis=1)
    print(f"[hernquist dopri5] {same.sum()} of {len(same)} orbits on the oracle's steps, max "
          f"{err[same].max():.2e}; others max {err[~same].max() if (~same).any() else 0:.2e}")
    assert err[same].max() < 1e-6
    assert (~same).sum() <= 3 and ((~same).sum() == 0 or err[~same].max() < 1e-3)
    # batch: 6 simulations with 3 Hernquist subhalos each, one orbit per simulation
    B, nsh = 6, 3
    mass, a, pos, vel, tc, tw = _subhalos(rng, B * nsh, 100.0, 900.0)
    batch = SubhaloLineBatch(mass.reshape(B, nsh), a.reshape(B, nsh), pos.reshape(B, nsh, 3),
                             vel.reshape(B, nsh, 3), tc.reshape(B, nsh), tw.reshape(B, nsh),
                             units=ssc.usys, profile="hernquist")
    wb = bound_orbits(rng, 1)[0]
    ys = integrate_orbit_batch(W.host_potential(), batch, wb, ts=[1000.0], t0=0.0, t1=1000.0)
    for b in range(B):
        sb = oracle.Sources(oracle.SRC_LINE_HERNQUIST, gm=G * mass[b * nsh:(b + 1) * nsh],
                            rs=a[b * nsh:(b + 1) * nsh], pos=pos[b * nsh:(b + 1) * nsh],
                            vel=vel[b * nsh:(b + 1) * nsh], t0=tc[b * nsh:(b + 1) * nsh],
                            t_window=tw[b * nsh:(b + 1) * nsh])
        _, yo, _ = oracle.dopri5(h, wb, 1000.0, t0=0.0, ts=[1000.0], src=sb)
        assert rel_err_rows(ys[b, 0].cpu().numpy()[None], yo[0]) < 1e-6
