"""Pins of the oracle's adaptive Dormand-Prince 5(4) (restating integrate_orbit with
solver=Dopri5, reference streamsculptor/main.py:125-163).

* tableau and dense-output weights == scipy's RK45 arrays (scipy/integrate/_ivp/rk.py);
* one step and its dense output == scipy's own RK step / RkDenseOutput on the same RHS;
* converged orbits == scipy DOP853 at rtol 1e-13, and the reference notebooks' printed orbits
  (tests.ipynb:302, 1717, 1747; StreamSubhaloExample.ipynb:35);
* controller invariants: dtmin floor, max_steps failure (+inf rows), tolerance convergence,
  backward integration, forced stops.
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.integrate._ivp.rk import RK45, RkDenseOutput, rk_step

import oracle

G = oracle.G


def nfw_host():
    return oracle.host([(oracle.COMP_NFW, [G * 1e12, 20.0])])


def gala_mw_host():
    # reference GalaMilkyWayPotential (potential.py:390-414): disk, bulge, nucleus, halo
    return oracle.host([(oracle.COMP_MN, [G * 6.8e10, 3.0, 0.28]),
                        (oracle.COMP_HERNQUIST, [G * 5e9, 1.0, 0.0]),
                        (oracle.COMP_HERNQUIST, [G * 1.71e9, 0.07, 0.0]),
                        (oracle.COMP_NFW, [G * 5.4e11, 15.62])])


def rhs_of(h, src=None):
    def f(t, y):
        _, g, _ = oracle.evaluate(h, y[:3], t=t, src=src)
        return np.concatenate([y[3:], -g[0]])
    return f


def test_tableau_matches_scipy_rk45():
    A, c, b, e, m = oracle.dp5_tableau()
    assert np.array_equal(A[:6, :5], RK45.A[:6, :5])
    assert np.array_equal(c[:6], RK45.C)
    assert np.array_equal(b[:6], RK45.B)
    assert np.array_equal(-e, RK45.E)  # scipy stores bhat - b
    # Shampine's midpoint weights = scipy's dense-output polynomial at s = 1/2
    s = 0.5
    mid = RK45.P @ np.array([s, s ** 2, s ** 3, s ** 4])
    np.testing.assert_allclose(m, mid, rtol=0, atol=4e-16)
    # consistency: rows of A sum to c; b sums to 1; error weights sum to 0
    np.testing.assert_allclose(A.sum(axis=1)[:6], c[:6], atol=1e-15)
    assert abs(b.sum() - 1.0) < 1e-15 and abs(e.sum()) < 1e-16


@pytest.mark.parametrize("h_step", [0.7, 5.0, 23.0])
def test_one_step_and_dense_output_match_scipy(h_step):
    h = gala_mw_host()
    f = rhs_of(h)
    y0 = np.array([12.0, 1.0, -0.5, -0.02, 0.17, 0.05])
    t = 3.0
    K = np.empty((RK45.n_stages + 1, 6))
    y1_ref, f1_ref = rk_step(f, t, y0, f(t, y0), h_step, RK45.A, RK45.B, RK45.C, K)
    err_ref = h_step * K.T @ (-RK45.E)
    s = np.array([0.1, 0.37, 0.5, 0.8, 1.0])
    yn, err, dense = oracle.dp5_step(h, t, h_step, y0, s)
    scale = np.abs(y1_ref) + 1e-3
    assert np.max(np.abs(yn - y1_ref) / scale) < 1e-14
    np.testing.assert_allclose(err, err_ref, rtol=1e-9, atol=1e-18)
    dense_ref = RkDenseOutput(t, t + h_step, y0, K.T @ RK45.P)
    for j, sj in enumerate(s):
        ref = dense_ref(t + sj * h_step)
        assert np.max(np.abs(dense[j] - ref) / scale) < 1e-13, (sj, dense[j], ref)


def test_linear_ode_stability_function():
    """Kepler-free check: in a Plummer core (r << r_s) the force is -w^2 x (up to O(r^2/r_s^2)),
    so one step of Dopri5 on the harmonic oscillator equals R(z) y0 with
    R(z) = sum_{k<=5} z^k/k! + z^6/600 (Dormand-Prince stability polynomial)."""
    gm, rs = G * 1e12, 1e6
    h = oracle.host([(oracle.COMP_PLUMMER, [gm, rs])])
    w2 = gm / rs ** 3
    w = np.sqrt(w2)
    y0 = np.array([1e-3, 0, 0, 0, 0, 0])
    hs = 0.3 / w
    yn, _, _ = oracle.dp5_step(h, 0.0, hs, y0)
    # 1-D oscillator [x, v]' = [[0, 1], [-w2, 0]] [x, v]; R(hJ) via the polynomial
    J = np.array([[0.0, 1.0], [-w2, 0.0]]) * hs
    R = np.eye(2)
    term = np.eye(2)
    for k in range(1, 6):
        term = term @ J / k
        R = R + term
    R = R + np.linalg.matrix_power(J, 6) / 600.0
    ref = R @ np.array([y0[0], y0[3]])
    np.testing.assert_allclose([yn[0], yn[3]], ref, rtol=1e-9)


def test_converged_orbit_matches_scipy_dop853():
    h = gala_mw_host()
    w0 = np.array([14.0, -2.0, 3.0, 0.03, 0.16, -0.06])
    ts = np.linspace(0, 1500, 31)
    ref = solve_ivp(rhs_of(h), (0, 1500), w0, method="DOP853", rtol=1e-13, atol=1e-13,
                    t_eval=ts)
    _, ys, st = oracle.dopri5(h, w0, 1500.0, ts=ts, rtol=1e-13, atol=1e-13, dtmin=None,
                              max_steps=10 ** 7)
    assert st[0, 2] == 0
    scale = np.abs(ref.y.T).max(axis=0)  # kpc for positions, kpc/Myr for velocities
    assert np.max(np.abs(ys[0] - ref.y.T) / scale) < 1e-9


def test_notebook_golden_orbits():
    h = nfw_host()
    # tests.ipynb:242-243,302 -- NFW(1e12, 20), sol.evaluate(30.0)
    w0 = np.array([20.0, 15.0, 20.0, 0.08, 0.1, -0.05])
    _, ys, _ = oracle.dopri5(h, w0, 3000.0, ts=[0.0, 30.0, 3000.0], rtol=1e-12, atol=1e-12,
                             dtmin=None, max_steps=10 ** 6)
    golden = np.array([21.9793661, 17.67654451, 18.10530132, 0.051923, 0.07815599, -0.07552167])
    # numpy printed 8 decimals (trailing zeros dropped): agree to half a unit of the last one
    np.testing.assert_allclose(ys[0, 1], golden, rtol=0, atol=6e-9)
    # tests.ipynb:1691,1747 -- sum of all saved rows at ts = linspace(0, 3000, 1000)
    w0 = np.array([20.0, 0.0, 20.0, 0.0, 0.2, 0.0])
    ts = np.linspace(0, 3000, 1000)
    _, ys, _ = oracle.dopri5(h, w0, 3000.0, ts=ts, rtol=1e-12, atol=1e-12, dtmin=None,
                             max_steps=10 ** 6)
    assert abs(ys.sum() / -936.42809302 - 1) < 1e-6
    # tests.ipynb:1700,1717 -- integrate_field row 1 (t = 3000/999)
    row1 = np.array([19.9947007, 0.600547553, 19.9947007, -3.52920125e-3, 0.199947006,
                     -3.52920125e-3])
    np.testing.assert_allclose(ys[0, 1], row1, rtol=3e-9)
    # StreamSubhaloExample.ipynb:35 -- GalaMilkyWay, back 3500 Myr from (20,0,20,0,.15,0);
    # the printed value comes from a Dopri8 run at rtol = atol = 1e-7 and dtmin = 0.3, which is
    # itself ~1e-5 from converged.
    w = np.array([20.0, 0.0, 20.0, 0.0, 0.15, 0.0])
    _, ys, _ = oracle.dopri5(gala_mw_host(), w, -3500.0, t0=0.0, ts=np.linspace(0, -3500, 1000),
                             rtol=1e-12, atol=1e-12, dtmin=None, max_steps=10 ** 6)
    golden = np.array([-7.23164146, -7.96692572, -10.81840286, 0.19182623, -0.20351324,
                       -0.01770436])
    assert np.max(np.abs(ys[0, -1] - golden) / np.abs(golden)) < 1e-4
    assert np.linalg.norm(ys[0, -1, :3] - golden[:3]) / np.linalg.norm(golden[:3]) < 2e-5


def test_tolerance_convergence_and_counts():
    h = gala_mw_host()
    w0 = np.array([[10.0, 0.0, 1.0, 0.0, 0.2, 0.02], [25.0, 3.0, -4.0, 0.01, 0.11, 0.03]])
    ref, _, _ = oracle.dopri5(h, w0, 2000.0, t0=0.0, rtol=1e-13, atol=1e-13, dtmin=None,
                              max_steps=10 ** 7)
    errs, steps = [], []
    for tol in (1e-6, 1e-8, 1e-10):
        out, _, st = oracle.dopri5(h, w0, 2000.0, t0=0.0, rtol=tol, atol=tol, dtmin=None,
                                   max_steps=10 ** 7)
        errs.append(np.abs(out - ref).max())
        steps.append(st[:, 0].sum())
    assert errs[0] > 10 * errs[1] > 100 * errs[2] * 0.1
    assert steps[0] < steps[1] < steps[2]
    # 5th-order method: 100x tighter tolerance costs ~100^(1/5) = 2.5x more steps
    assert 1.8 < steps[1] / steps[0] < 3.5


def test_dtmin_floor_and_max_steps_failure():
    h = gala_mw_host()
    w0 = np.array([[8.0, 0.0, 0.1, 0.0, 0.22, 0.0]])
    ts = np.linspace(0, 1000, 2001)
    _, ys, st = oracle.dopri5(h, w0, 1000.0, ts=ts, rtol=1e-12, atol=1e-12, dtmin=0.5,
                              max_steps=10 ** 6)
    # with a floor of 0.5 Myr at most 2000 steps fit into 1000 Myr, and every step at the floor
    # is accepted whatever its error
    assert st[0, 0] <= 2001 and st[0, 2] == 0
    out, ys, st = oracle.dopri5(h, w0, 1000.0, ts=ts, rtol=1e-12, atol=1e-12, dtmin=None,
                                max_steps=50)
    assert st[0, 2] == 1 and st[0, 0] + st[0, 1] == 50
    assert np.all(np.isinf(ys)) and np.all(np.isinf(out))


def test_backward_then_forward_returns():
    h = gala_mw_host()
    w0 = np.array([12.0, 0.0, 0.0, 0.0, 0.153, 0.051])
    back, _, _ = oracle.dopri5(h, w0, -2000.0, t0=0.0, rtol=1e-11, atol=1e-11, dtmin=None,
                               max_steps=10 ** 6)
    fwd, _, _ = oracle.dopri5(h, back, 0.0, t0=-2000.0, rtol=1e-11, atol=1e-11, dtmin=None,
                              max_steps=10 ** 6)
    np.testing.assert_allclose(fwd[0], w0, rtol=0, atol=5e-7)
    # reversing velocities and integrating forward is the same map as integrating backward
    flip = back.copy()
    flip[:, 3:] *= -1
    rev, _, _ = oracle.dopri5(h, w0 * np.array([1, 1, 1, -1, -1, -1]), 2000.0, t0=0.0,
                              rtol=1e-11, atol=1e-11, dtmin=None, max_steps=10 ** 6)
    np.testing.assert_allclose(rev, flip, rtol=0, atol=1e-12)


def test_saves_at_start_and_end_and_stops():
    h = gala_mw_host()
    w0 = np.array([[12.0, 0.0, 0.0, 0.0, 0.153, 0.051], [9.0, 1.0, 0.0, -0.05, 0.2, 0.0]])
    ts = np.array([-10.0, 0.0, 333.3, 1000.0, 1200.0])
    out, ys, st = oracle.dopri5(h, w0, 1000.0, ts=ts, t0=0.0)
    assert np.all(np.isnan(ys[:, 0])) and np.all(np.isnan(ys[:, 4]))
    np.testing.assert_array_equal(ys[:, 1], w0)
    np.testing.assert_array_equal(ys[:, 3], out)
    # forcing step ends at extra times moves the result by no more than the tolerance allows
    out2, _, st2 = oracle.dopri5(h, w0, 1000.0, ts=ts, t0=0.0, stops=np.arange(50.0, 1000.0, 50))
    assert np.abs(out2 - out).max() < 1e-3 * np.abs(out).max()
    assert st2[:, 0].sum() > st[:, 0].sum()
    # per-orbit start times (spray particles released at different times)
    out3, _, _ = oracle.dopri5(h, w0, 1000.0, t_start=[0.0, 400.0])
    alone, _, _ = oracle.dopri5(h, w0[1:], 1000.0, t0=400.0)
    np.testing.assert_array_equal(out3[1], alone[0])
