"""Pins of the oracle's adaptive (Dopri5) first-order response integration (CPU only).

* Step control on the whole state: with every source inactive and zero initial responses the
  responses stay zero, so the RMS error norm over the 6 + m D components is the base row's norm
  scaled by sqrt(6 / (6 + m D)).  The run must then take exactly the steps of the pinned orbit
  oracle with both tolerances divided by that factor (same accepted/rejected counts, same rows
  and saves).  A norm over the base row alone, or a sum instead of a mean, fails this.
* The ODE: at tight tolerances the mass block equals the central difference in eps of the
  pinned nonlinear Dopri5 orbit with the subhalo's potential scaled by eps (the reference's own
  test method, streamsculptor/examples/tests.ipynb:2959,3182), forwards and backwards in time,
  Plummer and Hernquist subhalos; the radius block equals the central difference of the mass
  block in the scale radius.
* Saves: rows saved at interior times equal separate integrations ending at those times.
"""
import numpy as np
import pytest

G = 4.498502151469553e-12


def _sources(oracle, kind, gm, rs, pos, vel, tc, tw):
    sk = oracle.SRC_LINE_HERNQUIST if kind == "hernquist" else oracle.SRC_LINE
    return oracle.Sources(sk, gm=gm, rs=rs, pos=pos, vel=vel, t0=tc, t_window=tw)


def _setup(seed=3, m=4):
    rng = np.random.default_rng(seed)
    gm = G * 10 ** rng.uniform(7, 8, m)
    rs = 1.62 * np.sqrt(gm / G / 1e8)
    pos = rng.normal(scale=6.0, size=(m, 3))
    vel = rng.normal(scale=0.1, size=(m, 3))
    tc = rng.uniform(-150.0, 150.0, m)
    tw = np.full(m, 1e4)
    w0 = np.array([[10.0, 1.0, 0.5, 0.02, 0.18, 0.03], [-7.0, 6.0, -1.0, -0.12, -0.1, 0.05]])
    return gm, rs, pos, vel, tc, tw, w0


@pytest.mark.parametrize("with_radius", [False, True])
def test_inactive_sources_follow_the_orbit_controller(oracle, with_radius):
    h = oracle.milky_way_host()
    gm, rs, pos, vel, tc, _, w0 = _setup()
    m = len(gm)
    src = _sources(oracle, "line", gm, rs, pos, vel, tc + 1e5, np.full(m, 1.0))  # never active
    N = 6 + m * (12 if with_radius else 6)
    k = np.sqrt(6.0 / N)
    ts = np.linspace(0.0, 800.0, 7)
    tol = dict(rtol=1e-7, atol=1e-7, dtmin=0.05)
    w, dw, ys, dys, st = oracle.perturb_dopri5(h, src, w0, 800.0, ts=ts, t0=0.0,
                                               with_radius=with_radius, **tol)
    wo, yso, sto = oracle.dopri5(h, w0, 800.0, ts=ts, t0=0.0, rtol=1e-7 / k, atol=1e-7 / k,
                                 dtmin=0.05)
    np.testing.assert_array_equal(st, sto)
    np.testing.assert_allclose(w, wo, rtol=1e-12, atol=0)
    np.testing.assert_allclose(ys, yso, rtol=1e-12, atol=0)
    assert not dw.any() and not dys.any()
    # the norm over the base row alone takes other steps
    wb, _, stb = oracle.dopri5(h, w0, 800.0, t0=0.0, **tol)
    assert (stb[:, 0] != st[:, 0]).any()


@pytest.mark.parametrize("kind,t1", [("line", 400.0), ("line", -400.0), ("hernquist", 400.0)])
def test_mass_block_is_derivative_of_the_flow(oracle, kind, t1):
    h = oracle.milky_way_host()
    gm, rs, pos, vel, tc, tw, w0 = _setup()
    tol = dict(rtol=1e-12, atol=1e-12, dtmin=None, max_steps=200_000)
    _, dw, _, _, st = oracle.perturb_dopri5(h, _sources(oracle, kind, gm, rs, pos, vel, tc, tw),
                                            w0, t1, t0=0.0, **tol)
    assert (st[:, 2] == 0).all()
    eps = 1e-4
    for j in range(len(gm)):
        out = []
        for sgn in (1, -1):
            g = np.zeros_like(gm)
            g[j] = sgn * eps * gm[j]
            wj, _, _ = oracle.dopri5(h, w0, t1, t0=0.0,
                                     src=_sources(oracle, kind, g, rs, pos, vel, tc, tw), **tol)
            out.append(wj)
        fd = (out[0] - out[1]) / (2 * eps)
        assert np.abs(dw[:, j, :] - fd).max() <= 1e-6 * np.abs(fd).max()


@pytest.mark.parametrize("kind", ["line", "hernquist"])
def test_radius_block_is_mixed_derivative(oracle, kind):
    h = oracle.milky_way_host()
    gm, rs, pos, vel, tc, tw, w0 = _setup()
    tol = dict(rtol=1e-12, atol=1e-12, dtmin=None, max_steps=200_000)
    _, dw, _, _, _ = oracle.perturb_dopri5(h, _sources(oracle, kind, gm, rs, pos, vel, tc, tw),
                                           w0, -300.0, t0=0.0, with_radius=True, **tol)
    for j in range(len(gm)):
        eta = 1e-4 * rs[j]
        res = []
        for sgn in (1, -1):
            r = rs.copy()
            r[j] += sgn * eta
            _, d, _, _, _ = oracle.perturb_dopri5(
                h, _sources(oracle, kind, gm, r, pos, vel, tc, tw), w0, -300.0, t0=0.0, **tol)
            res.append(d[:, j, :])
        fd = (res[0] - res[1]) / (2 * eta)
        assert np.abs(dw[:, j, 6:] - fd).max() <= 1e-5 * np.abs(fd).max()


def test_saves_match_integrations_to_each_time(oracle):
    h = oracle.milky_way_host()
    gm, rs, pos, vel, tc, tw, w0 = _setup()
    src = _sources(oracle, "line", gm, rs, pos, vel, tc, tw)
    tol = dict(rtol=1e-11, atol=1e-11, dtmin=None, max_steps=200_000)
    ts = np.array([0.0, -37.5, -120.0, -300.0])
    rng = np.random.default_rng(5)
    d0 = rng.normal(scale=1e-2, size=(2, len(gm), 12))
    _, _, ys, dys, _ = oracle.perturb_dopri5(h, src, w0, -300.0, ts=ts, t0=0.0, dw0=d0,
                                             with_radius=True, **tol)
    np.testing.assert_array_equal(ys[:, 0], w0)
    np.testing.assert_array_equal(dys[:, 0], d0)
    for jt, tj in enumerate(ts[1:], start=1):
        wj, dj, _, _, _ = oracle.perturb_dopri5(h, src, w0, tj, t0=0.0, dw0=d0, with_radius=True,
                                                **tol)
        assert np.abs(ys[:, jt] - wj).max() <= 1e-8 * np.abs(wj).max()
        assert np.abs(dys[:, jt] - dj).max() <= 1e-7 * np.abs(dj).max()
