"""Pins of the oracle's first-order response integration (CPU only).

The reference checks its perturbation output against the nonlinear stream differentiated in
the subhalo mass (streamsculptor/examples/tests.ipynb:2959,3182,2745).  Here: the response
equals the central difference in eps of the oracle's own leapfrog with the subhalo's mass
scaled by (1 + eps) -- an independent code path (direct force sums, no Hessians) -- to
O(eps^2); the radius block equals the central difference of the mass response in the scale
radius (Plummer and Hernquist subhalos; the oracle's Hernquist line source is itself pinned to
the Hernquist component); responses are linear in the initial response."""
import numpy as np
import pytest

G = 4.498502151469553e-12


def _setup(oracle, kind):
    rng = np.random.default_rng(8)
    h = oracle.milky_way_host()
    m = 5
    gm = G * 10 ** rng.uniform(7, 8, m)
    rs = 1.62 * np.sqrt(gm / G / 1e8)
    pos = rng.normal(scale=6.0, size=(m, 3))
    w0 = np.array([[10.0, 1.0, 0.5, 0.02, 0.18, 0.03], [-7.0, 6.0, -1.0, -0.12, -0.1, 0.05],
                   [3.0, -9.0, 2.0, 0.15, 0.05, -0.04]])
    if kind == "static":
        mk = lambda g, r: oracle.Sources(oracle.SRC_STATIC, gm=g, rs=r, pos=pos)
    else:
        vel = rng.normal(scale=0.1, size=(m, 3))
        tc = np.full(m, -150.0)
        tw = np.full(m, 1e4)
        sk = oracle.SRC_LINE_HERNQUIST if kind == "hernquist" else oracle.SRC_LINE
        mk = lambda g, r: oracle.Sources(sk, gm=g, rs=r, pos=pos, vel=vel, t0=tc, t_window=tw)
    return h, gm, rs, mk, w0


@pytest.mark.parametrize("kind", ["static", "line", "hernquist"])
def test_mass_response_is_derivative_of_the_map(oracle, kind):
    h, gm, rs, mk, w0 = _setup(oracle, kind)
    nsteps, t0, dt = 300, -300.0, 1.0
    base, dw = oracle.perturb_leapfrog(h, mk(gm, rs), w0, t0, dt, nsteps)
    ref0 = oracle.leapfrog(h, w0, t0, dt, nsteps)
    np.testing.assert_allclose(base, ref0, rtol=1e-13)
    for j in range(len(gm)):
        eps = 1e-4
        for sgn, store in ((1, "p"), (-1, "m")):
            g = np.zeros_like(gm)
            g[j] = sgn * eps * gm[j]
            w = oracle.leapfrog(h, w0, t0, dt, nsteps, src=mk(g, rs))
            if store == "p":
                wp = w
            else:
                wm = w
        fd = (wp - wm) / (2 * eps)
        assert np.abs(dw[:, j, :] - fd).max() <= 1e-6 * np.abs(fd).max()


@pytest.mark.parametrize("kind", ["static", "hernquist"])
def test_radius_block_is_mixed_derivative(oracle, kind):
    h, gm, rs, mk, w0 = _setup(oracle, kind)
    nsteps, t0, dt = 200, -200.0, 1.0
    _, dw = oracle.perturb_leapfrog(h, mk(gm, rs), w0, t0, dt, nsteps, with_radius=True)
    _, dm = oracle.perturb_leapfrog(h, mk(gm, rs), w0, t0, dt, nsteps)
    np.testing.assert_allclose(dw[:, :, :6], dm, rtol=1e-13, atol=1e-300)
    for j in range(len(gm)):
        eta = 1e-5 * rs[j]
        rp, rm = rs.copy(), rs.copy()
        rp[j] += eta
        rm[j] -= eta
        _, dp = oracle.perturb_leapfrog(h, mk(gm, rp), w0, t0, dt, nsteps)
        _, dmm = oracle.perturb_leapfrog(h, mk(gm, rm), w0, t0, dt, nsteps)
        fd = (dp[:, j, :] - dmm[:, j, :]) / (2 * eta)
        assert np.abs(dw[:, j, 6:] - fd).max() <= 1e-5 * np.abs(fd).max()


def test_responses_linear_in_initial_response(oracle):
    h, gm, rs, mk, w0 = _setup(oracle, "static")
    rng = np.random.default_rng(1)
    d1 = rng.normal(size=(3, 5, 6)) * 1e-3
    d2 = rng.normal(size=(3, 5, 6)) * 1e-3
    _, r0 = oracle.perturb_leapfrog(h, mk(gm, rs), w0, -100.0, 1.0, 100)
    _, r1 = oracle.perturb_leapfrog(h, mk(gm, rs), w0, -100.0, 1.0, 100, dw0=d1)
    _, r2 = oracle.perturb_leapfrog(h, mk(gm, rs), w0, -100.0, 1.0, 100, dw0=d2)
    _, r12 = oracle.perturb_leapfrog(h, mk(gm, rs), w0, -100.0, 1.0, 100, dw0=d1 + 2 * d2)
    np.testing.assert_allclose(r12 - r0, (r1 - r0) + 2 * (r2 - r0), rtol=1e-8,
                               atol=1e-12 * np.abs(r12).max())
