"""Pins of the oracle's potentials (CPU only).

Checks that do not reuse the oracle's formulas: sympy differentiation of the reference's
potential expressions (reference streamsculptor/potential.py:66-84,124-138), Poisson's equation
against the textbook densities, central finite differences, and the reference notebooks'
recorded values.
"""
import math

import numpy as np
import pytest
import sympy as sp

G = 4.498502151469553e-12


def test_G_from_constants(oracle):
    # CODATA 2018 G with the IAU 2015 nominal solar GM, IAU au/pc, Julian year (astropy 7).
    Gsi, GM = 6.6743e-11, 1.3271244e20
    pc = 149597870700.0 * 648000 / math.pi
    val = Gsi * (GM / Gsi) * (1e6 * 365.25 * 86400) ** 2 / (1e3 * pc) ** 3
    assert val == pytest.approx(G, rel=1e-15)
    assert oracle.G == G


def test_G_reproduces_custom_potential_notebook():
    # reference streamsculptor/examples/custom_potential.ipynb:132 prints
    # "-0.186204177133592 -0.19132325459167931" for its oscillating NFW at (1, 20, 10).
    om = 2 * np.pi / 500.0

    def osc(xyz, t):
        x, y, z = xyz
        z = z / (1 + 0.3 * np.cos(om * t))
        r = np.sqrt(x * x + y * y + z * z)
        m = 1e12 + 5e11 * np.sin(om * t)
        rs = 15.0 + 10.0 * np.sin(om * t)
        return -(G * m / r) * np.log(1 + r / rs)

    assert osc((1.0, 20.0, 10.0), -1000.0) == pytest.approx(-0.186204177133592, rel=1e-14)
    assert osc((1.0, 20.0, 10.0), -850.0) == pytest.approx(-0.19132325459167931, rel=1e-14)


def test_nfw_acceleration_notebook_value(oracle):
    # reference streamsculptor/examples/tests.ipynb:1671 (NFW m=1e12, r_s=20 at (1,2,3))
    h = oracle.host([(oracle.COMP_NFW, [G * 1e12, 20.0])])
    _, g, _ = oracle.evaluate(h, [1.0, 2.0, 3.0])
    np.testing.assert_allclose(-g[0], [-0.0011937, -0.00238739, -0.00358109], rtol=5e-6)


X, Y, Z = sp.symbols("x y z", real=True)


def _sym_phi(kind, p):
    r2 = X ** 2 + Y ** 2 + Z ** 2
    if kind == 1:  # MN, potential.py:69-72
        gm, a, b = p[:3]
        return -gm / sp.sqrt(X ** 2 + Y ** 2 + (sp.sqrt(Z ** 2 + b ** 2) + a) ** 2)
    if kind == 2:  # NFW, potential.py:81-84
        gm, rs = p[:2]
        m = sp.sqrt(r2) / rs
        return (-gm / rs) * sp.log(1 + m) / m
    if kind == 3:  # Hernquist, potential.py:135-138
        gm, c, soft = p[:3]
        return -gm / (sp.sqrt(r2 + soft) + c)
    if kind == 4:  # Plummer, potential.py:127-130
        gm, rs = p[:2]
        return -gm / sp.sqrt(r2 + rs ** 2)


CASES = [
    (1, [G * 6.8e10, 3.0, 0.28]),
    (1, [G * 1e10, 0.5, 2.0]),
    (2, [G * 5.4e11, 15.62]),
    (2, [G * 1e12, 20.0]),
    (3, [G * 5e9, 1.0, 0.0]),
    (3, [G * 1.71e9, 0.07, 0.01]),
    (4, [G * 1e4, 0.01]),
    (4, [G * 1e8, 1.62]),
]
POINTS = np.array([[8.0, 0.5, 0.3], [-3.0, 12.0, -4.0], [0.2, -0.1, 0.05], [25.0, -30.0, 12.0],
                   [1.0, 2.0, 3.0]])


@pytest.mark.parametrize("kind,p", CASES)
def test_components_against_sympy(oracle, kind, p):
    phi_s = _sym_phi(kind, [sp.Float(v, 30) for v in p])
    grad_s = [sp.diff(phi_s, v) for v in (X, Y, Z)]
    hess_s = [[sp.diff(g, v) for v in (X, Y, Z)] for g in grad_s]
    f_phi = sp.lambdify((X, Y, Z), phi_s, "mpmath")
    f_g = sp.lambdify((X, Y, Z), grad_s, "mpmath")
    f_h = sp.lambdify((X, Y, Z), hess_s, "mpmath")
    import mpmath
    mpmath.mp.dps = 30
    h = oracle.host([(kind, p)])
    phi, g, H = oracle.evaluate(h, POINTS)
    for i, x in enumerate(POINTS):
        a = [mpmath.mpf(float(c)) for c in x]
        assert phi[i] == pytest.approx(float(f_phi(*a)), rel=1e-13)
        np.testing.assert_allclose(g[i], np.array(f_g(*a), dtype=float), rtol=1e-12,
                                   atol=1e-14 * np.abs(g[i]).max())
        Hs = np.array([[float(e) for e in row] for row in f_h(*a)])
        np.testing.assert_allclose(H[i], Hs, rtol=1e-11, atol=1e-12 * np.abs(Hs).max())


def _density(kind, p, x):
    """Textbook densities (Binney & Tremaine) in units where the mass is G*m / G."""
    gm = p[0]
    r = np.linalg.norm(x)
    if kind == 1:
        a, b = p[1], p[2]
        R2, z = x[0] ** 2 + x[1] ** 2, x[2]
        zeta = np.sqrt(z * z + b * b)
        return (b * b * gm / (4 * np.pi)) * (a * R2 + (a + 3 * zeta) * (a + zeta) ** 2) / (
            (R2 + (a + zeta) ** 2) ** 2.5 * zeta ** 3)
    if kind == 2:
        rs = p[1]
        u = r / rs
        return gm / (4 * np.pi * rs ** 3) / (u * (1 + u) ** 2)
    if kind == 3:
        c = p[1]
        return gm * c / (2 * np.pi * r * (r + c) ** 3)
    if kind == 4:
        a = p[1]
        return 3 * gm / (4 * np.pi * a ** 3) * (1 + r * r / a ** 2) ** -2.5


@pytest.mark.parametrize("kind,p", [c for c in CASES if not (c[0] == 3 and c[1][2] != 0)])
def test_poisson(oracle, kind, p):
    h = oracle.host([(kind, p)])
    _, _, H = oracle.evaluate(h, POINTS)
    for i, x in enumerate(POINTS):
        lap = np.trace(H[i])
        assert lap == pytest.approx(4 * np.pi * _density(kind, p, x), rel=1e-9)


def test_finite_differences_and_combine(oracle):
    h = oracle.milky_way_host()
    rng = np.random.default_rng(3)
    pts = rng.normal(scale=10.0, size=(6, 3))
    phi, g, H = oracle.evaluate(h, pts)
    eps = 1e-5
    for k in range(3):
        d = np.zeros(3)
        d[k] = eps
        pp, gp, _ = oracle.evaluate(h, pts + d)
        pm, gm_, _ = oracle.evaluate(h, pts - d)
        np.testing.assert_allclose((pp - pm) / (2 * eps), g[:, k], rtol=1e-7)
        np.testing.assert_allclose((gp - gm_) / (2 * eps), H[:, :, k], rtol=1e-6,
                                   atol=1e-9 * np.abs(H).max())
    # Combine = sum of parts (reference potential.py:1285-1296)
    tot = np.zeros_like(g)
    for comp in [(1, [G * 6.8e10, 3.0, 0.28]), (3, [G * 5e9, 1.0, 0.0]), (2, [G * 5.4e11, 15.62])]:
        tot += oracle.evaluate(oracle.host([comp]), pts)[1]
    np.testing.assert_allclose(tot, g, rtol=1e-14)


def test_sources_superposition_and_symmetry(oracle):
    rng = np.random.default_rng(5)
    n = 7
    gm = G * 10 ** rng.uniform(6, 8, n)
    rs = 1.62 * np.sqrt(gm / G / 1e8)
    pos = rng.normal(scale=5, size=(n, 3))
    src = oracle.Sources(oracle.SRC_STATIC, gm=gm, rs=rs, pos=pos)
    h0 = oracle.host([])
    pts = rng.normal(scale=5, size=(4, 3))
    phi, g, H = oracle.evaluate(h0, pts, src=src)
    ref = np.zeros_like(g)
    for j in range(n):
        ref += oracle.evaluate(oracle.host([(4, [gm[j], rs[j]])]), pts - pos[j])[1]
    np.testing.assert_allclose(g, ref, rtol=1e-13)
    # a single source on the z axis pulls along z only
    one = oracle.Sources(oracle.SRC_STATIC, gm=[G * 1e8], rs=[0.5], pos=[[0, 0, 2.0]])
    _, g1, _ = oracle.evaluate(h0, [[0, 0, -1.0]], src=one)
    assert g1[0, 0] == 0 and g1[0, 1] == 0 and g1[0, 2] < 0
    # LINE: inactive outside the window, equal to STATIC at the line position inside it
    line = oracle.Sources(oracle.SRC_LINE, gm=gm, rs=rs, pos=pos, vel=np.ones((n, 3)) * 0.1,
                          t0=np.zeros(n), t_window=np.full(n, 50.0))
    _, gl, _ = oracle.evaluate(h0, pts, t=10.0, src=line)
    moved = oracle.Sources(oracle.SRC_STATIC, gm=gm, rs=rs, pos=pos + 1.0)
    np.testing.assert_allclose(gl, oracle.evaluate(h0, pts, src=moved)[1], rtol=1e-13)
    _, gout, _ = oracle.evaluate(h0, pts, t=60.0, src=line)
    assert np.all(gout == 0)


def test_hernquist_line_source_equals_shifted_component(oracle):
    """A LINE_HERNQUIST source at rest inside its window is a Hernquist component centred on it
    (the component's Phi, gradient and Hessian are pinned above); outside the window it is
    absent; a moving one sits at x0 + v (t - t0)."""
    import numpy as np
    rng = np.random.default_rng(21)
    G = 4.498502151469553e-12
    gm, a = G * 3e8, 0.7
    x0 = np.array([1.5, -2.0, 0.5])
    v = np.array([0.1, 0.05, -0.2])
    src = oracle.Sources(oracle.SRC_LINE_HERNQUIST, gm=np.array([gm]), rs=np.array([a]),
                         pos=x0[None], vel=v[None], t0=np.array([10.0]),
                         t_window=np.array([50.0]))
    empty = oracle.host([])
    comp = oracle.host([(oracle.COMP_HERNQUIST, [gm, a, 0.0])])
    x = rng.normal(scale=4.0, size=(40, 3))
    for t in (10.0, 35.0):
        c = x0 + v * (t - 10.0)
        p1, g1, H1 = oracle.evaluate(empty, x, t=t, src=src)
        p2, g2, H2 = oracle.evaluate(comp, x - c)
        np.testing.assert_allclose(p1, p2, rtol=1e-14)
        np.testing.assert_allclose(g1, g2, rtol=1e-13)
        np.testing.assert_allclose(H1, H2, rtol=1e-12, atol=1e-30)
    p3, g3, _ = oracle.evaluate(empty, x, t=70.0, src=src)
    assert np.all(p3 == 0.0) and np.all(g3 == 0.0)
