"""Spheres centred on a moving track (TRACK and TRACK_HERNQUIST sources: the kernel form of
the reference's ProgenitorPotential and TimeDepTranslatingPotential, potential.py:140-153,
448-462).  At grid row k a track source equals the analytic component of its profile -- whose
closed forms test_oracle_potentials.py pins by sympy differentiation -- evaluated at x - c_k;
a leapfrog in host + track equals the leapfrog in host + a static source at the same place
when the track does not move."""
import numpy as np
import pytest

G = 4.498502151469553e-12


@pytest.mark.parametrize("profile", ["plummer", "hernquist"])
def test_track_source_is_the_shifted_component(oracle, profile):
    rng = np.random.default_rng(5)
    gm, rs = G * 3e10, 2.5
    rows = rng.normal(scale=5.0, size=(7, 1, 3))
    kind = oracle.SRC_TRACK if profile == "plummer" else oracle.SRC_TRACK_HERNQUIST
    comp = oracle.COMP_PLUMMER if profile == "plummer" else oracle.COMP_HERNQUIST
    src = oracle.Sources(kind, gm=[gm], rs=[rs], pos=rows)
    none = oracle.host([])
    sphere = oracle.host([(comp, [gm, rs])])
    x = rng.normal(scale=8.0, size=(50, 3))
    for k in (0, 3, 6):
        phi, g, H = oracle.evaluate(none, x, step=k, src=src)
        phi0, g0, H0 = oracle.evaluate(sphere, x - rows[k, 0])
        np.testing.assert_allclose(phi, phi0, rtol=1e-14)
        np.testing.assert_allclose(g, g0, rtol=1e-13, atol=1e-16 * np.abs(g0).max())
        np.testing.assert_allclose(H, H0, rtol=1e-12, atol=1e-15 * np.abs(H0).max())


def test_resting_track_equals_static_source(oracle):
    h = oracle.milky_way_host()
    rng = np.random.default_rng(6)
    c = np.array([[3.0, -2.0, 1.0]])
    gm, rs = G * np.array([5e9]), np.array([1.5])
    track = oracle.Sources(oracle.SRC_TRACK, gm=gm, rs=rs, pos=np.repeat(c[None], 201, axis=0))
    static = oracle.Sources(oracle.SRC_STATIC, gm=gm, rs=rs, pos=c)
    w0 = np.hstack([rng.normal(scale=8.0, size=(20, 3)), rng.normal(scale=0.15, size=(20, 3))])
    a = oracle.leapfrog(h, w0, 0.0, 1.0, 200, src=track)
    b = oracle.leapfrog(h, w0, 0.0, 1.0, 200, src=static)
    np.testing.assert_array_equal(a, b)
