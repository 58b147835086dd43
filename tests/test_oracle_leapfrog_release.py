"""Pins of the oracle's leapfrog, PRNG and release model (CPU only)."""
import numpy as np
import pytest
from scipy.special import erfinv as sp_erfinv

G = 4.498502151469553e-12


# ------------------------------------------------------------------------------ leapfrog
def test_leapfrog_discrete_harmonic_oscillator(oracle):
    """Inside a Plummer core (r << r_s) the force is -omega^2 x; KDK is then the linear map
    M with cos(theta) = 1 - h^2 omega^2 / 2 and M^n = cos(n theta) I + sin(n theta)/sin(theta)
    (M - cos(theta) I) (Hairer, Lubich & Wanner, GNI sec. I.1)."""
    gm, rs = G * 1e12, 1.0e4
    h = oracle.host([(oracle.COMP_PLUMMER, [gm, rs])])
    om2 = gm / rs ** 3
    dt, n = 2.0e4, 500
    w0 = np.array([[0.05, -0.02, 0.03, 1e-6, 2e-6, -1e-6]])
    out = oracle.leapfrog(h, w0, 0.0, dt, n)
    c = 1 - dt * dt * om2 / 2
    M = np.array([[c, dt], [-dt * om2 * (1 - dt * dt * om2 / 4), c]])
    th = np.arccos(c)
    Mn = np.cos(n * th) * np.eye(2) + np.sin(n * th) / np.sin(th) * (M - c * np.eye(2))
    for k in range(3):
        exp = Mn @ np.array([w0[0, k], w0[0, 3 + k]])
        np.testing.assert_allclose([out[0, k], out[0, 3 + k]], exp, rtol=2e-7, atol=1e-12)


def test_leapfrog_time_reversible(oracle):
    h = oracle.milky_way_host()
    w0 = np.array([[12.0, 0.0, 0.0, 0.0, 0.153, 0.051], [8.0, 3.0, -1.0, -0.05, 0.2, 0.03]])
    fwd = oracle.leapfrog(h, w0, 0.0, 1.0, 2000)
    back = oracle.leapfrog(h, fwd, 2000.0, -1.0, 2000)
    np.testing.assert_allclose(back, w0, rtol=0, atol=1e-10)


def test_leapfrog_second_order_and_energy(oracle):
    h = oracle.milky_way_host()
    w0 = np.array([[12.0, 0.0, 0.0, 0.0, 0.153, 0.051]])
    T = 1000.0
    ref = oracle.leapfrog(h, w0, 0.0, T / 64000, 64000)
    errs = []
    for n in (1000, 2000, 4000):
        errs.append(np.abs(oracle.leapfrog(h, w0, 0.0, T / n, n) - ref).max())
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5
    _, tr = oracle.leapfrog(h, w0, 0.0, 1.0, 4000, traj=True)
    phi, _, _ = oracle.evaluate(h, tr[:, 0, :3])
    E = phi + 0.5 * (tr[:, 0, 3:] ** 2).sum(1)
    assert np.abs(E / E[0] - 1).max() < 1e-3


def test_leapfrog_start_steps_and_traj(oracle):
    h = oracle.milky_way_host()
    w0 = np.array([[12.0, 0.0, 0.0, 0.0, 0.153, 0.051]] * 3)
    out, tr = oracle.leapfrog(h, w0, -100.0, 1.0, 100, start=[0, 40, 100], traj=True)
    # a particle started at k integrates (nsteps - k) steps from its own start time
    ref = oracle.leapfrog(h, w0[1:2], -60.0, 1.0, 60)
    np.testing.assert_array_equal(out[1], ref[0])
    np.testing.assert_array_equal(out[2], w0[2])
    np.testing.assert_array_equal(tr[:41, 1], np.repeat(w0[1:2], 41, axis=0))


# ------------------------------------------------------------------------------ PRNG
def test_threefry_known_answers(oracle):
    # Random123 kat_vectors, threefry2x32 with 20 rounds
    assert oracle.threefry2x32(0, 0, 0, 0) == (0x6B200159, 0x99BA4EFE)
    assert oracle.threefry2x32(0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF) == (0x1CB996FC, 0xBB002BE7)
    assert oracle.threefry2x32(0x13198A2E, 0x03707344, 0x243F6A88, 0x85A308D3) == (0xC4923A9C, 0x483DF7A0)


@pytest.mark.parametrize("u", [1e-12, 0.1, -0.3, 0.49, 0.5, 0.7, 0.999, 1 - 1e-10, -(1 - 2 ** -53)])
def test_erfinv_against_scipy(oracle, u):
    assert oracle.erfinv(u) == pytest.approx(sp_erfinv(u), rel=4e-16, abs=1e-300)


def test_normal_draws_statistics(oracle):
    z = np.array([oracle.prng_normal(k) for k in range(1, 20001)])
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1) < 0.03
    r = oracle.randint5_1000(0)
    assert r.shape == (5,) and np.all((r >= 0) & (r < 1000))
    # release index 0 gives identical keys for all four draws (reference main.py:223-228)
    z0 = oracle.release_normals(0, 0)
    assert np.all(z0 == z0[0])


# ------------------------------------------------------------------------------ release
def test_release_geometry_spherical_circular(oracle):
    """Circular orbit in a Plummer sphere with all dispersions zero: positions and velocities
    follow from Omega^2 = phi'/r and phi'' in closed form (reference main.py:99-112,269-278)."""
    gm, a = G * 1e11, 2.0
    h = oracle.host([(oracle.COMP_PLUMMER, [gm, a])])
    r = 8.0
    d1 = gm * r / (r * r + a * a) ** 1.5
    d2 = gm * (a * a - 2 * r * r) / (r * r + a * a) ** 2.5
    vc = np.sqrt(r * d1)
    w = np.array([[r, 0, 0, 0, vc, 0]])
    msat = 1e5
    kv = [2.0, 0.3, 0.5, 0.1, 0.0, 0.0, 0.0, 0.0]
    pc, pf, vcl, vfa = oracle.release(h, w, [0.0], msat, seed=0, kvals=kv)
    om = vc / r
    rt = (G * msat / (om * om - d2)) ** (1 / 3)
    np.testing.assert_allclose(pf[0], [r + 2.0 * rt, 0, 0.5 * rt], rtol=1e-13)
    np.testing.assert_allclose(pc[0], [r - 2.0 * rt, 0, -0.5 * rt], rtol=1e-13)
    kvphi = 2.0 * 0.3
    np.testing.assert_allclose(vfa[0], [0, vc + kvphi * om * rt, 0.1 * om * rt], rtol=1e-13)
    np.testing.assert_allclose(vcl[0], [0, vc - kvphi * om * rt, -0.1 * om * rt], rtol=1e-13)


def test_release_jacobian_against_fd(oracle):
    h = oracle.milky_way_host()
    rng = np.random.default_rng(7)
    w = np.array([[12.0, 0.5, 1.0, 0.01, 0.153, 0.051], [-6.0, 9.0, 2.0, -0.15, -0.08, 0.04]])
    J = oracle.release_jacobian(h, w, 1e4, seed=0)
    nin = J.shape[2]
    assert nin == 6 + 2 * 3

    def rel(wv, hh):
        pc, pf, vc, vf = oracle.release(hh, wv, np.zeros(len(wv)), 1e4, seed=0)
        return np.concatenate([pc, pf, vc, vf], axis=1)

    for k in range(6):
        e = 1e-6 * max(1.0, abs(w[:, k]).max())
        wp, wm = w.copy(), w.copy()
        wp[:, k] += e
        wm[:, k] -= e
        fd = (rel(wp, h) - rel(wm, h)) / (2 * e)
        np.testing.assert_allclose(J[:, :, k], fd, rtol=2e-5, atol=1e-8 * np.abs(fd).max())
    comps = [(1, [G * 6.8e10, 3.0, 0.28]), (3, [G * 5e9, 1.0, 0.0]), (2, [G * 5.4e11, 15.62])]
    for c in range(3):
        for q in range(2):
            e = 1e-4 * comps[c][1][q]
            cp = [list(x[1]) for x in comps]
            cm = [list(x[1]) for x in comps]
            cp[c][q] += e
            cm[c][q] -= e
            fd = (rel(w, oracle.host([(comps[i][0], cp[i]) for i in range(3)])) -
                  rel(w, oracle.host([(comps[i][0], cm[i]) for i in range(3)]))) / (2 * e)
            # normwise: entries span many decades, the FD error is relative to the largest
            assert np.abs(J[:, :, 6 + 2 * c + q] - fd).max() < 1e-6 * np.abs(fd).max()
