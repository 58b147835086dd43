"""The spray stream of the reference's StreamSubhaloExample notebook, reproduced by the oracle.

The notebook (streamsculptor/examples/StreamSubhaloExample.ipynb) generates a stream in the
Gala Milky Way potential with seed 583, averages the rows of a patch -5 < y < -4 and prints
that mean integrated back 1000 Myr (`w_impact`, :179).  The patch holds ~250 of 9,998
particles, each placed by four Gaussian draws of the release model's key scheme (PRNGKey(seed)
-> randint -> PRNGKey(i r_k) -> normal, main.py:220-228,263-266), so the value pins the
Threefry key derivation, counter layout, 64-bit assembly, uniform -> normal map and randint
reduction end to end: another seed, or any other draw, moves it by 1e-3 to 1e-2 relative.

Tolerance: the notebook integrated with Dopri5 at rtol = atol = 1e-7 and dtmin = 0.3, whose
step sequence differs from ours in its last bits; single rows of such a stream differ from the
converged stream by up to 1.5e-4 relative (measured below), and one particle crossing the
patch edge moves the mean by about 0.29 kpc / 250 = 3e-4 relative in y.  The bound is 5e-4
relative per component (the rows' floor plus one edge crossing); the nearest of 29 other seeds
misses it by 1.8x, their median by 12x.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = 4.498502151469553e-12
TOL = 5e-4


def kat(name):
    with open(os.path.join(HERE, "reference_kats.json")) as f:
        return json.load(f)[name]


def mw(O):
    """GalaMilkyWayPotential (reference potential.py:390-414): disk, bulge, nucleus, halo."""
    return O.host([(O.COMP_MN, [G * 6.8e10, 3.0, 0.28]), (O.COMP_HERNQUIST, [G * 5e9, 1.0, 0.0]),
                   (O.COMP_HERNQUIST, [G * 1.71e9, 0.07, 0.0]), (O.COMP_NFW, [G * 5.4e11, 15.62])])


def w_impact(O, seed, kvals, tol=1e-7, dtmin=0.3):
    k = kat("spray_patch_w_impact")
    h = mw(O)
    ts = np.linspace(-3500.0, 0.0, 5000)
    msat = np.linspace(1e4, 0.0, 5000)
    lead, trail, st = O.stream_dopri5(h, np.array(k["prog_w0"]), ts, msat, seed, kvals=kvals,
                                      rtol=tol, atol=tol, dtmin=dtmin, max_steps=10 ** 6)
    assert st[:, 2].max() == 0
    stream = np.vstack([lead, trail])
    sel = (stream[:, 1] > -5) & (stream[:, 1] < -4)
    w, _, st = O.dopri5(h, stream[sel].mean(0), k["t_impact"], t0=0.0, rtol=1e-13, atol=1e-13,
                        dtmin=None, max_steps=10 ** 6)
    return w[0], stream, sel


def rel_err(w, ref):
    return np.abs(w - ref) / np.abs(ref)


def test_notebook_patch_reproduced(oracle):
    k = kat("spray_patch_w_impact")
    ref = np.array(k["value"])
    w, stream, sel = w_impact(oracle, k["seed"], np.array(k["kvals"]))
    assert 200 < sel.sum() < 300
    assert rel_err(w, ref).max() < TOL, rel_err(w, ref)
    # the tolerance floor: rows of the rtol 1e-7 stream against the converged stream
    _, conv, _ = w_impact(oracle, k["seed"], np.array(k["kvals"]), tol=1e-11, dtmin=None)
    d = np.abs(stream - conv).max(1) / np.abs(conv).max(1)
    assert d.max() < 3e-4 and np.median(d) < 1e-4


@pytest.mark.parametrize("case", ["other_seed", "current_kvals", "seed_plus_1000"])
def test_notebook_patch_discriminates(oracle, case):
    """Draws other than the reference's miss the printed value by far more than the bound."""
    k = kat("spray_patch_w_impact")
    ref = np.array(k["value"])
    seed, kv = k["seed"], np.array(k["kvals"])
    if case == "other_seed":
        seed += 1
    elif case == "seed_plus_1000":
        seed += 1000
    else:
        kv = None  # the defaults of the current reference code
    w, _, _ = w_impact(oracle, seed, kv)
    assert rel_err(w, ref).max() > 2 * TOL


def test_notebook_subhalo_potentials(oracle):
    """SubhaloLinePotential (Plummer) and SubhaloLinePotential_Custom(HernquistPotential) at
    (1, 2, 3), t = -850, for the notebook's subhalo built on the printed w_impact."""
    k = kat("subhalo_line_potentials")
    wi = np.array(kat("spray_patch_w_impact")["value"])
    none = oracle.host([])
    args = dict(gm=[G * 1e7], rs=[0.2], pos=[wi[:3]], vel=[wi[3:] + [0.02, 0.0, 0.02]],
                t0=[-1000.0], t_window=[250.0])
    for kind, val in ((oracle.SRC_LINE, k["plummer"]), (oracle.SRC_LINE_HERNQUIST, k["hernquist"])):
        src = oracle.Sources(kind, **args)
        phi, _, _ = oracle.evaluate(none, k["xyz"], t=k["t"], src=src)
        assert phi[0] == pytest.approx(val, rel=1e-6)
        # outside the time window the subhalo is absent
        phi, _, _ = oracle.evaluate(none, k["xyz"], t=-1300.0, src=src)
        assert phi[0] == 0.0


# ------------------------------------------------------------------------------ release + sources
def _tidal_radius_fd(O, h, src, x, v, msat, t=0.0, step=0):
    """(G Msat / (Omega^2 - rhat^T H rhat))^(1/3) with H by central differences of the gradient
    of host + sources (reference main.py:77-104)."""
    H = np.zeros((3, 3))
    for a in range(3):
        e = np.zeros(3)
        e[a] = 1e-5
        _, gp, _ = O.evaluate(h, x + e, t=t, step=step, src=src)
        _, gm, _ = O.evaluate(h, x - e, t=t, step=step, src=src)
        H[:, a] = (gp[0] - gm[0]) / 2e-5
    rh = x / np.linalg.norm(x)
    om = np.linalg.norm(np.cross(x, v)) / np.dot(x, x)
    return (G * msat / (om * om - rh @ H @ rh)) ** (1.0 / 3.0)


def test_release_sees_the_sources(oracle):
    """The release's tidal radius uses the Hessian of host + sources (STATIC, TRACK at its
    row, LINE at its time); with kvals (1, 0, 0, 0, 0...) the far particle sits at x + r_t r^."""
    h = oracle.milky_way_host()
    x, v = np.array([12.0, 1.0, 0.5]), np.array([0.01, 0.18, 0.05])
    w = np.concatenate([x, v])[None]
    msat = 2e4
    kv = [1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
    near = x + np.array([[0.6, 0.1, -0.2], [-0.4, 0.5, 0.3]])
    gm = G * np.array([3e8, 1e8])
    rs = np.array([0.3, 0.2])
    cases = [
        (oracle.Sources(oracle.SRC_STATIC, gm=gm, rs=rs, pos=near), 0.0, 0),
        (oracle.Sources(oracle.SRC_TRACK, gm=gm, rs=rs,
                        pos=np.stack([near + 5.0, near, near - 5.0])), 0.0, 1),
        (oracle.Sources(oracle.SRC_LINE, gm=gm, rs=rs, pos=near, vel=np.zeros((2, 3)),
                        t0=[-50.0, -50.0], t_window=[20.0, 20.0]), -45.0, 0),
        (oracle.Sources(oracle.SRC_LINE_HERNQUIST, gm=gm, rs=rs, pos=near, vel=np.zeros((2, 3)),
                        t0=[-50.0, -50.0], t_window=[20.0, 20.0]), -45.0, 0),
    ]
    rt_host = _tidal_radius_fd(oracle, h, None, x, v, msat)
    for src, t, step in cases:
        _, pf, _, _ = oracle.release(h, w, [t], msat, seed=0, kvals=kv, src=src, steps=[step])
        rt = np.linalg.norm(pf[0] - x)
        ref = _tidal_radius_fd(oracle, h, src, x, v, msat, t=t, step=step)
        assert rt == pytest.approx(ref, rel=1e-7)
        assert abs(rt / rt_host - 1) > 1e-2  # the sources matter here
    # a LINE subhalo outside its window leaves the host's tidal radius
    _, pf, _, _ = oracle.release(h, w, [-100.0], msat, seed=0, kvals=kv, src=cases[2][0])
    assert np.linalg.norm(pf[0] - x) == pytest.approx(rt_host, rel=1e-7)


# ------------------------------------------------------------------------------ controller start
def test_first_step_clamped_to_dtmin_and_accepted(oracle):
    """PIDController.init raises the starting step to dtmin and flags it: the first step of
    size dtmin is accepted whatever its error, so a run over exactly dtmin ends in one step even
    at a tolerance that step does not meet."""
    h = oracle.milky_way_host()
    w0 = np.array([8.0, 0.0, 0.5, 0.0, 0.22, 0.02])
    _, _, st = oracle.dopri5(h, w0, 0.3, t0=0.0, rtol=1e-14, atol=1e-14, dtmin=0.3,
                             max_steps=1)
    assert tuple(st[0]) == (1, 0, 0)


def test_first_step_clamped_to_dtmax(oracle):
    """At a loose tolerance the starting-step heuristic (0.1 < h < 0.5 here: with dtmax = 0.5
    the run takes 21 steps) exceeds dtmax = 0.1; clamped, the run takes exactly
    (t1 - t0) / dtmax accepted steps and no rejected one (unclamped: fewer than 100)."""
    h = oracle.milky_way_host()
    w0 = np.array([8.0, 0.0, 0.5, 0.0, 0.22, 0.02])
    _, _, st = oracle.dopri5(h, w0, 10.0, t0=0.0, rtol=1e-3, atol=1e-3, dtmin=None, dtmax=0.5,
                             max_steps=10 ** 4)
    assert tuple(st[0]) == (21, 0, 0)
    _, _, st = oracle.dopri5(h, w0, 10.0, t0=0.0, rtol=1e-3, atol=1e-3, dtmin=None, dtmax=0.1,
                             max_steps=10 ** 4)
    assert tuple(st[0]) == (100, 0, 0)
