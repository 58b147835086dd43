"""Pins of the oracle's discrete leapfrog adjoint (CPU only): every gradient against central
finite differences of the loss evaluated through the oracle's forward leapfrog (an independent
code path: no Hessians, no reverse sweep).  The loss mixes a terminal term and a linear
functional of the intermediate rows (the forcing input)."""
import numpy as np
import pytest

G = 4.498502151469553e-12
COMPS = [(1, [G * 6.8e10, 3.0, 0.28]), (3, [G * 5e9, 1.0, 0.0]), (2, [G * 5.4e11, 15.62])]


def _setup(oracle, with_src=True):
    rng = np.random.default_rng(10)
    nsteps, t0, dt = 150, -150.0, 1.0
    w0 = np.array([[10.0, 1.0, 0.5, 0.02, 0.18, 0.03], [-7.0, 6.0, -1.0, -0.12, -0.1, 0.05],
                   [3.0, -9.0, 2.0, 0.15, 0.05, -0.04]])
    start = np.array([0, 30, 71], dtype=np.int32)
    m = 4
    gm = G * 10 ** rng.uniform(7.5, 8.5, m)
    rs = 1.62 * np.sqrt(gm / G / 1e8)
    sw = np.hstack([rng.normal(scale=8, size=(m, 3)), rng.normal(scale=0.1, size=(m, 3))])
    _, tr = oracle.leapfrog(oracle.host(COMPS), sw, t0, dt, nsteps, traj=True)
    track = np.ascontiguousarray(tr[:, :, :3])
    tgt = rng.normal(scale=5, size=(3, 3))
    C = rng.normal(scale=1e-2, size=(nsteps + 1, 3, 6))  # linear functional of every row
    return dict(nsteps=nsteps, t0=t0, dt=dt, w0=w0, start=start, gm=gm, rs=rs, track=track,
                tgt=tgt, C=C, with_src=with_src)


def _loss(oracle, S, comps=COMPS, w0=None, gm=None, track=None):
    w0 = S["w0"] if w0 is None else w0
    gm = S["gm"] if gm is None else gm
    track = S["track"] if track is None else track
    src = oracle.Sources(oracle.SRC_TRACK, gm=gm, rs=S["rs"], pos=track) if S["with_src"] else None
    out, tr = oracle.leapfrog(oracle.host(comps), w0, S["t0"], S["dt"], S["nsteps"], start=S["start"],
                              src=src, traj=True)
    return ((out[:, :3] - S["tgt"]) ** 2).sum() + (S["C"] * tr).sum()


def _adjoint(oracle, S):
    src = oracle.Sources(oracle.SRC_TRACK, gm=S["gm"], rs=S["rs"], pos=S["track"]) if S["with_src"] else None
    out = oracle.leapfrog(oracle.host(COMPS), S["w0"], S["t0"], S["dt"], S["nsteps"], start=S["start"], src=src)
    lamT = np.zeros((3, 6))
    lamT[:, :3] = 2 * (out[:, :3] - S["tgt"])
    # rows before a particle's start hold its initial row: their forcing acts on w0 directly
    F = S["C"].copy()
    for i, s in enumerate(S["start"]):
        F[s, i] += S["C"][:s, i].sum(0)
        F[:s, i] = 0.0
    return oracle.adjoint_leapfrog(oracle.host(COMPS), S["w0"], S["t0"], S["dt"], S["nsteps"], lamT,
                                   start=S["start"], src=src, forcing=F)


@pytest.mark.parametrize("with_src", [True, False])
def test_adjoint_initial_rows_fd(oracle, with_src):
    S = _setup(oracle, with_src)
    A = _adjoint(oracle, S)
    for i in range(3):
        for c in range(6):
            e = 1e-6 * (1.0 if c < 3 else 0.01)
            wp, wm = S["w0"].copy(), S["w0"].copy()
            wp[i, c] += e
            wm[i, c] -= e
            fd = (_loss(oracle, S, w0=wp) - _loss(oracle, S, w0=wm)) / (2 * e)
            assert A["lam0"][i, c] == pytest.approx(fd, rel=2e-5, abs=1e-6 * abs(A["lam0"]).max())


def test_adjoint_host_parameters_fd(oracle):
    S = _setup(oracle)
    A = _adjoint(oracle, S)
    for c in range(3):
        for q in range(2):
            e = 1e-5 * COMPS[c][1][q]
            cp = [(k, list(p)) for k, p in COMPS]
            cm = [(k, list(p)) for k, p in COMPS]
            cp[c][1][q] += e
            cm[c][1][q] -= e
            fd = (_loss(oracle, S, comps=cp) - _loss(oracle, S, comps=cm)) / (2 * e)
            assert A["g_host"][c, q] == pytest.approx(fd, rel=1e-4, abs=1e-8 * abs(fd) + 1e-12)


def test_adjoint_source_mass_and_position_fd(oracle):
    S = _setup(oracle)
    A = _adjoint(oracle, S)
    for j in range(4):
        e = 1e-5 * S["gm"][j]
        gp, gmn = S["gm"].copy(), S["gm"].copy()
        gp[j] += e
        gmn[j] -= e
        fd = (_loss(oracle, S, gm=gp) - _loss(oracle, S, gm=gmn)) / (2 * e)
        assert A["g_gm"][j] == pytest.approx(fd, rel=1e-4)
    rng = np.random.default_rng(0)
    for _ in range(6):
        k, j, c = int(rng.integers(1, S["nsteps"])), int(rng.integers(0, 4)), int(rng.integers(0, 3))
        e = 1e-4
        tp, tm = S["track"].copy(), S["track"].copy()
        tp[k, j, c] += e
        tm[k, j, c] -= e
        fd = (_loss(oracle, S, track=tp) - _loss(oracle, S, track=tm)) / (2 * e)
        assert A["g_pos"][k, j, c] == pytest.approx(fd, rel=1e-4, abs=1e-9)
