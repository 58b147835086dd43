"""The oracle of the surrogate integrator and its discrete adjoint (oracle/surrogate.py):
(1) with every cell flagged it is the exact clump sum, so forward and adjoint must equal the
plain oracle (oracle.leapfrog / oracle.adjoint_leapfrog with the clumps as static sources, the
latter pinned by finite differences in test_oracle_adjoint.py), forcing included;
(2) with the smooth float64 MLP and a mix of flagged and unflagged cells the adjoint must match
central finite differences of the forward map in the initial rows and the halo parameters."""
import numpy as np

import oracle as O
import oracle.mlp as OM
import oracle.surrogate as OS

G = 4.498502151469553e-12


def _host(m_halo=5.4e11, rs_halo=15.62):
    return O.host([(O.COMP_MN, [G * 6.8e10, 3.0, 0.28]), (O.COMP_HERNQUIST, [G * 5e9, 1.0, 0.0]),
                   (O.COMP_NFW, [G * m_halo, rs_halo])])


def _clumps(rng, n=12):
    gm = G * 10 ** rng.uniform(7, 8.5, n)
    rs = 1.62 * np.sqrt(gm / G / 1e8)
    pos = rng.normal(scale=7.0, size=(n, 3))
    return O.Sources(O.SRC_STATIC, gm=gm, rs=rs, pos=pos)


def _rows(rng, n, rmin=6.0, rmax=12.0):
    r = rng.uniform(rmin, rmax, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    t = np.cross(u, rng.normal(size=(n, 3)))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    v = 0.2 * t * rng.uniform(0.7, 1.0, n)[:, None]
    return np.hstack([r[:, None] * u, v])


def _model(h, clumps, rng, flags=None, cells=4):
    p = OM.init_params(5) + rng.normal(scale=0.05, size=OM.NPARAM).astype(np.float32)
    half = np.full(3, 16.0)
    cell = 2 * half / cells
    if flags is None:
        flags = (rng.uniform(size=(cells,) * 3) < 0.3).astype(np.uint8)
    return OS.Model(h, clumps, p, center=np.zeros(3), half=half, g_scale=2e-4,
                    grid_lo=-half, grid_cell=cell, grid_n=(cells,) * 3, flags=flags)


def test_all_flagged_equals_plain_adjoint():
    rng = np.random.default_rng(0)
    h = _host()
    clumps = _clumps(rng)
    M = _model(h, clumps, rng, flags=np.ones((4, 4, 4), np.uint8))
    n, T, dt = 7, 90, 1.0
    w0 = _rows(rng, n)
    start = np.array([0, 0, 5, 30, 61, 89, 90], dtype=np.int32)
    w, _ = OS.leapfrog(M, w0, dt, T, start=start)
    ref = O.leapfrog(h, w0, 0.0, dt, T, start=start, src=clumps)
    np.testing.assert_allclose(w, ref, rtol=1e-12, atol=1e-13)
    lam_T = rng.normal(size=(n, 6))
    forcing = np.zeros((T + 1, n, 6))
    forcing[::7] = rng.normal(size=forcing[::7].shape)
    A = OS.adjoint(M, w0, dt, T, lam_T, start=start, forcing=forcing)
    B = O.adjoint_leapfrog(h, w0, 0.0, dt, T, lam_T, start=start, src=clumps, forcing=forcing)
    np.testing.assert_allclose(A["lam0"], B["lam0"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(A["g_host"], B["g_host"], rtol=1e-10, atol=1e-20)


def test_smooth_adjoint_matches_finite_differences():
    rng = np.random.default_rng(2)
    clumps = _clumps(rng)
    h = _host()
    M = _model(h, clumps, rng, cells=8)
    n, T, dt = 6, 80, 1.0
    w0 = _rows(rng, n)
    start = np.array([0, 0, 10, 25, 40, 79], dtype=np.int32)
    lam_T = rng.normal(size=(n, 6))

    def loss(w0_, mh=5.4e11, rsh=15.62):
        M.h = _host(mh, rsh)
        w, _ = OS.leapfrog(M, w0_, dt, T, start=start, emulate=False)
        return float((w * lam_T).sum())

    # both branches of the clump force are used along these orbits
    _, _, X = OS.leapfrog(M, w0, dt, T, start=start, emulate=False, traj=True)
    fb = M.fallback(X.reshape(-1, 3))
    assert 0.05 < fb.mean() < 0.95
    M.h = h
    A = OS.adjoint(M, w0, dt, T, lam_T, start=start, emulate=False)
    fd = np.zeros((n, 6))
    for i in range(n):
        for c in range(6):
            eps = 1e-6 * (1.0 if c < 3 else 0.1)
            wp, wm = w0.copy(), w0.copy()
            wp[i, c] += eps
            wm[i, c] -= eps
            fd[i, c] = (loss(wp) - loss(wm)) / (2 * eps)
    err = np.abs(A["lam0"] - fd).max() / np.abs(fd).max()
    assert err < 1e-6, err
    for k, (mh, rsh) in enumerate([(5.4e11 * (1 + 1e-6), 15.62), (5.4e11, 15.62 * (1 + 1e-6))]):
        mh2, rsh2 = 2 * 5.4e11 - mh, 2 * 15.62 - rsh
        step = (mh - 5.4e11) * G if k == 0 else rsh - 15.62
        d = (loss(w0, mh, rsh) - loss(w0, mh2, rsh2)) / (2 * step)
        assert abs(A["g_host"][2, k] - d) < 1e-6 * abs(d), (k, A["g_host"][2, k], d)
