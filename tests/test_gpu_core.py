"""GPU parity of the core path against the oracle: potential evaluation, fixed-step leapfrog
(with STATIC / TRACK / LINE Plummer sources), release model and the spray stream.

Tolerances: evaluation 1e-12 relative; fixed-step final phase-space rows <= 1e-9 relative
(normwise per row); release conditions 1e-12.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

G = 4.498502151469553e-12


def rel_err_rows(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = np.asarray(b)
    num = np.linalg.norm(a - b, axis=-1)
    den = np.maximum(np.linalg.norm(b, axis=-1), 1e-300)
    return (num / den).max()


def bound_orbits(rng, n, rmin=6.0, rmax=30.0):
    """Bound, non-plunging initial rows: radius uniform in [rmin, rmax], speed 0.6-1.0 of a
    220 km/s-like circular speed in a random tangential direction plus a small radial part.
    Orbits that graze the centre amplify last-bit rounding chaotically, which no FP64
    implementation can hold at 1e-9; the stream workloads stay far from it."""
    r = rng.uniform(rmin, rmax, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    t = np.cross(u, rng.normal(size=(n, 3)))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    speed = rng.uniform(0.6, 1.0, n) * 0.2
    v = speed[:, None] * t + rng.normal(scale=0.02, size=(n, 3))
    return np.hstack([r[:, None] * u, v])


def _host_pair(oracle, ssc):
    from streamsculptor_amd import workloads as W
    comps = W.host_components()
    return oracle.host(comps), W.host_potential()


def _sources(rng, n=24, spread=8.0):
    gm = G * 10 ** rng.uniform(6, 8, n)
    rs = 1.62 * np.sqrt(gm / G / 1e8)
    pos = rng.normal(scale=spread, size=(n, 3))
    return gm, rs, pos


def test_eval_parity(oracle, ssc):
    from streamsculptor_amd.potential import GalaMilkyWayPotential, Potential_Combine, PlummerClumps
    rng = np.random.default_rng(0)
    pts = rng.normal(scale=10.0, size=(1000, 3))
    mw = GalaMilkyWayPotential(units=ssc.usys)
    h = oracle.host([(1, [G * 6.8e10, 3.0, 0.28]), (3, [G * 5e9, 1.0, 0.0]),
                     (3, [G * 1.71e9, 0.07, 0.0]), (2, [G * 5.4e11, 15.62])])
    gm, rs, pos = _sources(rng)
    pot = Potential_Combine([mw, PlummerClumps(gm / G, rs, pos, units=ssc.usys)], units=ssc.usys)
    src = oracle.Sources(oracle.SRC_STATIC, gm=gm, rs=rs, pos=pos)
    phi_o, g_o, H_o = oracle.evaluate(h, pts, src=src)
    phi = pot.potential(pts, 0.0).cpu().numpy()
    g = pot.gradient(pts, 0.0).cpu().numpy()
    H = pot.jacobian_force(pts, 0.0).cpu().numpy()
    np.testing.assert_allclose(phi, phi_o, rtol=1e-12)
    assert np.abs(g - g_o).max() <= 1e-12 * np.abs(g_o).max()
    assert np.abs(H - H_o).max() <= 1e-11 * np.abs(H_o).max()
    # acceleration = -gradient; density from the Laplacian; scalar point keeps shape (3,)
    a = mw.acceleration([1.0, 2.0, 3.0], 0.0)
    assert tuple(a.shape) == (3,)
    # reference NFW value (streamsculptor/examples/tests.ipynb:1671)
    from streamsculptor_amd.potential import NFWPotential
    an = NFWPotential(1e12, 20.0, units=ssc.usys).acceleration([1.0, 2.0, 3.0], 0.0).cpu().numpy()
    np.testing.assert_allclose(an, [-0.0011937, -0.00238739, -0.00358109], rtol=5e-6)


@pytest.mark.parametrize("kind", ["none", "static", "track", "line"])
def test_leapfrog_parity(oracle, ssc, kind):
    from streamsculptor_amd.potential import (Potential_Combine, PlummerClumps,
                                              SubhaloLinePotential, LiveSubhaloPotential)
    from streamsculptor_amd import workloads as W
    rng = np.random.default_rng(11)
    h, host = _host_pair(oracle, ssc)
    n, nsteps, t0, dt = 700, 1500, -1500.0, 1.0
    w0 = bound_orbits(rng, n)
    start = np.sort(rng.integers(0, nsteps + 10, n)).astype(np.int32)  # ragged, some never start
    src_o, pot = None, host
    if kind == "static":
        gm, rs, pos = _sources(rng, n=64)
        src_o = oracle.Sources(oracle.SRC_STATIC, gm=gm, rs=rs, pos=pos)
        pot = Potential_Combine([host, PlummerClumps(gm / G, rs, pos, units=ssc.usys)], units=ssc.usys)
    elif kind == "track":
        sub_w0, m, rs = W.subhalo_ics(n=64, seed=2)
        live = LiveSubhaloPotential.from_orbits(host, sub_w0, m, rs, t0, dt, nsteps)
        _, tr = oracle.leapfrog(h, sub_w0, t0, dt, nsteps, traj=True)
        assert rel_err_rows(live.orbits.reshape(-1, 6), tr.reshape(-1, 6)) < 1e-10
        track = live.track.cpu().numpy()  # identical track on both sides
        src_o = oracle.Sources(oracle.SRC_TRACK, gm=G * m, rs=rs, pos=track)
        pot = Potential_Combine([host, LiveSubhaloPotential(m, rs, live.track, t0, dt, units=ssc.usys)],
                                units=ssc.usys)
    elif kind == "line":
        gm, rs, pos = _sources(rng, n=32)
        vel = rng.normal(scale=0.2, size=(32, 3))
        tc = rng.uniform(t0, 0, 32)
        tw = rng.uniform(50, 400, 32)
        src_o = oracle.Sources(oracle.SRC_LINE, gm=gm, rs=rs, pos=pos, vel=vel, t0=tc, t_window=tw)
        pot = Potential_Combine([host, SubhaloLinePotential(gm / G, rs, pos, vel, tc, tw, units=ssc.usys)],
                                units=ssc.usys)
    ref = oracle.leapfrog(h, w0, t0, dt, nsteps, start=start, src=src_o)
    out, _ = pot._leapfrog(w0, t0, dt, nsteps, start=start)
    err = rel_err_rows(out, ref)
    o = out.cpu().numpy()
    per_row = np.linalg.norm(o - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"leapfrog[{kind}] row rel err median {np.median(per_row):.2e} "
          f"p99 {np.percentile(per_row, 99):.2e} max {err:.2e}")
    assert err <= 1e-9


@pytest.mark.parametrize("path", ["wave_per_orbit", "two_per_lane", "one_per_lane", "group_8",
                                  "group_16", "group_4"])
def test_leapfrog_launch_paths(oracle, ssc, path):
    """The launch shapes of ssc_leapfrog (one wave per orbit for few orbits in a large source
    table; two particles per lane for large batches; G lanes per orbit for mid-sized batches with
    sources; one particle per lane otherwise) against the oracle."""
    from streamsculptor_amd.potential import Potential_Combine, PlummerClumps
    from streamsculptor_amd import workloads as W
    rng = np.random.default_rng(21)
    h, host = _host_pair(oracle, ssc)
    nsrc = {"wave_per_orbit": 512, "one_per_lane": 12, "group_16": 64}.get(path, 48)
    pos, m, rs = W.clump_table(n=nsrc, seed=1)
    # sizes that select each shape (two per lane from 393,216 orbits; G from the wave target)
    n, nsteps = {"wave_per_orbit": (7, 800), "two_per_lane": (400001, 40),
                 "one_per_lane": (3001, 300), "group_8": (3001, 300), "group_16": (30001, 150),
                 "group_4": (100001, 60)}[path]
    w0 = bound_orbits(rng, n)
    start = np.sort(rng.integers(0, nsteps, n)).astype(np.int32)
    src_o = oracle.Sources(oracle.SRC_STATIC, gm=G * m, rs=rs, pos=pos)
    pot = Potential_Combine([host, PlummerClumps(m, rs, pos, units=ssc.usys)], units=ssc.usys)
    save = np.array([0, nsteps // 3, nsteps], dtype=np.int32)
    out, ys = pot._leapfrog(w0, -float(nsteps), 1.0, nsteps, start=start, save_idx=save)
    ref, tr = oracle.leapfrog(h, w0, -float(nsteps), 1.0, nsteps, start=start, src=src_o, traj=True)
    err = rel_err_rows(out, ref)
    print(f"path {path}: n={n} max row rel err {err:.2e}")
    assert err <= 1e-9
    assert rel_err_rows(ys.reshape(-1, 6), tr[save].reshape(-1, 6)) <= 1e-9


def test_integrate_orbit_saves_and_backward(oracle, ssc):
    from streamsculptor_amd import workloads as W
    h, host = _host_pair(oracle, ssc)
    ts = np.array([0.0, -1000.0, -2500.0, -4000.0])
    sol = host.integrate_orbit(w0=W.PROG_TODAY, ts=ts, t0=0.0, t1=-4000.0, solver="leapfrog", dt0=1.0)
    _, tr = oracle.leapfrog(h, W.PROG_TODAY[None], 0.0, -1.0, 4000, traj=True)
    exp = tr[[0, 1000, 2500, 4000], 0, :]
    assert tuple(sol.ys.shape) == (4, 6)
    assert rel_err_rows(sol.ys, exp) <= 1e-9
    # batch form
    w = np.stack([W.PROG_TODAY, W.PROG_TODAY * 0.9])
    solb = host.integrate_orbit_batch_vmapped(w0=w, ts=np.linspace(0, 500, 6), solver="leapfrog", dt0=0.5)
    _, trb = oracle.leapfrog(h, w, 0.0, 0.5, 1000, traj=True)
    assert tuple(solb.ys.shape) == (2, 6, 6)
    assert rel_err_rows(solb.ys.permute(1, 0, 2).reshape(-1, 6), trb[::200].reshape(-1, 6)) <= 1e-9
    with pytest.raises(ValueError):
        host.integrate_orbit(w0=W.PROG_TODAY, ts=[0.0, 10.25], solver="leapfrog", dt0=1.0)
    with pytest.raises(NotImplementedError):
        host.integrate_orbit(w0=W.PROG_TODAY, ts=[0.0, 10.0], solver=ssc.Dopri8())


def test_release_parity(oracle, ssc):
    h, host = _host_pair(oracle, ssc)
    rng = np.random.default_rng(4)
    n = 3000
    r = rng.uniform(5, 30, n)
    u = rng.normal(size=(n, 3))
    x = r[:, None] * u / np.linalg.norm(u, axis=1, keepdims=True)
    v = rng.normal(scale=0.15, size=(n, 3))
    w = np.hstack([x, v])
    ref = oracle.release(h, w, np.zeros(n), 1e4, seed=0)
    ts = np.zeros(n)
    gms = torch.full((n,), G * 1e4, dtype=torch.float64, device="cuda")
    idx = torch.arange(n, dtype=torch.int32, device="cuda")
    out = host._release(torch.as_tensor(w, device="cuda"), idx, gms, 0, 1.0)
    for o, e in zip(out, ref):
        np.testing.assert_allclose(o.cpu().numpy(), e, rtol=1e-12, atol=1e-13 * np.abs(e).max())
    # single-release API with an explicit index and custom kvals
    kv = [2.0, 0.3, 0.1, 0.1, 0.5, 0.5, 0.5, 0.5]
    pl, pt, vl, vt = host.release_model(x=x[7], v=v[7], Msat=1e4, i=7, t=0.0, seed_num=3, kval_arr=kv)
    refk = oracle.release(h, w[:8], np.zeros(8), 1e4, seed=3, kvals=kv)
    np.testing.assert_allclose(pl.cpu().numpy(), refk[0][7], rtol=1e-12)
    np.testing.assert_allclose(vt.cpu().numpy(), refk[3][7], rtol=1e-12)


def test_plain_stream_parity_full_size(oracle, ssc):
    """The plain spray stream at full size: 4000 steps, 8000 particles, seed 0."""
    from streamsculptor_amd import workloads as W
    h, host = _host_pair(oracle, ssc)
    back = host.integrate_orbit(w0=W.PROG_TODAY, ts=[0.0, -4000.0], t0=0.0, t1=-4000.0,
                                solver="leapfrog", dt0=1.0).ys[-1]
    prog_w0 = back.cpu().numpy()
    ts = W.release_times()
    lead, trail = host.gen_stream_vmapped(ts=ts, prog_w0=prog_w0, Msat=W.MSAT, seed_num=0,
                                          solver="leapfrog", dt0=1.0)
    assert tuple(lead.shape) == (4000, 6) and tuple(trail.shape) == (4000, 6)
    ref_lead, ref_trail = oracle.stream_leapfrog(h, prog_w0, ts, W.MSAT, 0, 1.0)
    e1, e2 = rel_err_rows(lead, ref_lead), rel_err_rows(trail, ref_trail)
    print(f"plain stream rel err lead {e1:.3e} trail {e2:.3e}")
    assert e1 <= 1e-9 and e2 <= 1e-9
    # ICs: the stream's release conditions match the oracle release at the oracle's orbit
    pc, pf, vc, vf = host.gen_stream_ics(ts=ts, prog_w0=prog_w0, Msat=W.MSAT, seed_num=0,
                                         solver="leapfrog", dt0=1.0)
    _, prog = oracle.leapfrog(h, prog_w0[None], ts[0], 1.0, 4000, traj=True)
    rc = oracle.release(h, prog[:, 0, :], ts, W.MSAT, 0)
    assert rel_err_rows(pc, rc[0]) < 1e-11 and rel_err_rows(vf, rc[3]) < 1e-9


def test_edge_cases(oracle, ssc):
    h, host = _host_pair(oracle, ssc)
    empty = torch.empty(0, 6, dtype=torch.float64, device="cuda")
    out, _ = host._leapfrog(empty, 0.0, 1.0, 10)
    assert out.shape[0] == 0
    w = np.array([[12.0, 0.0, 0.0, 0.0, 0.153, 0.051]])
    out0, _ = host._leapfrog(w, 0.0, 1.0, 0)
    np.testing.assert_array_equal(out0.cpu().numpy(), w)
    outs, _ = host._leapfrog(w, 0.0, 1.0, 5, start=[9])
    np.testing.assert_array_equal(outs.cpu().numpy(), w)
    # a bad descriptor is rejected with the library's message
    from streamsculptor_amd import _lib
    bad = _lib.HostPotential()
    bad.n_components = 1
    bad.c[0].kind = 99
    with pytest.raises(RuntimeError, match="unknown component kind"):
        wt = torch.as_tensor(w, device="cuda")
        _lib.call("ssc_leapfrog", _lib.ctypes.byref(bad), None, _lib.ptr(wt), None, 1, 0.0, 1.0,
                  3, _lib.ptr(wt.clone()), None, 0, None)
    # custom potentials cannot be lowered
    from streamsculptor_amd.potential import CustomPotential
    with pytest.raises(NotImplementedError):
        CustomPotential(lambda x, t: 0.0, units=ssc.usys).gradient([1.0, 0, 0], 0.0)
