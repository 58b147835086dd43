"""GPU: the reference notebook's spray stream (StreamSubhaloExample.ipynb) through the drop-in
API, the release model in potentials with point sources, and the controller's first step.

The notebook value pins the release draws end to end (see test_oracle_spray_notebook.py for
the tolerance: per-component 5e-4 relative, the truncation floor of the notebook's
rtol = 1e-7 stream plus one particle crossing the patch edge).
"""
import json
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = 4.498502151469553e-12
TOL = 5e-4


def kat(name):
    with open(os.path.join(HERE, "reference_kats.json")) as f:
        return json.load(f)[name]


def test_notebook_patch_on_gpu(oracle, ssc):
    from streamsculptor_amd import potential as P
    k = kat("spray_patch_w_impact")
    ref = np.array(k["value"])
    pot = P.GalaMilkyWayPotential(units=ssc.usys)
    ts = np.linspace(-3500.0, 0.0, 5000)
    msat = np.linspace(1e4, 0.0, 5000)
    print("notebook stream: 9998 particles, Dopri5 rtol 1e-7, dtmin 0.3, seed 583", flush=True)
    lead, trail = pot.gen_stream_scan(ts=ts, prog_w0=np.array(k["prog_w0"]), Msat=msat,
                                      seed_num=k["seed"], solver=ssc.Dopri5(),
                                      kval_arr=np.array(k["kvals"]), max_steps=10 ** 5)
    stream = torch.cat([lead, trail]).cpu().numpy()
    sel = (stream[:, 1] > -5) & (stream[:, 1] < -4)
    avg = stream[sel].mean(0)
    w = pot.integrate_orbit(w0=avg, ts=[0.0, k["t_impact"]], t0=0.0, t1=k["t_impact"],
                            solver=ssc.Dopri5(), rtol=1e-13, atol=1e-13, dtmin=None,
                            max_steps=10 ** 6).ys[-1].cpu().numpy()
    err = np.abs(w - ref) / np.abs(ref)
    print(f"patch rows {sel.sum()}, w_impact rel err {err.max():.2e}")
    assert err.max() < TOL, err
    # the same stream from the oracle: same controller, same draws
    h = oracle.host([(oracle.COMP_MN, [G * 6.8e10, 3.0, 0.28]),
                     (oracle.COMP_HERNQUIST, [G * 5e9, 1.0, 0.0]),
                     (oracle.COMP_HERNQUIST, [G * 1.71e9, 0.07, 0.0]),
                     (oracle.COMP_NFW, [G * 5.4e11, 15.62])])
    rl, rt, _ = oracle.stream_dopri5(h, np.array(k["prog_w0"]), ts, msat, k["seed"],
                                     kvals=np.array(k["kvals"]), rtol=1e-7, atol=1e-7, dtmin=0.3,
                                     max_steps=10 ** 5)
    d = np.abs(stream - np.vstack([rl, rt])).max(1) / np.abs(np.vstack([rl, rt])).max(1)
    print(f"rows vs oracle: median {np.median(d):.1e} p99 {np.percentile(d, 99):.1e} "
          f"max {d.max():.1e}")
    assert np.median(d) < 1e-9 and np.percentile(d, 99) < 1e-6


def test_notebook_subhalo_potentials_on_gpu(ssc):
    from streamsculptor_amd import potential as P
    k = kat("subhalo_line_potentials")
    wi = np.array(kat("spray_patch_w_impact")["value"])
    kw = dict(subhalo_x0=np.array([wi[:3]]), subhalo_v=np.array([wi[3:] + [0.02, 0.0, 0.02]]),
              subhalo_t0=np.array([-1000.0]), t_window=np.array([250.0]), units=ssc.usys)
    pl = P.SubhaloLinePotential(m=np.array([1e7]), a=np.array([0.2]), **kw)
    hq = P.SubhaloLinePotential_Custom(
        pot=P.HernquistPotential(m=np.array([1e7]), r_s=np.array([0.2]), units=ssc.usys), **kw)
    assert float(pl.potential(np.array(k["xyz"]), k["t"])) == pytest.approx(k["plummer"], rel=1e-6)
    assert float(hq.potential(np.array(k["xyz"]), k["t"])) == pytest.approx(k["hernquist"],
                                                                            rel=1e-6)
    assert float(hq.potential(np.array(k["xyz"]), -1300.0)) == 0.0


def test_release_with_sources_on_gpu(oracle, ssc):
    """release_model / gen_stream_ics in host + point sources use the Hessian of the whole
    potential, as the oracle does (STATIC, TRACK at the release's row, LINE at its time)."""
    from streamsculptor_amd import potential as P
    from streamsculptor_amd import workloads as W
    h = oracle.milky_way_host()
    host = W.host_potential()
    rng = np.random.default_rng(9)
    n = 64
    x = np.array([12.0, 1.0, 0.5]) + rng.normal(scale=0.3, size=(n, 3))
    v = np.array([0.01, 0.18, 0.05]) + rng.normal(scale=0.01, size=(n, 3))
    w = np.hstack([x, v])
    near = np.array([[12.6, 1.1, 0.3], [11.6, 1.5, 0.8]])
    m = np.array([3e8, 1e8])
    rs = np.array([0.3, 0.2])
    ts = -45.0 + np.arange(n, dtype=np.float64) * 0.5
    track = np.stack([near + 0.01 * k for k in range(n)])
    cases = {
        "static": (P.PlummerClumps(m, rs, near, units=ssc.usys),
                   oracle.Sources(oracle.SRC_STATIC, gm=G * m, rs=rs, pos=near), None),
        "track": (P.LiveSubhaloPotential(m, rs, torch.as_tensor(track, device="cuda"),
                                         float(ts[0]), 0.5, units=ssc.usys),
                  oracle.Sources(oracle.SRC_TRACK, gm=G * m, rs=rs, pos=track), np.arange(n)),
        "line": (P.SubhaloLinePotential(m, rs, near, np.zeros((2, 3)), np.array([-40.0, -35.0]),
                                        np.array([8.0, 8.0]), units=ssc.usys),
                 oracle.Sources(oracle.SRC_LINE, gm=G * m, rs=rs, pos=near, vel=np.zeros((2, 3)),
                                t0=[-40.0, -35.0], t_window=[8.0, 8.0]), None),
        "line_hernquist": (
            P.SubhaloLinePotentialCustom_fromFunc(P.HernquistPotential, m, rs, near,
                                                  np.zeros((2, 3)), np.array([-40.0, -35.0]),
                                                  np.array([8.0, 8.0]), units=ssc.usys),
            oracle.Sources(oracle.SRC_LINE_HERNQUIST, gm=G * m, rs=rs, pos=near,
                           vel=np.zeros((2, 3)), t0=[-40.0, -35.0], t_window=[8.0, 8.0]), None),
    }
    ref_host = oracle.release(h, w, ts, 2e4, seed=1)
    for name, (srcpot, osrc, steps) in cases.items():
        pot = P.Potential_Combine([host, srcpot], units=ssc.usys)
        out = pot.release_model(x=x, v=v, Msat=2e4, i=np.arange(n), t=ts, seed_num=1)
        ref = oracle.release(h, w, ts, 2e4, seed=1, src=osrc, steps=steps)
        for o, e in zip(out, ref):
            np.testing.assert_allclose(o.cpu().numpy(), e, rtol=1e-12,
                                       atol=1e-13 * np.abs(e).max(), err_msg=name)
        moved = np.abs(out[1].cpu().numpy() - ref_host[1]).max()
        print(f"release with {name} sources: moved by up to {moved:.2e} kpc")
        assert moved > 1e-4, name


def test_first_step_clamped_on_gpu(ssc):
    from streamsculptor_amd import adaptive
    from streamsculptor_amd import workloads as W
    host = W.host_potential()
    h, _ = host._descriptor()
    w0 = torch.tensor([[8.0, 0.0, 0.5, 0.0, 0.22, 0.02]], dtype=torch.float64, device="cuda")
    _, _, st = adaptive.run(h, None, w0, 0.3, t0=0.0, rtol=1e-14, atol=1e-14, dtmin=0.3,
                            max_steps=1)
    assert tuple(st[0].tolist()) == (1, 0, 0)
    _, _, st = adaptive.run(h, None, w0, 10.0, t0=0.0, rtol=1e-3, atol=1e-3, dtmin=None,
                            dtmax=0.1, max_steps=10 ** 4)
    assert tuple(st[0].tolist()) == (100, 0, 0)
