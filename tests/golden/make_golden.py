#!/usr/bin/env python3
"""Writes the committed golden fixtures of this directory.

* ``reference_kats.json`` -- known-answer values printed in the reference repository's own
  notebooks (cited cell by cell) and published cipher test vectors.  These are transcribed
  numbers, not computed here.
* ``oracle_small.npz`` -- inputs and outputs of the CPU oracle on small seeded cases of every
  kernel family (spray release, leapfrog stream, Dopri5, first-order responses, discrete
  adjoint, MLP forward).  GPU tests compare the HIP path with these vectors without running
  the oracle; CPU tests check that the oracle still reproduces them.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from oracle import mlp as OM  # noqa: E402

G = O.G

REFERENCE_KATS = {
    "G_kpc3_per_Msun_Myr2": {
        "value": 4.498502151469553e-12,
        "source": "gala/astropy G in (kpc, Myr, Msun); reproduces "
                  "streamsculptor/examples/custom_potential.ipynb:132"},
    "nfw_acceleration": {
        "potential": {"kind": "NFW", "m": 1e12, "r_s": 20.0},
        "xyz": [1.0, 2.0, 3.0],
        "value": [-0.0011937, -0.00238739, -0.00358109],
        "printed_decimals": 8,
        "source": "streamsculptor/examples/tests.ipynb:1671 (hamiltonian_field.term, cell 58)"},
    "nfw_orbit_evaluate_30": {
        "potential": {"kind": "NFW", "m": 1e12, "r_s": 20.0},
        "w0": [20.0, 15.0, 20.0, 0.08, 0.1, -0.05], "t0": 0.0, "t": 30.0,
        "value": [21.9793661, 17.67654451, 18.10530132, 0.051923, 0.07815599, -0.07552167],
        "printed_decimals": 8,
        "source": "streamsculptor/examples/tests.ipynb:242-243,302 (Dopri8, rtol = atol = 1e-7)"},
    "nfw_orbit_ys_sum": {
        "potential": {"kind": "NFW", "m": 1e12, "r_s": 20.0},
        "w0": [20.0, 0.0, 20.0, 0.0, 0.2, 0.0], "ts": "linspace(0, 3000, 1000)",
        "value": -936.42809302,
        "source": "streamsculptor/examples/tests.ipynb:1691,1747 (rtol = atol = 1e-6)"},
    "nfw_field_row1": {
        "w0": [20.0, 0.0, 20.0, 0.0, 0.2, 0.0], "t": 3000.0 / 999.0,
        "value": [19.9947007, 0.600547553, 19.9947007, -3.52920125e-3, 0.199947006,
                  -3.52920125e-3],
        "source": "streamsculptor/examples/tests.ipynb:1700,1717"},
    "gala_mw_back_3500": {
        "potential": "GalaMilkyWayPotential (disk, bulge, nucleus, halo)",
        "w_today": [20.0, 0.0, 20.0, 0.0, 0.15, 0.0], "t1": -3500.0,
        "value": [-7.23164146, -7.96692572, -10.81840286, 0.19182623, -0.20351324,
                  -0.01770436],
        "note": "Dopri8 at rtol = atol = 1e-7, dtmin = 0.3; about 1e-5 from converged",
        "source": "streamsculptor/examples/StreamSubhaloExample.ipynb:35"},
    "spray_patch_w_impact": {
        "potential": "GalaMilkyWayPotential (disk, bulge, nucleus, halo)",
        "prog_w0": [-7.23164146, -7.96692572, -10.81840286, 0.19182623, -0.20351324,
                    -0.01770436],
        "ts": "linspace(-3500, 0, 5000)", "Msat": "linspace(1e4, 0, 5000)", "seed": 583,
        "kvals": [2.0, 0.3, 0.0, 0.0, 0.5, 0.5, 0.5, 0.5],
        "stream": "gen_stream_scan, Dopri5 rtol = atol = 1e-7, dtmin = 0.3",
        "patch": "lead and trail rows with -5 < y < -4, mean row",
        "t_impact": -1000.0,
        "value": [-18.84511332, 4.23873108, -17.32878089, -0.04416233, -0.15074649,
                  -0.07956093],
        "note": "prog_w0 is the printed initial condition (StreamSubhaloExample.ipynb:35); the "
                "kvals are the vintage kept in the comment at streamsculptor/main.py:214, with "
                "which the notebook was run (the current defaults miss by 1.5e-3); the mean "
                "row is integrated back 1000 Myr with Dopri8 at rtol = atol = 1e-7",
        "source": "streamsculptor/examples/StreamSubhaloExample.ipynb:80,89-90 (stream), "
                  ":156,186-190 (patch, back integration), :179 (printed value)"},
    "subhalo_line_potentials": {
        "xyz": [1.0, 2.0, 3.0], "t": -850.0,
        "subhalo": "m = 1e7, r_s = 0.2, x0 = w_impact[:3], v = w_impact[3:] + (0.02, 0, 0.02), "
                   "t0 = -1000, t_window = 250 (w_impact = spray_patch_w_impact.value)",
        "plummer": -1.053815931966235e-06,
        "hernquist": -1.048913077256255e-06,
        "note": "plummer: SubhaloLinePotential; hernquist: SubhaloLinePotential_Custom with "
                "HernquistPotential(m, r_s); the subhalo row follows from the 8-digit printed "
                "w_impact, so both values hold to about 1e-7 relative",
        "source": "streamsculptor/examples/StreamSubhaloExample.ipynb:193-204,237 (subhalo), "
                  ":228-229 (printed values)"},
    "threefry2x32_20": {
        "vectors": [[[0, 0], [0, 0], [0x6B200159, 0x99BA4EFE]],
                    [[0xFFFFFFFF, 0xFFFFFFFF], [0xFFFFFFFF, 0xFFFFFFFF],
                     [0x1CB996FC, 0xBB002BE7]],
                    [[0x13198A2E, 0x03707344], [0x243F6A88, 0x85A308D3],
                     [0xC4923A9C, 0x483DF7A0]]],
        "layout": "[key, counter, output]",
        "source": "Random123 known-answer tests (Salmon et al., SC'11)"},
}


def host_components():
    return [(O.COMP_MN, [G * 6.8e10, 3.0, 0.28]), (O.COMP_HERNQUIST, [G * 5e9, 1.0, 0.0]),
            (O.COMP_NFW, [G * 5.4e11, 15.62])]


def bound_rows(rng, n, rmin=6.0, rmax=30.0):
    r = rng.uniform(rmin, rmax, n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    t = np.cross(u, rng.normal(size=(n, 3)))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    v = rng.uniform(0.6, 1.0, n)[:, None] * 0.2 * t + rng.normal(scale=0.02, size=(n, 3))
    return np.hstack([r[:, None] * u, v])


def oracle_vectors():
    h = O.host(host_components())
    rng = np.random.default_rng(20261202)
    out = {}
    # spray stream: progenitor (12,0,0,0,0.153,0.051) integrated back 200 x 1 Myr, 2 particles
    # per release step, seed 0, leapfrog
    prog_today = np.array([12.0, 0.0, 0.0, 0.0, 0.153, 0.051])
    back = O.leapfrog(h, prog_today, 0.0, -1.0, 200)[0]
    ts = -200.0 + np.arange(201.0)
    lead, trail = O.stream_leapfrog(h, back, ts, 1e4, 0, 1.0)
    out.update(stream_prog_w0=back, stream_ts=ts, stream_lead=lead, stream_trail=trail)
    # release model on 8 perturbed progenitor rows (seed 3)
    rows = np.tile(back, (8, 1)) + np.random.default_rng(7).normal(scale=0.1, size=(8, 6))
    pc, pf, vc, vf = O.release(h, rows, ts[:8], 1e4, 3)
    out.update(release_prog=rows, release_pc=pc, release_pf=pf, release_vc=vc, release_vf=vf)
    # leapfrog with 16 static Plummer sources, ragged starts
    gm = G * 10 ** rng.uniform(6, 8, 16)
    rs = 1.62 * np.sqrt(gm / G / 1e8)
    spos = rng.normal(scale=8.0, size=(16, 3))
    src = O.Sources(O.SRC_STATIC, gm=gm, rs=rs, pos=spos)
    w0 = bound_rows(rng, 64)
    start = np.sort(rng.integers(0, 300, 64)).astype(np.int32)
    lf = O.leapfrog(h, w0, -300.0, 1.0, 300, start=start, src=src)
    out.update(lf_w0=w0, lf_start=start, lf_gm=gm, lf_rs=rs, lf_pos=spos, lf_out=lf)
    # Dopri5 (rtol = atol = 1e-7, dtmin = 0.3) saved at 5 times, no sources
    dw0 = bound_rows(rng, 16)
    dts = np.array([0.0, 100.0, 333.0, 700.0, 1000.0])
    _, dys, dst = O.dopri5(h, dw0, 1000.0, t0=0.0, ts=dts)
    out.update(dp_w0=dw0, dp_ts=dts, dp_ys=dys, dp_stats=dst)
    # first-order responses: 8 particles x 4 static subhalos x 150 steps, mass and radius blocks
    pgm = G * 10 ** rng.uniform(6.5, 8, 4)
    prs = 1.62 * np.sqrt(pgm / G / 1e8)
    ppos = rng.normal(scale=6.0, size=(4, 3))
    pert = O.Sources(O.SRC_STATIC, gm=pgm, rs=prs, pos=ppos)
    rw0 = bound_rows(rng, 8)
    rdw0 = rng.normal(scale=1e-3, size=(8, 4, 12))
    rw, rdw = O.perturb_leapfrog(h, pert, rw0, -150.0, 1.0, 150, dw0=rdw0, with_radius=True)
    out.update(resp_w0=rw0, resp_dw0=rdw0, resp_gm=pgm, resp_rs=prs, resp_pos=ppos, resp_w=rw,
               resp_dw=rdw)
    # discrete adjoint: 12 particles, 6 static sources, 120 steps, random terminal adjoint
    agm = G * 10 ** rng.uniform(6.5, 8, 6)
    ars = 1.62 * np.sqrt(agm / G / 1e8)
    apos = rng.normal(scale=6.0, size=(6, 3))
    asrc = O.Sources(O.SRC_STATIC, gm=agm, rs=ars, pos=apos)
    aw0 = bound_rows(rng, 12)
    astart = np.sort(rng.integers(0, 40, 12)).astype(np.int32)
    lam_T = rng.normal(size=(12, 6))
    adj = O.adjoint_leapfrog(h, aw0, -120.0, 1.0, 120, lam_T, start=astart, src=asrc)
    out.update(adj_w0=aw0, adj_start=astart, adj_gm=agm, adj_rs=ars, adj_pos=apos,
               adj_lamT=lam_T, adj_lam0=adj["lam0"], adj_ghost=adj["g_host"],
               adj_ggm=adj["g_gm"])
    # MLP forward: parameters oracle.mlp.init_params(11) (not stored: 0.5 MB), 40 inputs in
    # [-1, 1]^3, BF16 operand emulation
    params = OM.init_params(11)
    x = np.random.default_rng(12).uniform(-1, 1, size=(40, 3)).astype(np.float32)
    y, _ = OM.forward(params, x)
    out.update(mlp_param_seed=np.array(11), mlp_param_sum=np.array(float(np.sum(params))),
               mlp_x=x, mlp_y=np.asarray(y, dtype=np.float32))
    return out


def main():
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(REFERENCE_KATS, f, indent=1)
    vec = oracle_vectors()
    np.savez_compressed(os.path.join(HERE, "oracle_small.npz"), **vec)
    size = os.path.getsize(os.path.join(HERE, "oracle_small.npz"))
    print(f"wrote reference_kats.json and oracle_small.npz ({size / 1024:.0f} KiB, "
          f"{len(vec)} arrays)")


if __name__ == "__main__":
    main()
