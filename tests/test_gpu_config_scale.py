"""GPU parity at the full sizes of the measured workloads (the other GPU tests run reduced sizes):

* live subhalos + adjoint: the discrete adjoint with all 64 subhalos of subhalo_ics(seed=2)
  on their orbits (TRACK), 2000 particles with ragged starts over 1200 steps, against the
  oracle adjoint; and the full 1e5-particle, 4000-step stream loss gradient (halo mass and
  scale radius, the subhalo masses that matter most, progenitor coordinates) against central
  differences of the GPU forward;
* first-order responses: 1e5 spray particles x 64 live-subhalo responses over 4000 steps, a
  strided sample against the oracle;
* clumpy halo, plain: 1e6 spray particles x 512 clumps over 4000 steps, a strided sample against
  the oracle;
* clumpy halo, surrogate + adjoint: a network trained on the 512-clump table along the orbits of
  24 progenitors with streams of 2000-6000 steps; sampled particles' final rows and adjoints
  against the oracle's BF16-emulating surrogate integrator and adjoint, and the progenitors'
  plain adjoint against the oracle.

Fixed-step rows are held to 1e-9 relative; the surrogate comparisons, between two BF16
network implementations whose accumulation orders differ, to the bounds stated there.
"""
import numpy as np
import pytest
import torch

import oracle.surrogate as OS

pytestmark = pytest.mark.gpu

G = 4.498502151469553e-12


def _rows_rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return np.linalg.norm(a - b, axis=-1) / np.linalg.norm(b, axis=-1)


def _nrel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_adjoint_64_live_subhalos_vs_oracle(oracle, ssc):
    from streamsculptor_amd import _lib
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.adjoint import LeapfrogAdjoint
    from streamsculptor_amd.main import SourceSet
    from test_gpu_core import bound_orbits
    rng = np.random.default_rng(64)
    host = W.host_potential()
    h = oracle.host(W.host_components())
    n, nsteps, t0, dt = 2000, 1200, -1200.0, 1.0
    sw0, m, rs = W.subhalo_ics(n=64, seed=2)
    _, tr = oracle.leapfrog(h, sw0, t0, dt, nsteps, traj=True)
    pos = np.ascontiguousarray(tr[:, :, :3])
    src = SourceSet(_lib.SRC_TRACK, gm=G * m, rs=rs, pos=pos, grid=(t0, dt))
    so = oracle.Sources(oracle.SRC_TRACK, gm=G * m, rs=rs, pos=pos)
    w0 = bound_orbits(rng, n)
    start = np.sort(rng.integers(0, nsteps + 3, n)).astype(np.int32)
    lamT = rng.normal(size=(n, 6))
    F = rng.normal(scale=1e-3, size=(nsteps + 1, n, 6))
    for i, s in enumerate(start):
        F[:s, i] = 0.0
    print("[config scale] adjoint, 64 live subhalos, 2000 particles x 1200 steps", flush=True)
    adj = LeapfrogAdjoint(host, w0, t0, dt, nsteps, start=start, checkpoint_every=100, src=src)
    wT = adj.forward()
    out = adj.backward(lamT, forcing=F)
    ref = oracle.adjoint_leapfrog(h, w0, t0, dt, nsteps, lamT, start=start, src=so, forcing=F)
    fwd = oracle.leapfrog(h, w0, t0, dt, nsteps, start=start, src=so)
    assert _rows_rel(wT, fwd).max() < 1e-9
    e = {k: _nrel(out[k], ref[k]) for k in ("lam0", "g_host", "g_gm", "g_pos")}
    print("[config scale] adjoint vs oracle " + " ".join(f"{k} {v:.2e}" for k, v in e.items()))
    assert all(v < 1e-9 for v in e.values()), e


def test_stream_loss_gradient_full_size(ssc):
    """The live-subhalo stream loss at full size (1e5 particles, 4000 steps, 64 subhalos,
    checkpoints every 100 steps): adjoint gradient against central differences of the GPU
    forward for the halo mass and scale radius, the three subhalo masses with the largest
    effect on the loss, and two progenitor coordinates.

    The halo parameters move the subhalos' orbits (two of them pass within 2 kpc of the
    centre), and particles that meet a displaced subhalo respond nonlinearly: the central
    difference's error falls as h^2 from 160 % at a relative step of 1e-6 to 0.3 % at 1e-8
    (measured; with the subhalo orbits frozen it matches the adjoint to 1e-7 at every step), so
    the halo derivatives use Richardson-extrapolated differences at relative steps of 4e-9 and
    2e-9."""
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.adjoint import StreamModel, stream_loss
    ts = W.release_times()
    sw0, m, rs = W.subhalo_ics()
    model = StreamModel(ts, sw0, rs, W.MSAT, W.particles_per_step(100_000), seed=0,
                        checkpoint_every=100)
    host = W.host_potential()
    prog_w0 = host.integrate_orbit(w0=W.PROG_TODAY, ts=[0.0, ts[0]], t0=0.0, t1=float(ts[0]),
                                   solver="leapfrog", dt0=W.DT).ys[-1].cpu().numpy()
    halo0 = np.array([W.HOST["m_halo"], W.HOST["rs_halo"]])
    with torch.no_grad():
        target = model.forward(halo0[0] * 1.05, halo0[1] * 0.97, m, prog_w0)["wT"][:, :3].clone()
    halo = torch.tensor(halo0, dtype=torch.float64, device="cuda", requires_grad=True)
    ms = torch.tensor(m, dtype=torch.float64, device="cuda", requires_grad=True)
    pw = torch.tensor(prog_w0, dtype=torch.float64, device="cuda", requires_grad=True)
    print("[config scale] stream loss gradient, 1e5 particles x 4000 steps", flush=True)
    L = stream_loss(halo, ms, pw, model, target)
    L.backward()
    L0 = float(L.item())

    def loss(hh, mm, ww):
        with torch.no_grad():
            d = model.forward(hh[0], hh[1], mm, ww)["wT"][:, :3] - target
            return float((d * d).sum().item())

    def central(f, x, k, eps):
        xp, xm = x.copy(), x.copy()
        xp[k] += eps
        xm[k] -= eps
        return (f(xp) - f(xm)) / (2 * eps)

    checks = []
    for k in range(2):
        f1, f2 = (central(lambda hh: loss(hh, m, prog_w0), halo0, k, halo0[k] * q)
                  for q in (4e-9, 2e-9))
        checks.append((f"halo[{k}]", halo.grad[k].item(), (4 * f2 - f1) / 3, abs(L0) / halo0[k]))
    gm = ms.grad.cpu().numpy()
    for j in np.argsort(-np.abs(gm * m))[:3]:
        fd = central(lambda mm: loss(halo0, mm, prog_w0), m, j, m[j] * 1e-4)
        checks.append((f"m[{j}]", gm[j], fd, 0.0))
    for c, eps in ((0, 1e-6), (4, 1e-8)):
        fd = central(lambda ww: loss(halo0, m, ww), prog_w0, c, eps)
        checks.append((f"prog[{c}]", pw.grad[c].item(), fd, abs(L0) / abs(prog_w0[c])))
    for name, a, b, scale in checks:
        print(f"[config scale] grad {name}: adjoint {a:.9e} fd {b:.9e} rel "
              f"{abs(a - b) / abs(b):.2e}")
    for name, a, b, scale in checks:
        assert abs(a - b) <= 2e-5 * max(abs(b), 1e-3 * scale), name


def test_responses_full_size_sample(oracle, ssc):
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.fields import perturb_leapfrog
    from streamsculptor_amd.potential import LiveSubhaloPotential
    from streamsculptor_amd.streams import spray_particles
    host = W.host_potential()
    h = oracle.host(W.host_components())
    ts = W.release_times()
    sw0, m, rs = W.subhalo_ics()
    live = LiveSubhaloPotential.from_orbits(host, sw0, m, rs, float(ts[0]), W.DT, W.T_STEPS)
    prog_w0 = host.integrate_orbit(w0=W.PROG_TODAY, ts=[0.0, ts[0]], t0=0.0, t1=float(ts[0]),
                                   solver="leapfrog", dt0=W.DT).ys[-1]
    w0, start, _, (t0, dt, nsteps, _) = spray_particles(host, host, ts, prog_w0, W.MSAT, 0,
                                                       W.particles_per_step(100_000), W.DT)
    print(f"[config scale] responses, {w0.shape[0]} particles x 64 subhalos x 4000 steps",
          flush=True)
    out, dw = perturb_leapfrog(host, live._kernel_sources(), w0, None, start, t0, dt, nsteps,
                               False)
    every = 200
    tens = live._kernel_sources().tensors()
    so = oracle.Sources(oracle.SRC_TRACK, gm=tens["gm"].cpu().numpy(),
                        rs=tens["rs"].cpu().numpy(), pos=tens["pos"].cpu().numpy())
    ws = w0[::every].cpu().numpy()
    rw, rdw = oracle.perturb_leapfrog(h, so, ws, t0, dt, nsteps, start=start[::every])
    e_w = _rows_rel(out[::every], rw)
    g = dw[::every].cpu().numpy()
    e_dw = np.linalg.norm((g - rdw).reshape(len(ws), -1), axis=1) / np.linalg.norm(
        rdw.reshape(len(ws), -1), axis=1)
    print(f"[config scale] responses sample of {len(ws)}: rows {e_w.max():.2e}, responses "
          f"{e_dw.max():.2e}")
    assert e_w.max() < 1e-9 and e_dw.max() < 1e-9


def test_clumps_plain_full_size_sample(oracle, ssc):
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.potential import PlummerClumps, Potential_Combine
    from streamsculptor_amd.streams import spray_particles
    host = W.host_potential()
    h = oracle.host(W.host_components())
    pos, m, rs = W.clump_table()
    pot = Potential_Combine([host, PlummerClumps(m, rs, pos, units=ssc.usys)], units=ssc.usys)
    ts = W.release_times()
    prog_w0 = pot.integrate_orbit(w0=W.PROG_TODAY, ts=[0.0, ts[0]], t0=0.0, t1=float(ts[0]),
                                  solver="leapfrog", dt0=W.DT).ys[-1]
    w0, start, _, (t0, dt, nsteps, _) = spray_particles(pot, pot, ts, prog_w0, W.MSAT, 0,
                                                       W.particles_per_step(1_000_000), W.DT)
    print(f"[config scale] clumps plain, {w0.shape[0]} particles x 512 clumps x 4000 steps",
          flush=True)
    out, _ = pot._leapfrog(w0, t0, dt, nsteps, start=start)
    every = 2000
    so = oracle.Sources(oracle.SRC_STATIC, gm=G * m, rs=rs, pos=pos)
    ref = oracle.leapfrog(h, w0[::every].cpu().numpy(), t0, dt, nsteps, start=start[::every],
                          src=so)
    e = _rows_rel(out[::every], ref)
    print(f"[config scale] clumps plain sample of {len(ref)}: median {np.median(e):.2e} "
          f"max {e.max():.2e}")
    assert e.max() < 1e-9


def test_surrogate_ensemble_512_clumps_vs_oracle(oracle, ssc):
    """The clumpy-halo ensemble with a network trained on the 512-clump table (24 progenitors,
    streams of 2000-6000 steps, 10 release steps each).  Sampled particles: final rows against
    the oracle's surrogate integrator (BF16-emulating network, the same flags) -- two network
    implementations that round the same operands but accumulate in different orders, whose
    differences grow along orbits in the clumpy halo: median below 1 pc, every row below
    30 pc; the adjoints lam0 of the same particles against the oracle's surrogate adjoint
    normwise within 2e-2 (median row below 1e-3).  The progenitors' plain adjoint, which sees
    the exact clump sum, against the oracle adjoint to 1e-9."""
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.ensemble import SurrogateStreamEnsemble
    from streamsculptor_amd.potential import PlummerClumps
    from streamsculptor_amd.surrogate import ClumpSurrogate
    host = W.host_potential()
    h = oracle.host(W.host_components())
    cpos, cm, crs = W.clump_table(512, seed=1)
    clumps = PlummerClumps(cm, crs, cpos, units=ssc.usys)
    lengths, prog = W.ensemble_setup(24, seed=4)
    ens = SurrogateStreamEnsemble(None, lengths, n_release=10, msat=W.MSAT, seed=0,
                                  checkpoint_every=100, clumps=clumps)
    print(f"[config scale] surrogate ensemble: 24 progenitors, lengths {lengths.min()}-"
          f"{lengths.max()}, training", flush=True)
    P = ens.training_positions(ens.host(5.4e11, 15.62), prog, n_sample=1 << 12)
    sur = ClumpSurrogate.train(host, clumps, P, iters=3000, n_val=1 << 16)
    ens.sur = sur
    parts = ens.particles(ens.host(5.4e11, 15.62), prog)
    w0 = parts["w0"]
    start = np.concatenate([ens.rel_step.reshape(-1)] * 2).astype(np.int32)
    order = np.argsort(start, kind="stable")
    T = ens.T
    wT, _ = sur.leapfrog(w0[torch.as_tensor(order, device="cuda")], ens.t0, 1.0, T,
                         start=start[order])
    sel = order[np.linspace(0, len(order) - 1, 6).astype(np.int64)]
    sc = sur.scaling
    om = OS.Model(h, oracle.Sources(oracle.SRC_STATIC, gm=G * cm, rs=crs, pos=cpos),
                  sur.mlp.params.cpu().numpy(), sc.center, sc.half, sc.g_scale, sur.grid_lo,
                  sur.grid_cell, sur.grid_n, sur.flags.cpu().numpy())
    w0s = w0[torch.as_tensor(sel, device="cuda")].cpu().numpy()
    ref, _ = OS.leapfrog(om, w0s, 1.0, T, start=start[sel])
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    got = wT[torch.as_tensor(inv[sel], device="cuda")].cpu().numpy()
    dpos = np.linalg.norm(got[:, :3] - ref[:, :3], axis=1)
    print(f"[config scale] surrogate rows vs oracle (kpc): {dpos}", flush=True)
    assert np.median(dpos) < 1e-3 and dpos.max() < 3e-2
    # adjoints of the sampled particles
    lam = np.random.default_rng(3).normal(size=(len(sel), 6))
    sa = sur.adjoint(w0s, ens.t0, 1.0, T, start=start[sel], checkpoint_every=100)
    sa.forward()
    g = sa.backward(torch.as_tensor(lam, device="cuda"))
    A = OS.adjoint(om, w0s, 1.0, T, lam, start=start[sel])
    e_rows = np.linalg.norm(g["lam0"].cpu().numpy() - A["lam0"], axis=1) / np.linalg.norm(
        A["lam0"], axis=1)
    print(f"[config scale] surrogate adjoint lam0 vs oracle per row: {e_rows}", flush=True)
    assert np.median(e_rows) < 1e-3 and _nrel(g["lam0"], A["lam0"]) < 2e-2
    # the progenitors' plain adjoint with the exact 512-clump sum
    from streamsculptor_amd.adjoint import LeapfrogAdjoint, _with_sources
    src = clumps._kernel_sources()
    total = _with_sources(ens.host(5.4e11, 15.62), src)
    lamp = np.random.default_rng(4).normal(size=(24, 6))
    pa = LeapfrogAdjoint(total, prog, ens.t0, 1.0, T, start=ens.prog_start, checkpoint_every=100,
                         src=src)
    pa.forward()
    gp = pa.backward(lamp)
    so = oracle.Sources(oracle.SRC_STATIC, gm=G * cm, rs=crs, pos=cpos)
    C = oracle.adjoint_leapfrog(h, prog, ens.t0, 1.0, T, lamp, start=ens.prog_start, src=so)
    e = {k: _nrel(gp[k], C[k]) for k in ("lam0", "g_host", "g_gm", "g_pos")}
    print("[config scale] progenitor adjoint vs oracle " +
          " ".join(f"{k} {v:.2e}" for k, v in e.items()))
    assert all(v < 1e-9 for v in e.values()), e
