"""Differentiable stream ensembles in the clumpy halo with the learned clump force
(streamsculptor_amd.ensemble): several progenitors of different stream ages on one grid, the
release model, particles through the surrogate adjoint, progenitor orbits through the plain
adjoint forced at the release rows.

(1) With every cell flagged (exact clump sum: the map is smooth FP64) the gradient with respect
    to the halo mass and scale radius and to progenitor coordinates must match central finite
    differences of the GPU forward to 1e-7 relative, or to 1e-7 of the loss per unit parameter
    (|L| / |theta|) for a component that happens to be small (measured: 5e-9 relative; the halo
    derivatives use Richardson-extrapolated differences, whose error is a few 1e-10 where the
    plain central difference's O(h^2) term is 2e-8).  The target's offsets are seeded, so the
    loss is the same on every run.
(2) With a mixed flag map the whole gradient must match the oracle pipeline (oracle.leapfrog
    progenitors, oracle.release and its Jacobian, the BF16-emulating surrogate adjoint of
    oracle/surrogate.py, oracle.adjoint_leapfrog for the progenitors) to 2e-5 normwise
    (measured: about 5e-6; the BF16 rounding of the two network implementations differs).
"""
import numpy as np
import pytest
import torch

import oracle.surrogate as OS
from test_gpu_core import bound_orbits
from test_gpu_surrogate_adjoint import _setup

pytestmark = pytest.mark.gpu

G = 4.498502151469553e-12
LENGTHS = [120, 200, 160, 90, 200]
R = 30


def _model(ssc, oracle, kind):
    from streamsculptor_amd.ensemble import SurrogateStreamEnsemble
    sur, om, h, osrc, rng = _setup(ssc, oracle, kind, seed=11)
    ens = SurrogateStreamEnsemble(sur, LENGTHS, n_release=R, msat=2e4, seed=5,
                                  checkpoint_every=50)
    prog = bound_orbits(rng, len(LENGTHS), rmin=7.0, rmax=11.0)
    return ens, sur, om, osrc, prog, rng


def test_gradient_matches_finite_differences(ssc, oracle):
    from streamsculptor_amd.ensemble import ensemble_loss
    ens, sur, om, osrc, prog, rng = _model(ssc, oracle, "all")
    st = ens.forward(5.4e11, 15.62, prog)
    gen = torch.Generator(device=st["wT"].device).manual_seed(7)
    target = st["wT"][:, :3] + 0.05 * torch.randn(st["wT"][:, :3].shape, generator=gen,
                                                  dtype=torch.float64, device=st["wT"].device)
    halo = torch.tensor([5.4e11, 15.62], dtype=torch.float64, device="cuda", requires_grad=True)
    pw = torch.tensor(prog, dtype=torch.float64, device="cuda", requires_grad=True)
    L = ensemble_loss(halo, pw, ens, target)
    L.backward()

    def loss(mh, rsh, p):
        s = ens.forward(mh, rsh, p)
        d = s["wT"][:, :3] - target
        return float((d * d).sum().item())

    L0 = float(L.item())
    checks = []
    for k, (dm, dr) in enumerate([(5.4e11 * 2e-6, 0.0), (0.0, 15.62 * 2e-6)]):
        # Richardson extrapolation of central differences at h and h/2: the loss's third
        # derivative in the halo radius makes the plain central difference's O(h^2) error
        # about 2e-8 at h = 1e-6 r_s (measured), against a few 1e-10 extrapolated
        fds = [(loss(5.4e11 + dm / q, 15.62 + dr / q, prog) -
                loss(5.4e11 - dm / q, 15.62 - dr / q, prog)) / (2 * (dm + dr) / q) for q in (1, 2)]
        fd = (4 * fds[1] - fds[0]) / 3
        checks.append((f"halo[{k}]", halo.grad[k].item(), fd, abs(L0) / (5.4e11, 15.62)[k]))
    for j, c in [(0, 0), (1, 4), (3, 2), (4, 5)]:
        eps = 1e-6 if c < 3 else 1e-7
        pp, pm = prog.copy(), prog.copy()
        pp[j, c] += eps
        pm[j, c] -= eps
        fd = (loss(5.4e11, 15.62, pp) - loss(5.4e11, 15.62, pm)) / (2 * eps)
        checks.append((f"prog[{j},{c}]", pw.grad[j, c].item(), fd, abs(L0) / abs(prog[j, c])))
    for name, a, b, scale in checks:
        print(f"{name}: adjoint {a:.10e} fd {b:.10e} rel {abs(a - b) / abs(b):.2e} "
              f"(L/theta {scale:.2e})")
        assert abs(a - b) <= 1e-7 * max(abs(b), scale), name


def test_gradient_matches_oracle_pipeline(ssc, oracle):
    from streamsculptor_amd import workloads as W
    ens, sur, om, osrc, prog, rng = _model(ssc, oracle, "mixed")
    st = ens.forward(5.4e11, 15.62, prog)
    wT = st["wT"].cpu().numpy()
    target = wT[:, :3] + 0.05 * rng.normal(size=(wT.shape[0], 3))
    lam_T = np.zeros_like(wT)
    lam_T[:, :3] = 2.0 * (wT[:, :3] - target)
    gm_h, grs_h, g_w0 = ens.backward(st, torch.as_tensor(lam_T, device="cuda"))

    # oracle pipeline on the same inputs
    S, T = len(LENGTHS), ens.T
    h = oracle.host(W.host_components())
    pstart = ens.prog_start
    _, tr = oracle.leapfrog(h, prog, ens.t0, 1.0, T, start=pstart, src=osrc, traj=True)
    lead, trail, jacs, prow = [], [], [], []
    for j in range(S):
        rows = tr[ens.rel_step[j], j]
        pc, pf, vc, vf = oracle.release(h, rows, np.zeros(R), 2e4, 5)
        lead.append(np.hstack([pc, vc]))
        trail.append(np.hstack([pf, vf]))
        jacs.append(oracle.release_jacobian(h, rows, 2e4, 5))
        prow.append(rows)
    w0 = np.vstack(lead + trail)
    start = np.concatenate([ens.rel_step.reshape(-1)] * 2)
    wref, _ = OS.leapfrog(om, w0, 1.0, T, start=start)
    dpos = np.linalg.norm(wref[:, :3] - wT[:, :3], axis=1)
    print(f"final positions vs oracle: median {np.median(dpos):.2e} max {dpos.max():.2e} kpc")
    assert np.median(dpos) < 1e-7 and dpos.max() < 5e-4
    A = OS.adjoint(om, w0, 1.0, T, lam_T, start=start)
    nr = S * R
    lam0 = A["lam0"]
    g_host = A["g_host"].copy()
    # release VJP: output order (pos_close, pos_far, vel_close, vel_far)
    gout = np.hstack([lam0[:nr, :3], lam0[nr:, :3], lam0[:nr, 3:], lam0[nr:, 3:]])
    jac = np.concatenate(jacs)
    gin = np.einsum("no,noi->ni", gout, jac)
    F = np.zeros((T + 1, S, 6))
    rr = ens.rel_step.reshape(-1)
    ss = np.repeat(np.arange(S), R)
    np.add.at(F, (rr, ss), gin[:, :6])
    nc = h.n_components
    g_host += gin[:, 6:].sum(0).reshape(nc, 2)
    C = oracle.adjoint_leapfrog(h, prog, ens.t0, 1.0, T, np.zeros((S, 6)), start=pstart, src=osrc,
                                forcing=F)
    g_host += C["g_host"]
    ref = np.concatenate([[G * g_host[2, 0], g_host[2, 1]], C["lam0"].reshape(-1)])
    got = np.concatenate([[gm_h.item(), grs_h.item()], g_w0.cpu().numpy().reshape(-1)])
    scale = np.abs(ref)
    print("halo grad gpu", got[:2], "oracle", ref[:2])
    e_halo = np.abs(got[:2] - ref[:2]) / scale[:2]
    e_prog = np.linalg.norm(got[2:] - ref[2:]) / np.linalg.norm(ref[2:])
    print(f"halo rel {e_halo}, progenitor rows normwise {e_prog:.2e}")
    assert (e_halo < 2e-5).all() and e_prog < 2e-5
