"""Discrete adjoint of the surrogate leapfrog (ssc_surrogate_adjoint_forward / _backward).

(1) Every cell flagged: the clump force is the exact sum, so lam0 and the host-parameter
    gradients must equal the oracle's plain adjoint (oracle.adjoint_leapfrog, pinned by finite
    differences) and the plain GPU adjoint, to 1e-9.
(2) Mixed and empty flag maps: against the oracle of the surrogate adjoint with the
    BF16-emulating MLP (oracle/surrogate.py).  Both sides round the same operands to BF16 but
    accumulate in different orders and the GPU reconstructs positions by walking the leapfrog
    backwards, so the comparison is normwise at 2e-5 (measured: about 1e-6).
Ragged starts (including rows that start at a checkpoint and at the last step), a checkpoint
interval that does not divide the step count, and per-row forcing are exercised.  The forward
with checkpoints must reproduce ssc_surrogate_leapfrog bit for bit.
"""
import numpy as np
import pytest
import torch

import oracle.mlp as OM
import oracle.surrogate as OS
from test_gpu_core import bound_orbits

pytestmark = pytest.mark.gpu

G = 4.498502151469553e-12


def _np(t):
    return t.detach().cpu().numpy()


def _nrel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _setup(ssc, oracle, flags_kind, seed=3, cells=8, nclump=24):
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.mlp import MLP
    from streamsculptor_amd.potential import PlummerClumps
    from streamsculptor_amd.surrogate import ClumpSurrogate, Scaling
    rng = np.random.default_rng(seed)
    host = W.host_potential()
    cpos = rng.normal(scale=7.0, size=(nclump, 3))
    cm = 10 ** rng.uniform(7, 8.5, nclump)
    crs = 1.62 * np.sqrt(cm / 1e8)
    clumps = PlummerClumps(cm, crs, cpos, units=ssc.usys)
    p = OM.init_params(5) + rng.normal(scale=0.05, size=OM.NPARAM).astype(np.float32)
    half = np.full(3, 16.0)
    cell = 2 * half / cells
    if flags_kind == "all":
        flags = np.ones((cells,) * 3, np.uint8)
    elif flags_kind == "none":
        flags = np.zeros((cells,) * 3, np.uint8)
    else:
        flags = (rng.uniform(size=(cells,) * 3) < 0.3).astype(np.uint8)
    sc = Scaling((0.0, 0.0, 0.0), tuple(half), 1.0, 2e-4)
    sur = ClumpSurrogate(host, clumps, MLP(params=p), sc, torch.as_tensor(flags).cuda(),
                         tuple(-half), tuple(cell), (cells,) * 3)
    h = oracle.host(W.host_components())
    osrc = oracle.Sources(oracle.SRC_STATIC, gm=G * cm, rs=crs, pos=cpos)
    om = OS.Model(h, osrc, p, np.zeros(3), half, 2e-4, -half, cell, (cells,) * 3, flags)
    return sur, om, h, osrc, rng


def _case(rng, n=200, nsteps=150):
    w0 = bound_orbits(rng, n, rmin=6.0, rmax=12.0)
    start = np.sort(rng.integers(0, nsteps, n)).astype(np.int32)
    start[:3] = 0
    start[-2:] = nsteps            # never move
    start[50] = 40                 # starts exactly at a checkpoint row
    start = np.sort(start)
    lam_T = rng.normal(size=(n, 6))
    forcing = np.zeros((nsteps + 1, n, 6))
    forcing[::13] = 0.1 * rng.normal(size=forcing[::13].shape)
    return w0, start, lam_T, forcing


def test_forward_with_checkpoints_matches_plain_surrogate(ssc, oracle):
    sur, *_, rng = _setup(ssc, oracle, "mixed")
    w0, start, _, _ = _case(rng)
    A = sur.adjoint(w0, 0.0, 1.0, 150, start=start, checkpoint_every=40)
    wa = A.forward()
    wb, cnt = sur.leapfrog(w0, 0.0, 1.0, 150, start=start)
    assert torch.equal(wa, wb)
    assert int(A.fallback_forward.item()) == int(cnt.item()) > 0


def test_all_flagged_equals_plain_adjoint(ssc, oracle):
    from streamsculptor_amd.adjoint import LeapfrogAdjoint
    sur, om, h, osrc, rng = _setup(ssc, oracle, "all")
    w0, start, lam_T, forcing = _case(rng)
    A = sur.adjoint(w0, 0.0, 1.0, 150, start=start, checkpoint_every=40)
    A.forward()
    R = A.backward(lam_T, forcing=forcing)
    B = oracle.adjoint_leapfrog(h, w0, 0.0, 1.0, 150, lam_T, start=start, src=osrc,
                                forcing=forcing)
    from streamsculptor_amd.potential import Potential_Combine
    tot = Potential_Combine([sur.host, sur.clumps], units=ssc.usys)
    P = LeapfrogAdjoint(tot, w0, 0.0, 1.0, 150, start=start, checkpoint_every=40)
    P.forward()
    C = P.backward(lam_T, forcing=forcing, want_sources=False)
    e_or = _nrel(_np(R["lam0"]), B["lam0"])
    e_gh = _nrel(_np(R["g_host"]), B["g_host"])
    e_gpu = _nrel(_np(R["lam0"]), _np(C["lam0"]))
    print(f"all flagged: lam0 vs oracle {e_or:.2e}, g_host {e_gh:.2e}, vs plain GPU {e_gpu:.2e}")
    assert e_or < 1e-9 and e_gh < 1e-9 and e_gpu < 1e-9
    assert int(R["fallback"].item()) > 0


@pytest.mark.parametrize("kind", ["mixed", "none"])
def test_matches_oracle_surrogate_adjoint(ssc, oracle, kind):
    sur, om, h, osrc, rng = _setup(ssc, oracle, kind)
    w0, start, lam_T, forcing = _case(rng)
    A = sur.adjoint(w0, 0.0, 1.0, 150, start=start, checkpoint_every=40)
    A.forward()
    R = A.backward(lam_T, forcing=forcing)
    O = OS.adjoint(om, w0, 1.0, 150, lam_T, start=start, forcing=forcing, emulate=True)
    e0 = _nrel(_np(R["lam0"]), O["lam0"])
    eh = _nrel(_np(R["g_host"]), O["g_host"])
    row = np.linalg.norm(_np(R["lam0"]) - O["lam0"], axis=1) / np.linalg.norm(O["lam0"], axis=1)
    print(f"{kind}: lam0 {e0:.2e} (row median {np.median(row):.2e}, max {row.max():.2e}), "
          f"g_host {eh:.2e}, fallback rows {int(R['fallback'].item())}")
    assert e0 < 2e-5 and eh < 2e-5
    if kind == "none":
        assert int(R["fallback"].item()) == 0
