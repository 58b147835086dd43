"""The HIP path against the committed golden vectors (tests/golden/oracle_small.npz): no oracle
is run, the fixture is what the results are compared against.  Tolerances as in the other
parity tests: fixed-step rows 1e-9 relative, adaptive 1e-6, release 1e-12, MLP outputs 2e-3
(BF16 operands)."""
import os

import numpy as np
import pytest
import torch

from test_gpu_core import rel_err_rows

pytestmark = pytest.mark.gpu

G = 4.498502151469553e-12
F = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                         "oracle_small.npz"))


def _nrel(a, b):
    a = a.detach().cpu().double().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_stream_and_release(ssc):
    from streamsculptor_amd import workloads as W
    pot = W.host_potential()
    print("[golden] stream", flush=True)
    lead, trail = pot.gen_stream_vmapped(ts=F["stream_ts"], prog_w0=F["stream_prog_w0"],
                                         Msat=1e4, seed_num=0, solver="leapfrog", dt0=1.0)
    assert rel_err_rows(lead, F["stream_lead"]) < 1e-9
    assert rel_err_rows(trail, F["stream_trail"]) < 1e-9
    rows = torch.as_tensor(F["release_prog"], device="cuda")
    gms = torch.full((8,), G * 1e4, dtype=torch.float64, device="cuda")
    idx = torch.arange(8, dtype=torch.int32, device="cuda")
    out = pot._release(rows, idx, gms, 3, 1.0)
    for a, k in zip(out, ("release_pc", "release_pf", "release_vc", "release_vf")):
        assert rel_err_rows(a, F[k]) < 1e-12, k


def test_leapfrog_with_sources(ssc):
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.potential import PlummerClumps, Potential_Combine
    pot = Potential_Combine([W.host_potential(),
                             PlummerClumps(F["lf_gm"] / G, F["lf_rs"], F["lf_pos"],
                                           units=ssc.usys)], units=ssc.usys)
    out, _ = pot._leapfrog(F["lf_w0"], -300.0, 1.0, 300, start=F["lf_start"])
    assert rel_err_rows(out, F["lf_out"]) < 1e-9


def test_dopri5(ssc):
    from streamsculptor_amd import workloads as W
    sol = W.host_potential().integrate_orbit(w0=F["dp_w0"], ts=F["dp_ts"], solver="dopri5")
    assert rel_err_rows(sol.ys.reshape(-1, 6), F["dp_ys"].reshape(-1, 6)) < 1e-6


def test_responses(ssc):
    from streamsculptor_amd import _lib
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.fields import perturb_leapfrog
    from streamsculptor_amd.main import SourceSet
    src = SourceSet(_lib.SRC_STATIC, gm=F["resp_gm"], rs=F["resp_rs"], pos=F["resp_pos"])
    w, dw = perturb_leapfrog(W.host_potential(), src, F["resp_w0"], F["resp_dw0"], None, -150.0,
                             1.0, 150, True)
    assert rel_err_rows(w, F["resp_w"]) < 1e-9
    assert _nrel(dw, F["resp_dw"]) < 1e-9


def test_adjoint(ssc):
    from streamsculptor_amd import _lib
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.adjoint import LeapfrogAdjoint
    from streamsculptor_amd.main import SourceSet
    src = SourceSet(_lib.SRC_STATIC, gm=F["adj_gm"], rs=F["adj_rs"], pos=F["adj_pos"])
    adj = LeapfrogAdjoint(W.host_potential(), F["adj_w0"], -120.0, 1.0, 120,
                          start=F["adj_start"], checkpoint_every=50, src=src)
    adj.forward()
    out = adj.backward(F["adj_lamT"])
    assert _nrel(out["lam0"], F["adj_lam0"]) < 1e-9
    assert _nrel(out["g_host"], F["adj_ghost"]) < 1e-9
    assert _nrel(out["g_gm"], F["adj_ggm"]) < 1e-9


def test_mlp_forward(ssc):
    import oracle.mlp as OM
    from streamsculptor_amd.mlp import MLP
    net = MLP(params=OM.init_params(int(F["mlp_param_seed"])))
    assert _nrel(net.forward(F["mlp_x"]), F["mlp_y"]) < 2e-3
    assert _nrel(net.infer(F["mlp_x"]), F["mlp_y"]) < 2e-3
