"""The committed golden fixtures (tests/golden/): reference notebook values and oracle vectors.

CPU: the oracle reproduces every stored vector (a regression pin of the restatement) and the
reference known-answer values read from reference_kats.json.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixtures():
    return np.load(os.path.join(HERE, "oracle_small.npz"))


def kats():
    with open(os.path.join(HERE, "reference_kats.json")) as f:
        return json.load(f)


def test_oracle_reproduces_vectors(oracle):
    import sys
    sys.path.insert(0, HERE)
    import make_golden
    fresh = make_golden.oracle_vectors()
    F = fixtures()
    assert set(F.files) == set(fresh)
    for k in F.files:
        a, b = F[k], fresh[k]
        if a.dtype.kind in "iu":
            np.testing.assert_array_equal(a, b, err_msg=k)
        else:
            # OpenMP reductions of the adjoint's parameter sums may change order
            np.testing.assert_allclose(b, a, rtol=1e-11, atol=1e-300, err_msg=k)


def test_reference_kats_with_oracle(oracle):
    K = kats()
    G = oracle.G
    assert K["G_kpc3_per_Msun_Myr2"]["value"] == G
    nfw = oracle.host([(oracle.COMP_NFW, [G * 1e12, 20.0])])
    k = K["nfw_acceleration"]
    _, g, _ = oracle.evaluate(nfw, k["xyz"])
    np.testing.assert_allclose(-g[0], k["value"], rtol=0, atol=5e-9)
    k = K["nfw_orbit_evaluate_30"]
    _, ys, _ = oracle.dopri5(nfw, k["w0"], 3000.0, t0=0.0, ts=[0.0, k["t"]], rtol=1e-12,
                             atol=1e-12, dtmin=None, max_steps=10 ** 6)
    np.testing.assert_allclose(ys[0, 1], k["value"], rtol=0, atol=6e-9)
    k = K["nfw_orbit_ys_sum"]
    _, ys, _ = oracle.dopri5(nfw, k["w0"], 3000.0, t0=0.0, ts=np.linspace(0, 3000, 1000),
                             rtol=1e-12, atol=1e-12, dtmin=None, max_steps=10 ** 6)
    assert abs(ys.sum() / k["value"] - 1) < 1e-6
    np.testing.assert_allclose(ys[0, 1], K["nfw_field_row1"]["value"], rtol=3e-9)
    for key, ctr, outv in K["threefry2x32_20"]["vectors"]:
        assert oracle.threefry2x32(*key, *ctr) == tuple(outv)


def test_mlp_fixture_parameters_are_the_seeded_init():
    import oracle.mlp as OM
    F = fixtures()
    p = OM.init_params(int(F["mlp_param_seed"]))
    assert float(np.sum(p)) == pytest.approx(float(F["mlp_param_sum"]), rel=1e-12)
