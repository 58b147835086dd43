"""Per-batch derivative files in the reference's layout (generate_derivs.py:201-204): a dict
with pert_out = [[lead, dlead], [trail, dtrail]], r_s_root and ImpactFrameParams, written as
<path>/<index>.npy and read back unchanged (CPU only)."""
import numpy as np
import pytest
import torch

from streamsculptor_amd.derivs import load_derivs_batch, save_derivs_batch


def test_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    T, nsh = 7, 3
    pert_out = [[torch.tensor(rng.normal(size=(T, 6))), torch.tensor(rng.normal(size=(T, nsh, 12)))],
                [torch.tensor(rng.normal(size=(T, 6))), torch.tensor(rng.normal(size=(T, nsh, 12)))]]
    rs = rng.uniform(0.1, 1.0, nsh)
    ifp = {"bImpact": rng.uniform(size=nsh), "tImpact": -rng.uniform(size=nsh) * 1000}
    name = save_derivs_batch(str(tmp_path), 12, pert_out, rs, ifp)
    assert name.endswith("/12.npy")
    with pytest.raises(ValueError):
        load_derivs_batch(str(tmp_path), 12)  # pickled files are refused unless trusted
    d = load_derivs_batch(str(tmp_path), 12, trusted=True)
    assert set(d) == {"pert_out", "r_s_root", "ImpactFrameParams"}
    for arm in range(2):
        for k in range(2):
            np.testing.assert_array_equal(d["pert_out"][arm][k], pert_out[arm][k].numpy())
    assert d["pert_out"][0][1].shape == (T, nsh, 12)
    np.testing.assert_array_equal(d["r_s_root"], rs)
    np.testing.assert_array_equal(d["ImpactFrameParams"]["tImpact"], ifp["tImpact"])
    # no impact parameters
    save_derivs_batch(str(tmp_path), 13, pert_out, rs)
    assert load_derivs_batch(str(tmp_path), 13, trusted=True)["ImpactFrameParams"] is None
