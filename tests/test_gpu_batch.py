"""GPU tests of batched simulations (grouped sources): every simulation of a batch equals the
same simulation run on its own, and the grouped Dopri5 kernel matches the oracle."""
import numpy as np
import pytest
import torch

from test_gpu_core import bound_orbits, rel_err_rows

pytestmark = pytest.mark.gpu

G = 4.498502151469553e-12


def _batch(ssc, B=5, nsh=2, seed=4):
    from streamsculptor_amd.batch import SubhaloLineBatch
    rng = np.random.default_rng(seed)
    return SubhaloLineBatch(m=10 ** rng.uniform(6.5, 8.0, (B, nsh)),
                            a=rng.uniform(0.05, 0.3, (B, nsh)),
                            subhalo_x0=rng.normal(scale=8.0, size=(B, nsh, 3)),
                            subhalo_v=rng.normal(scale=0.15, size=(B, nsh, 3)),
                            subhalo_t0=rng.uniform(-800, -50, (B, nsh)),
                            t_window=rng.uniform(50, 300, (B, nsh)), units=ssc.usys)


def test_grouped_dopri5_matches_oracle(oracle, ssc):
    from streamsculptor_amd import adaptive
    from streamsculptor_amd import workloads as W
    bt = _batch(ssc, B=7, nsh=3)
    rng = np.random.default_rng(9)
    n = 700
    w0 = bound_orbits(rng, n)
    group = np.sort(rng.integers(0, bt.B, n)).astype(np.int32)
    t_start = rng.uniform(-900, -100, n)
    h = W.host_potential()._host_only()
    g = torch.as_tensor(group, device="cuda")
    out, _, st = adaptive.run(h, bt.desc(g), w0, 0.0, t_start=t_start)
    so = oracle.Sources(oracle.SRC_LINE, gm=G * bt.m.reshape(-1), rs=bt.a.reshape(-1),
                        pos=bt.subhalo_x0.reshape(-1, 3), vel=bt.subhalo_v.reshape(-1, 3),
                        t0=bt.subhalo_t0.reshape(-1), t_window=bt.t_window.reshape(-1),
                        group_size=bt.nsh, group=group)
    out_o, _, st_o = oracle.dopri5(oracle.host(W.host_components()), w0, 0.0, t_start=t_start,
                                   src=so)
    assert rel_err_rows(out, out_o) < 1e-6
    assert st.cpu().numpy()[:, 2].max() == 0
    # out-of-range group indices fail that orbit only (status 2), in bounds
    bad = g.clone()
    bad[3] = bt.B
    bad[5] = -1
    out_b, _, st_b = adaptive.run(h, bt.desc(bad), w0, 0.0, t_start=t_start)
    st_b = st_b.cpu().numpy()
    assert st_b[3, 2] == 2 and st_b[5, 2] == 2 and torch.isnan(out_b[3]).all()
    keep = np.setdiff1d(np.arange(n), [3, 5])
    np.testing.assert_array_equal(out_b.cpu().numpy()[keep], out.cpu().numpy()[keep])


def test_stream_batch_equals_single_runs(ssc):
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.batch import gen_stream_batch_with_pert
    from streamsculptor_amd.streamhelpers import (gen_stream_vmapped_with_pert,
                                                  gen_stream_vmapped_with_pert_fixed_prog)
    base = W.host_potential()
    bt = _batch(ssc)
    ts = np.concatenate([np.linspace(-1000.0, -200.0, 150), [0.0]])
    kw = dict(rtol=1e-6, atol=1e-6, dtmin=0.05)
    for fixed in (False, True):
        lead, trail = gen_stream_batch_with_pert(base, bt, ts, W.PROG_TODAY, 1e4, 11,
                                                 fixed_prog=fixed, **kw)
        assert lead.shape == (bt.B, 150, 6)
        fn = gen_stream_vmapped_with_pert_fixed_prog if fixed else gen_stream_vmapped_with_pert
        for b in (0, 3):
            l1, t1 = fn(base, bt[b], ts, W.PROG_TODAY, 1e4, 11, solver="dopri5", **kw)
            # grouped sources read the same records in the same order
            assert rel_err_rows(lead[b], l1.cpu().numpy()) < 1e-9
            assert rel_err_rows(trail[b], t1.cpu().numpy()) < 1e-9
