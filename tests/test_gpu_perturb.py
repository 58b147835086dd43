"""GPU parity of the first-order response kernel (fields interface) against the oracle:
final base rows and every response vector <= 1e-9 relative (normwise per particle)."""
import numpy as np
import pytest
import torch

from test_gpu_core import bound_orbits

pytestmark = pytest.mark.gpu

G = 4.498502151469553e-12


def _rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a
    n = a.shape[0]
    a, b = a.reshape(n, -1), b.reshape(n, -1)
    return (np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)).max()


@pytest.mark.parametrize("kind,radius", [("track", False), ("static", False), ("line", True),
                                         ("track", True), ("hernquist", False),
                                         ("hernquist", True)])
def test_perturb_parity(oracle, ssc, kind, radius):
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd.fields import perturb_leapfrog
    from streamsculptor_amd.main import SourceSet
    from streamsculptor_amd import _lib
    rng = np.random.default_rng(12)
    host = W.host_potential()
    h = oracle.host(W.host_components())
    n, nsteps, t0, dt = 900, 600, -600.0, 1.0
    m = 64 if kind == "track" else 37
    w0 = bound_orbits(rng, n)
    start = np.sort(rng.integers(0, nsteps + 5, n)).astype(np.int32)
    if kind == "track":
        sw0, mm, rs = W.subhalo_ics(n=m, seed=2)
        _, tr = oracle.leapfrog(h, sw0, t0, dt, nsteps, traj=True)
        pos = np.ascontiguousarray(tr[:, :, :3])
        src = SourceSet(_lib.SRC_TRACK, gm=G * mm, rs=rs, pos=pos, grid=(t0, dt))
        so = oracle.Sources(oracle.SRC_TRACK, gm=G * mm, rs=rs, pos=pos)
    else:
        gm = G * 10 ** rng.uniform(6, 8, m)
        rs = 1.62 * np.sqrt(gm / G / 1e8)
        pos = rng.normal(scale=10.0, size=(m, 3))
        if kind == "static":
            src = SourceSet(_lib.SRC_STATIC, gm=gm, rs=rs, pos=pos)
            so = oracle.Sources(oracle.SRC_STATIC, gm=gm, rs=rs, pos=pos)
        else:
            vel = rng.normal(scale=0.2, size=(m, 3))
            tc = rng.uniform(t0, 0, m)
            tw = rng.uniform(50, 300, m)
            kk, ko = ((_lib.SRC_LINE_HERNQUIST, oracle.SRC_LINE_HERNQUIST) if kind == "hernquist"
                      else (_lib.SRC_LINE, oracle.SRC_LINE))
            src = SourceSet(kk, gm=gm, rs=rs, pos=pos, vel=vel, t0=tc, t_window=tw)
            so = oracle.Sources(ko, gm=gm, rs=rs, pos=pos, vel=vel, t0=tc, t_window=tw)
    dd = 12 if radius else 6
    dw0 = rng.normal(scale=1e-3, size=(n, m, dd))
    w, dw = perturb_leapfrog(host, src, w0, dw0, start, t0, dt, nsteps, radius)
    wr, dwr = oracle.perturb_leapfrog(h, so, w0, t0, dt, nsteps, dw0=dw0, start=start,
                                      with_radius=radius)
    e1, e2 = _rel(w, wr), _rel(dw, dwr)
    print(f"perturb[{kind}, radius={radius}] base {e1:.2e} responses {e2:.2e}", flush=True)
    assert e1 <= 1e-9 and e2 <= 1e-9
    # the split form (base orbits once per particle, through segments of rows) with segments
    # of 37 rows, whose boundaries fall inside particles' lives: the same arithmetic, bit for
    # bit; the form that recomputes the base orbit in every lane sums the source forces in
    # another order: the same to the last bits
    w2, dw2 = perturb_leapfrog(host, src, w0, dw0, start, t0, dt, nsteps, radius, segment_rows=37)
    np.testing.assert_array_equal(w2.cpu().numpy(), w.cpu().numpy())
    np.testing.assert_array_equal(dw2.cpu().numpy(), dw.cpu().numpy())
    w3, dw3 = perturb_leapfrog(host, src, w0, dw0, start, t0, dt, nsteps, radius, split=False)
    e3, e4 = _rel(w3, w.cpu().numpy()), _rel(dw3, dw.cpu().numpy())
    print(f"perturb[{kind}] split vs per-lane base orbits: base {e3:.2e} responses {e4:.2e}")
    assert e3 <= 1e-11 and e4 <= 1e-9


def test_integrate_field_api(oracle, ssc):
    """integrate_field with MassRadiusPerturbation_OTF: pytree in, pytree ys out, saved at ts."""
    from streamsculptor_amd import workloads as W
    from streamsculptor_amd import fields
    from streamsculptor_amd.potential import SubhaloLinePotential
    rng = np.random.default_rng(3)
    host = W.host_potential()
    h = oracle.host(W.host_components())
    m = 6
    gm = G * 10 ** rng.uniform(7, 8, m)
    rs = 1.62 * np.sqrt(gm / G / 1e8)
    pos, vel = rng.normal(scale=8, size=(m, 3)), rng.normal(scale=0.1, size=(m, 3))
    tc, tw = np.zeros(m), np.full(m, 1e5)
    pert = SubhaloLinePotential(gm / G, rs, pos, vel, tc, tw, units=ssc.usys)

    class Gen:  # the attributes integrate_field reads from a perturbation generator
        potential_base = host
        potential_perturbation = pert
        potential_structural = True

    w0 = np.array([10.0, 1.0, 0.5, 0.02, 0.18, 0.03])
    d0 = np.zeros((m, 12))
    ts = np.array([0.0, 100.0, 250.0])
    sol = fields.integrate_field(w0=[w0, d0], ts=ts, field=fields.MassRadiusPerturbation_OTF(Gen),
                                 solver="leapfrog", dt0=0.5)
    assert tuple(sol.ys[0].shape) == (3, 6) and tuple(sol.ys[1].shape) == (3, m, 12)
    so = oracle.Sources(oracle.SRC_LINE, gm=gm, rs=rs, pos=pos, vel=vel, t0=tc, t_window=tw)
    wr, dwr = oracle.perturb_leapfrog(h, so, w0[None], 0.0, 0.5, 500, dw0=d0[None], with_radius=True)
    assert _rel(sol.ys[0][-1:], wr) <= 1e-9
    assert _rel(sol.ys[1][-1:], dwr) <= 1e-9
    # hamiltonian_field goes through the orbit integrator
    solh = fields.integrate_field(w0=w0, ts=ts, field=fields.hamiltonian_field(host),
                                  solver="leapfrog", dt0=0.5)
    np.testing.assert_allclose(solh.ys[-1].cpu().numpy(), wr[0], rtol=1e-9)
    with pytest.raises(NotImplementedError):
        fields.integrate_field(w0=w0, ts=ts, field=fields.CustomField(lambda t, c, a: c),
                               solver="leapfrog", dt0=0.5)
