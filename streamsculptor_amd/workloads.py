"""Synthetic inputs of the stream workloads (seeded numpy; no GPU needed to build them).

* ``plain``       -- 3-component host (MN 6.8e10/3/0.28 + Hernquist 5e9/1 + NFW 5.4e11/15.62),
                     progenitor today at (12, 0, 0) kpc, (0, 0.153, 0.051) kpc/Myr, Msat = 1e4,
                     integrated back 4000 x 1 Myr; 2 spray particles per step (seed 0): 8000
                     particles, 16,004,000 particle-steps.
* ``subhalos``    -- 64 Plummer subhalos on their own host orbits: positions uniform on the
                     r = 30 kpc sphere, isotropic Gaussian velocities (sigma = 0.15 kpc/Myr),
                     masses log-uniform in [1e6, 1e8] Msun, r_s = 1.62 (M/1e8)^0.5 kpc (seed 2);
                     1e5 spray particles = 25 per release step.
* ``clumps``      -- 512 static Plummer clumps, radii drawn from the NFW(r_s = 15.62) enclosed
                     mass within 60 kpc, isotropic (seed 1), masses log-uniform in [1e5, 1e7],
                     r_s = 1.62 (M/1e8)^0.5; 1e6 spray particles = 250 per release step.
* ``gd1``         -- a batch of perturbed GD-1-like streams as the reference's GPU example runs
                     them (examples/GD1_and_SubhaloFitting_GPU.ipynb, cells 7, 18-27): Milky Way
                     host with nucleus, stream age 2500 Myr, stripping times
                     linspace(-2500, -300, 2000) plus the observation time 0 (4000 particles),
                     Msat = 1e4, seed 532, Dopri5 at rtol = atol = 1e-6 and dtmin = 0.05; each
                     simulation has one straight-line Plummer subhalo (t_window = 100 Myr) drawn
                     from the example's prior and placed at impact parameter b from the
                     progenitor's orbit at the impact time (seed 8942).
* ``ensemble``    -- the clumps potential with 2000 progenitors of stream lengths 2000-6000
                     steps (ensemble_setup, seed 4), 1000 release steps each (2 particles per
                     step): 4e6 particles on one 6000-step grid, differentiated with respect
                     to the halo (m, r_s) and every progenitor's initial row.
"""
from __future__ import annotations

import numpy as np

G = 4.498502151469553e-12

HOST = dict(m_disk=6.8e10, a_disk=3.0, b_disk=0.28, m_bulge=5e9, c_bulge=1.0,
            m_halo=5.4e11, rs_halo=15.62)
PROG_TODAY = np.array([12.0, 0.0, 0.0, 0.0, 0.153, 0.051])
MSAT = 1e4
PROG_RS = 0.01
T_STEPS = 4000
DT = 1.0


def host_components(m_halo=HOST["m_halo"], rs_halo=HOST["rs_halo"]):
    """(kind, params) list of the host with G*m first (kinds: 1 MN, 3 Hernquist, 2 NFW)."""
    return [(1, [G * HOST["m_disk"], HOST["a_disk"], HOST["b_disk"]]),
            (3, [G * HOST["m_bulge"], HOST["c_bulge"], 0.0]),
            (2, [G * m_halo, rs_halo])]


def host_potential(m_halo=HOST["m_halo"], rs_halo=HOST["rs_halo"]):
    from .potential import (HernquistPotential, MiyamotoNagaiDisk, NFWPotential,
                            Potential_Combine)
    from .main import usys
    return Potential_Combine([
        MiyamotoNagaiDisk(HOST["m_disk"], HOST["a_disk"], HOST["b_disk"], units=usys),
        HernquistPotential(HOST["m_bulge"], HOST["c_bulge"], units=usys),
        NFWPotential(m_halo, rs_halo, units=usys)], units=usys)


def release_times(nsteps=T_STEPS, dt=DT):
    """ts = -nsteps*dt .. 0; the last entry is the observation time."""
    return -nsteps * dt + dt * np.arange(nsteps + 1)


def subhalo_ics(n=64, seed=2, r_shell=30.0, sigma_v=0.15, mmin=1e6, mmax=1e8):
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(n, 3))
    pos = r_shell * u / np.linalg.norm(u, axis=1, keepdims=True)
    vel = rng.normal(scale=sigma_v, size=(n, 3))
    m = 10 ** rng.uniform(np.log10(mmin), np.log10(mmax), n)
    rs = 1.62 * np.sqrt(m / 1e8)
    return np.hstack([pos, vel]), m, rs


def nfw_mass_profile(r, rs):
    x = r / rs
    return np.log1p(x) - x / (1 + x)


def clump_table(n=512, seed=1, rmax=60.0, rs_halo=HOST["rs_halo"], mmin=1e5, mmax=1e7):
    """Clump positions drawn from the NFW density inside rmax (inverse CDF of the enclosed
    mass, tabulated), isotropic directions, log-uniform masses."""
    rng = np.random.default_rng(seed)
    grid = np.linspace(0.0, rmax, 200001)
    cdf = nfw_mass_profile(grid, rs_halo) / nfw_mass_profile(rmax, rs_halo)
    r = np.interp(rng.uniform(size=n), cdf, grid)
    u = rng.normal(size=(n, 3))
    pos = r[:, None] * u / np.linalg.norm(u, axis=1, keepdims=True)
    m = 10 ** rng.uniform(np.log10(mmin), np.log10(mmax), n)
    rs = 1.62 * np.sqrt(m / 1e8)
    return pos, m, rs


def particles_per_step(total, nsteps=T_STEPS):
    per = total // nsteps
    if per * nsteps != total:
        raise ValueError("particle count must be a multiple of the number of release steps")
    return per


def spray_layout(per_step):
    """Arms and seeds of the particles released at one step: ceil(per/2) seeds, lead and trail
    alternating; at odd steps an odd count drops a lead instead of a trail so the two arms
    stay balanced.  Returns a list (per step parity) of [(seed_offset, arm)] with arm 0 = lead
    (close), 1 = trail (far)."""
    nseed = (per_step + 1) // 2
    even, odd = [], []
    for s in range(nseed):
        even += [(s, 0), (s, 1)]
        odd += [(s, 1), (s, 0)]
    return even[:per_step], odd[:per_step]


def particle_steps(per_step, nsteps=T_STEPS):
    return per_step * nsteps * (nsteps + 1) // 2


# GD-1-like present-day point: the example's spline of the unperturbed stream at phi1 = -30
# (examples/GD1_and_SubhaloFitting_GPU.ipynb, cell 17 output).
GD1_TODAY = np.array([-12.19942006, 0.4483408, 6.52725897, -0.07686926, -0.2278697, -0.1179158])
GD1_AGE = 2500.0


def gd1_ts(n_strip=2000, t_dissolve=-300.0):
    return np.concatenate([np.linspace(-GD1_AGE, t_dissolve, n_strip), [0.0]])


def gd1_subhalo_prior(n, seed=8942):
    """Per-simulation subhalo parameters from the example's prior (cell 24): log10 m in [6, 7],
    log10 a in [log10 0.01, log10 0.2], log10 t_impact in [1, log10 2000], log10 b in
    [-5, log10 0.5], v_impact in [-0.4, 0.4] kpc/Myr; directions isotropic."""
    rng = np.random.default_rng(seed)
    m = 10 ** rng.uniform(6.0, 7.0, n)
    a = 10 ** rng.uniform(np.log10(0.01), np.log10(0.2), n)
    t_imp = 10 ** rng.uniform(1.0, np.log10(2000.0), n)
    b = 10 ** rng.uniform(-5.0, np.log10(0.5), n)
    v = rng.uniform(-0.4, 0.4, n)
    nb = rng.normal(size=(n, 3))
    nb /= np.linalg.norm(nb, axis=1, keepdims=True)
    nv = rng.normal(size=(n, 3))
    nv /= np.linalg.norm(nv, axis=1, keepdims=True)
    return dict(m=m, a=a, t_impact=t_imp, b=b, v=v, dir_b=nb, dir_v=nv)


def ensemble_setup(n_streams=2000, seed=4, lmin=2000, lmax=6000, rmin=10.0, rmax=25.0):
    """Progenitors of the stream ensemble: stream lengths (grid steps) uniform in [lmin, lmax]
    and initial rows at each progenitor's own start (t = -length): radius uniform in
    [rmin, rmax], isotropic direction, a tangential velocity of 0.75-1.0 times the circular
    speed of the spherically averaged host (disk as a point mass) in a random tangential
    direction, plus a 0.02 kpc/Myr Gaussian part."""
    rng = np.random.default_rng(seed)
    lengths = rng.integers(lmin, lmax + 1, n_streams)
    r = rng.uniform(rmin, rmax, n_streams)
    u = rng.normal(size=(n_streams, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    t = np.cross(u, rng.normal(size=(n_streams, 3)))
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    m_enc = (HOST["m_halo"] * nfw_mass_profile(r, HOST["rs_halo"])
             + HOST["m_bulge"] * r ** 2 / (r + HOST["c_bulge"]) ** 2 + HOST["m_disk"])
    vc = np.sqrt(G * m_enc / r)
    v = (rng.uniform(0.75, 1.0, n_streams) * vc)[:, None] * t
    v += rng.normal(scale=0.02, size=(n_streams, 3))
    return lengths, np.hstack([r[:, None] * u, v])


def gala_mw_components():
    """reference GalaMilkyWayPotential (potential.py:390-414): disk, bulge, nucleus, halo."""
    return [(1, [G * HOST["m_disk"], HOST["a_disk"], HOST["b_disk"]]),
            (3, [G * HOST["m_bulge"], HOST["c_bulge"], 0.0]),
            (3, [G * 1.71e9, 0.07, 0.0]),
            (2, [G * HOST["m_halo"], HOST["rs_halo"]])]
