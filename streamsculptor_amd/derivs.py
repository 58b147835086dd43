"""Batch production of perturbation derivatives saved per batch (the output format and loop body
of reference streamsculptor/generate_derivs.py:153-214).

Each batch of straight-line subhalos (unit mass, as the reference sets ``m = 1`` so that the
mass block is the derivative per unit subhalo mass) is turned into first-order responses of
every stream particle by ``GenerateMassRadiusPerturbation.compute_perturbation_OTF`` and
written as ``path/<index>.npy`` holding the dict the reference saves:

    {"pert_out": [[lead (T, 6), dlead (T, Nsh, 12)], [trail (T, 6), dtrail (T, Nsh, 12)]],
     "r_s_root": r_s (Nsh,), "ImpactFrameParams": dict or None}

with numpy arrays in place of JAX arrays (``numpy.save`` of a dict: a pickled 0-d object array,
read back with ``load_derivs_batch``).  The subhalo parameters come from the caller; the
reference draws them with its impact sampler (GenerateImpactParams.ImpactGenerator).
"""
from __future__ import annotations

import os

import numpy as np
import torch

from .perturbative import GenerateMassRadiusPerturbation
from .potential import (HernquistPotential, PlummerPotential,
                        SubhaloLinePotentialCustom_dRadius_fromFunc,
                        SubhaloLinePotentialCustom_fromFunc)


def _np(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    if isinstance(x, dict):
        return {k: _np(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [_np(v) for v in x]
    return np.asarray(x) if x is not None else None


def save_derivs_batch(path, index, pert_out, r_s_root, impact_frame_params=None):
    """Write one batch as ``path/<index>.npy`` in the reference's layout
    (generate_derivs.py:201-204); returns the file name."""
    d = dict(pert_out=_np(pert_out), r_s_root=_np(r_s_root),
             ImpactFrameParams=_np(impact_frame_params))
    fname = os.path.join(path, str(int(index)))
    np.save(fname, d, allow_pickle=True)
    return fname + ".npy"


def load_derivs_batch(path, index, trusted=False):
    """Read a batch written by ``save_derivs_batch``.  The reference's layout is a pickled dict,
    and unpickling runs code stored in the file: pass ``trusted=True`` only for files this code
    (or another trusted writer) produced; otherwise the call raises."""
    fname = os.path.join(path, f"{int(index)}.npy")
    if not trusted:
        raise ValueError(f"{fname} holds a pickled dict (the reference's batch layout); loading "
                         "it runs code stored in the file -- pass trusted=True for a file you "
                         "wrote yourself")
    return np.load(fname, allow_pickle=True).item()


def generate_derivs_batches(potential_base, BaseStreamModel, batches, path, func=HernquistPotential,
                            t_window=150.0, save_iter_start=0, solver=None, units=None, **tol):
    """The reference's batch loop (generate_derivs.py:153-204) with caller-supplied subhalos.

    batches: iterable of dicts with ``r_s`` (Nsh,), ``x0`` (Nsh, 3), ``v`` (Nsh, 3), ``t0``
    (Nsh,) and optionally ``t_window`` and ``ImpactFrameParams``; each becomes unit-mass
    straight-line subhalos of profile ``func`` (Hernquist as in the reference, or Plummer),
    their mass and radius responses are integrated with ``solver`` (default: the base model's)
    and the batch is saved as ``path/<save_iter_start + i>.npy``.  Returns the file names."""
    if func not in (HernquistPotential, PlummerPotential):
        raise NotImplementedError("subhalo profile must be HernquistPotential or PlummerPotential")
    os.makedirs(path, exist_ok=True)
    names = []
    for i, b in enumerate(batches):
        rs = np.asarray(b["r_s"], dtype=np.float64).reshape(-1)
        nsh = rs.shape[0]
        tw = np.broadcast_to(np.asarray(b.get("t_window", t_window), dtype=np.float64),
                             (nsh,)).copy()
        args = (func, np.ones(nsh), rs, np.asarray(b["x0"], dtype=np.float64).reshape(nsh, 3),
                np.asarray(b["v"], dtype=np.float64).reshape(nsh, 3),
                np.asarray(b["t0"], dtype=np.float64).reshape(nsh), tw)
        sub = SubhaloLinePotentialCustom_fromFunc(*args, units=units)
        struct = SubhaloLinePotentialCustom_dRadius_fromFunc(*args, units=units)
        gen = GenerateMassRadiusPerturbation(potential_base, sub, struct,
                                             BaseStreamModel=BaseStreamModel, solver=solver,
                                             units=units, **tol)
        pert_out = gen.compute_perturbation_OTF(solver=solver, **tol)
        names.append(save_derivs_batch(path, save_iter_start + i, pert_out, rs,
                                       b.get("ImpactFrameParams")))
    return names
