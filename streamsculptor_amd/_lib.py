"""ctypes binding of libssc_mi355x.so (the C-ABI declared in include/ssc_mi355x.h).

The library is the only compute path: if it is missing, or no ROCm GPU is visible, every
entry point raises instead of falling back to another implementation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libssc_mi355x.so")

G_USYS = 4.498502151469553e-12  # G in kpc^3 / (Msun Myr^2)

COMP_MIYAMOTO_NAGAI, COMP_NFW, COMP_HERNQUIST, COMP_PLUMMER = 1, 2, 3, 4
SRC_NONE, SRC_STATIC, SRC_TRACK, SRC_LINE, SRC_LINE_HERNQUIST, SRC_TRACK_HERNQUIST = 0, 1, 2, 3, 4, 5
TRACK_KINDS = (SRC_TRACK, SRC_TRACK_HERNQUIST)
MAX_COMPONENTS = 8
MAX_SOURCES = 1024


class Component(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("p", ctypes.c_double * 4)]


class HostPotential(ctypes.Structure):
    _fields_ = [("n_components", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("c", Component * MAX_COMPONENTS)]


class Sources(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("n", ctypes.c_int32), ("n_track", ctypes.c_int32),
                ("group_size", ctypes.c_int32),
                ("gm", ctypes.c_void_p), ("rs", ctypes.c_void_p), ("pos", ctypes.c_void_p),
                ("vel", ctypes.c_void_p), ("t0", ctypes.c_void_p), ("t_window", ctypes.c_void_p),
                ("group", ctypes.c_void_p)]


class Dopri5Options(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("atol", ctypes.c_double), ("dtmin", ctypes.c_double),
                ("dtmax", ctypes.c_double), ("dt0", ctypes.c_double),
                ("max_steps", ctypes.c_int32), ("n_stops", ctypes.c_int32),
                ("stops", ctypes.c_void_p)]


_vp, _i64, _i32, _d = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double

# name -> argtypes; every entry point returns int (0 = success)
SIGNATURES = {
    "ssc_eval": [_vp, _vp, _vp, _i64, _d, _i32, _vp, _vp, _vp, _vp],
    "ssc_leapfrog": [_vp, _vp, _vp, _vp, _i64, _d, _d, _i32, _vp, _vp, _i32, _vp, _vp],
    "ssc_dopri5": [_vp, _vp, _vp, _vp, _i64, _d, _d, _vp, _i32, _vp, _vp, _vp, _vp, _vp],
    "ssc_dopri5_dense": [_vp, _vp, _vp, _vp, _i64, _d, _d, _vp, _vp, _vp, _i32, _vp],
    "ssc_dopri5_orbits": [_vp, _vp, _vp, _vp, _vp, _i64, _vp, _i32, _vp, _vp, _vp, _vp],
    "ssc_perturb_leapfrog": [_vp, _vp, _vp, _vp, _vp, _i64, _d, _d, _i32, _i32, _vp, _vp, _vp,
                             _i64],
    "ssc_perturb_dopri5": [_vp, _vp, _vp, _vp, _vp, _i64, _d, _d, _vp, _i32, _i32, _vp, _vp, _vp,
                           _vp, _vp, _vp],
    "ssc_mlp_forward": [_vp, _vp, _i64, _vp, _vp, _i64],
    "ssc_mlp_loss_backward": [_vp, _vp, _vp, _vp, ctypes.c_float, _i64, _vp, _vp, _vp, _vp, _i64],
    "ssc_mlp_loss_backward_dev": [_vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _i64],
    "ssc_mlp_backward": [_vp, _vp, _vp, _i64, _vp, _vp, _i64],
    "ssc_mlp_adam": [_vp, _vp, _vp, _vp, _i64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                     ctypes.c_float, _i32, ctypes.c_float],
    "ssc_mlp_pack": [_vp, _vp],
    "ssc_mlp_infer": [_vp, _vp, _i64, _vp],
    "ssc_mlp_pack_vjp": [_vp, _vp],
    "ssc_mlp_input_vjp": [_vp, _vp, _vp, _vp, _i64, _vp],
    "ssc_surrogate_leapfrog": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _d, _d, _i32, _vp, _vp],
    "ssc_release_seeds": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, ctypes.c_int64, _i32, _vp, _vp,
                          _vp],
    "ssc_release_vjp": [_vp, _vp, _vp, _vp, _i32, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                        _vp],
    "ssc_release_vjp_noalloc": [_vp, _vp, _vp, _vp, _i32, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp,
                                _vp, _vp],
    "ssc_adjoint_forward": [_vp, _vp, _vp, _vp, _i64, _d, _d, _i32, _i32, _vp, _vp, _i64],
    "ssc_adjoint_checkpoints_from_rows": [_vp, _i64, _i32, _i32, _i32, _vp, _i64],
    "ssc_adjoint_backward": [_vp, _vp, _vp, _i64, _d, _d, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp,
                             _vp, _i64],
    "ssc_adjoint_backward_range": [_vp, _vp, _vp, _i64, _d, _d, _i32, _i32, _i32, _i32, _vp, _vp,
                                   _vp, _vp, _vp, _vp, _vp, _i64],
    "ssc_release": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, ctypes.c_int64, _vp, _vp, _vp, _vp,
                    _vp],
    "ssc_surrogate_adjoint_forward": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _d, _d, _i32, _i32,
                                      _vp, _vp, _vp, _i64],
    "ssc_surrogate_adjoint_backward": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _d, _i32, _i32,
                                       _vp, _vp, _vp, _vp, _vp, _vp, _i64],
}

_lib = None


class ExtensionMissing(RuntimeError):
    pass


def load():
    """Load the shared library (no GPU needed to load it)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ExtensionMissing(
                f"{LIB_PATH} is not built: run `python -m streamsculptor_amd.build` "
                "(hipcc, gfx950).  There is no CPU fallback.")
        L = ctypes.CDLL(LIB_PATH)
        L.ssc_last_error.restype = ctypes.c_char_p
        L.ssc_version.restype = ctypes.c_char_p
        for name in ("ssc_mlp_num_params", "ssc_mlp_packed_bytes", "ssc_mlp_packed_vjp_bytes"):
            getattr(L, name).restype = ctypes.c_int64
            getattr(L, name).argtypes = []
        L.ssc_mlp_workspace_bytes.restype = ctypes.c_int64
        L.ssc_mlp_workspace_bytes.argtypes = [ctypes.c_int64]
        L.ssc_surrogate_adjoint_workspace_bytes.restype = ctypes.c_int64
        L.ssc_surrogate_adjoint_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int32,
                                                            ctypes.c_int32]
        L.ssc_perturb_workspace_bytes.restype = ctypes.c_int64
        L.ssc_perturb_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int32]
        L.ssc_adjoint_workspace_bytes.restype = ctypes.c_int64
        L.ssc_adjoint_workspace_bytes.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                                  ctypes.c_int32]
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("streamsculptor_amd needs a ROCm GPU (gfx950); none is visible. "
                           "There is no CPU fallback.")
    return load()


def device():
    require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def call(name, *args):
    """Call a C-ABI entry point on the current torch stream; raise on a non-zero status."""
    L = require_gpu()
    rc = getattr(L, name)(*args, stream_handle())
    if rc != 0:
        raise RuntimeError(f"{name} failed ({rc}): {L.ssc_last_error().decode()}")


class _Ptr(ctypes.c_void_p):
    """A device pointer that keeps its tensor alive for as long as the pointer object lives,
    i.e. until the ctypes call it is passed to returns.  Without this, a temporary built inside
    the argument list is freed as soon as its address is taken and the caching allocator can
    hand the same block to the next temporary.  (Reuse after the call is safe: the allocator
    orders it after the work already enqueued on the stream.)"""


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    p = _Ptr(t.data_ptr())
    p._keep = t
    return p


def f64(x, dev=None):
    """Any array-like -> contiguous float64 tensor on the GPU."""
    dev = dev or device()
    if isinstance(x, torch.Tensor):
        return x.detach().to(device=dev, dtype=torch.float64).contiguous()
    if isinstance(x, np.ndarray) and not x.flags.writeable:
        x = np.array(x, dtype=np.float64)  # read-only views (broadcasts): copy on the host first
    return torch.as_tensor(x, dtype=torch.float64, device=dev).contiguous()


def i32(x, dev=None):
    dev = dev or device()
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.int32).contiguous()
    return torch.as_tensor(x, dtype=torch.int32, device=dev).contiguous()
