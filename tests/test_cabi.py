"""The C-ABI library loads and exports every entry point declared in include/*.h (CPU only:
no compute calls)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s*(ssc_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared()
    for n in ("ssc_eval", "ssc_leapfrog", "ssc_release", "ssc_last_error", "ssc_version"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from streamsculptor_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from streamsculptor_amd.build import build
        build(verbose=False)
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"
    L.ssc_version.restype = ctypes.c_char_p
    assert b"gfx950" in L.ssc_version()
    # every declared compute entry point has a ctypes signature in the binding
    for n in _declared():
        if n not in ("ssc_last_error", "ssc_version", "ssc_mlp_num_params",
                     "ssc_mlp_packed_bytes", "ssc_mlp_workspace_bytes",
                     "ssc_adjoint_workspace_bytes"):
            assert n in _lib.SIGNATURES, n


def test_struct_layout_matches_header():
    from streamsculptor_amd import _lib
    assert ctypes.sizeof(_lib.Component) == 40
    assert ctypes.sizeof(_lib.HostPotential) == 8 + 8 * 40
    assert ctypes.sizeof(_lib.Sources) == 16 + 6 * 8


def test_product_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import streamsculptor_amd as S
    from streamsculptor_amd.potential import NFWPotential
    with pytest.raises(RuntimeError, match="GPU"):
        NFWPotential(1e12, 20.0, units=S.usys).gradient([1.0, 2.0, 3.0], 0.0)
