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
                     "ssc_mlp_packed_bytes", "ssc_mlp_packed_vjp_bytes", "ssc_mlp_workspace_bytes",
                     "ssc_perturb_workspace_bytes",
                     "ssc_adjoint_workspace_bytes", "ssc_surrogate_adjoint_workspace_bytes"):
            assert n in _lib.SIGNATURES, n


_LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "ssc_mi355x.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("ssc_component %zu\nssc_host_potential %zu\nssc_sources %zu\n"
         "ssc_dopri5_options %zu\n", sizeof(ssc_component), sizeof(ssc_host_potential),
         sizeof(ssc_sources), sizeof(ssc_dopri5_options));
  F(ssc_sources, group_size) F(ssc_sources, gm) F(ssc_sources, t_window) F(ssc_sources, group)
  F(ssc_dopri5_options, dt0) F(ssc_dopri5_options, max_steps) F(ssc_dopri5_options, stops)
  F(ssc_host_potential, c) F(ssc_component, p)
  return 0;
}
"""


def test_struct_layout_matches_header(tmp_path):
    """The ctypes mirrors have the C compiler's sizes and offsets for every descriptor."""
    import subprocess
    from streamsculptor_amd import _lib
    src = tmp_path / "layout.c"
    src.write_text(_LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True,
                                                       capture_output=True, text=True)
               .stdout.splitlines())
    mirror = {"ssc_component": _lib.Component, "ssc_host_potential": _lib.HostPotential,
              "ssc_sources": _lib.Sources, "ssc_dopri5_options": _lib.Dopri5Options}
    for name, cls in mirror.items():
        assert ctypes.sizeof(cls) == int(got[name]), name
    for key, val in got.items():
        if "." in key:
            struct, field = key.split(".")
            assert getattr(mirror[struct], field).offset == int(val), key


def test_product_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    import streamsculptor_amd as S
    from streamsculptor_amd.potential import NFWPotential
    with pytest.raises(RuntimeError, match="GPU"):
        NFWPotential(1e12, 20.0, units=S.usys).gradient([1.0, 2.0, 3.0], 0.0)
