"""ssc_release_vjp_noalloc (static device scratch, no allocation) against ssc_release_vjp (caller
workspace), called through the C-ABI (collected last: it must pass alone and after the others): two calls in a row on one GPU, the second with fewer
releases, must give the same dL/d(progenitor rows) and dL/d(host parameters) bit for bit (same
kernels, same reduction order); a size above the static maximum is refused."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

G = 4.498502151469553e-12


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    r = rng.uniform(6.0, 25.0, n)
    u = rng.normal(size=(n, 3))
    x = r[:, None] * u / np.linalg.norm(u, axis=1, keepdims=True)
    v = rng.normal(scale=0.15, size=(n, 3))
    dev = "cuda"
    w = torch.as_tensor(np.hstack([x, v]), dtype=torch.float64, device=dev).contiguous()
    idx = torch.arange(n, dtype=torch.int32, device=dev)
    gms = torch.full((n,), G * 2e4, dtype=torch.float64, device=dev)
    cot = [torch.as_tensor(rng.normal(size=(n, 3)), dtype=torch.float64, device=dev)
           for _ in range(4)]
    return w, idx, gms, cot


def _call_both(L, h, n, seed):
    from streamsculptor_amd import _lib
    w, idx, gms, cot = _inputs(n, seed)
    nc = h.n_components
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    outs = {}
    for name in ("ssc_release_vjp", "ssc_release_vjp_noalloc"):
        g_prog = torch.full((n, 6), np.nan, dtype=torch.float64, device="cuda")
        g_host = torch.full((nc, 2), np.nan, dtype=torch.float64, device="cuda")
        args = [ctypes.byref(h), w.data_ptr(), idx.data_ptr(), gms.data_ptr(), n, 7, None]
        args += [c.data_ptr() for c in cot] + [g_prog.data_ptr(), g_host.data_ptr()]
        if name == "ssc_release_vjp":
            scratch = torch.empty(n, 2 * nc, dtype=torch.float64, device="cuda")
            args.append(scratch.data_ptr())
        rc = getattr(L, name)(*args, stream)
        assert rc == 0, L.ssc_last_error().decode()
        torch.cuda.synchronize()
        outs[name] = (g_prog.cpu().numpy(), g_host.cpu().numpy())
    return outs


def test_noalloc_matches_workspace_entry_point(ssc):
    from streamsculptor_amd import _lib
    from streamsculptor_amd import workloads as W
    L = _lib.require_gpu()
    h, _ = W.host_potential()._descriptor()
    for n, seed in ((5000, 1), (700, 2)):  # the second call is smaller than the first
        print(f"[no-alloc] release VJP, nrel = {n}", flush=True)
        o = _call_both(L, h, n, seed)
        a, b = o["ssc_release_vjp"], o["ssc_release_vjp_noalloc"]
        assert np.all(np.isfinite(a[0])) and np.all(np.isfinite(a[1]))
        np.testing.assert_array_equal(b[0], a[0])
        np.testing.assert_array_equal(b[1], a[1])


def test_noalloc_rejects_sizes_above_its_maximum(ssc):
    from streamsculptor_amd import _lib
    from streamsculptor_amd import workloads as W
    L = _lib.require_gpu()
    h, _ = W.host_potential()._descriptor()
    t = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = t.data_ptr()
    rc = L.ssc_release_vjp_noalloc(ctypes.byref(h), p, None, p, 65537, 0, None, p, p, p, p, p, p,
                                   None)
    assert rc == -4
    assert "exceeds the static scratch" in L.ssc_last_error().decode()
