"""GPU parity of the BF16 MFMA MLP kernels against the BF16-emulating oracle.

Tolerances: both sides round the same operands to BF16 but accumulate in different orders, so
isolated activations can round to neighbouring BF16 values; outputs and gradients are compared
normwise at 2e-3 relative, Adam (pure fp32 elementwise) at 1e-6.
"""
import numpy as np
import pytest
import torch

import oracle.mlp as OM

pytestmark = pytest.mark.gpu


def _nrel(a, b):
    a = a.detach().cpu().double().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _case(n=3000, seed=0):
    rng = np.random.default_rng(seed)
    p = OM.init_params(seed)
    p = p + rng.normal(scale=0.02, size=p.shape).astype(np.float32)  # non-zero biases
    x = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    t = rng.normal(size=(n, 4)).astype(np.float32)
    w = rng.uniform(0.2, 2.0, n).astype(np.float32)
    return p, x, t, w


def test_forward_generic_and_tile_match_oracle(ssc):
    from streamsculptor_amd.mlp import MLP
    p, x, _, _ = _case()
    net = MLP(params=p)
    y_ref, _ = OM.forward(p, x, emulate=True)
    yg = net.forward(x)
    yt = net.infer(x)
    e1, e2, e3 = _nrel(yg, y_ref), _nrel(yt, y_ref), _nrel(yt, yg.double().cpu().numpy())
    print(f"mlp forward: generic {e1:.2e} tile {e2:.2e} tile-vs-generic {e3:.2e}")
    assert e1 < 2e-3 and e2 < 2e-3 and e3 < 2e-3
    # ragged batch (not a multiple of the 32-point tile)
    yr = net.infer(x[:77])
    assert _nrel(yr, y_ref[:77]) < 2e-3


def test_loss_and_gradients_match_oracle(ssc):
    from streamsculptor_amd.mlp import MLP
    p, x, t, w = _case(n=4096, seed=1)
    net = MLP(params=p)
    loss, y = net.loss_backward(x, t, w)
    lo, go, yo = OM.loss_and_grad(p, x, t, w, emulate=True)
    assert abs(loss.item() - lo) <= 1e-3 * abs(lo)
    g = net.grads.cpu().double().numpy()
    off = 0
    for name, s in OM.SHAPES:
        k = int(np.prod(s))
        e = np.linalg.norm(g[off:off + k] - go[off:off + k]) / max(np.linalg.norm(go[off:off + k]), 1e-30)
        print(f"grad {name}: {e:.2e}")
        assert e < 2e-3, name
        off += k
    # the vector-Jacobian entry point gives the same gradient for dy = dL/dy
    wsum = w.sum()
    dy = 2 * w[:, None] * (yo - t) / wsum
    gv = net.vjp(x, dy.astype(np.float32)).cpu().double().numpy()
    assert np.linalg.norm(gv - go) / np.linalg.norm(go) < 3e-3


def test_device_weight_sum_equals_host_weight_sum(ssc):
    """ssc_mlp_loss_backward_dev (the minibatch weight sum read from device memory) and
    ssc_mlp_loss_backward (the same sum passed by value) give the same loss and gradients,
    bit for bit: a training loop can drop the per-iteration host round trip."""
    import ctypes
    from streamsculptor_amd import _lib
    from streamsculptor_amd.mlp import MLP
    p, x, t, w = _case(n=4096, seed=2)
    net = MLP(params=p)
    loss_d, y_d = net.loss_backward(x, t, w)
    g_d = net.grads.clone()
    xt, tt, wt = (torch.as_tensor(a, device="cuda") for a in (x, t, w))
    wsum = float(wt.sum().item())
    y = torch.empty_like(y_d)
    loss = torch.empty_like(loss_d)
    ws = net._workspace(4096)
    _lib.call("ssc_mlp_loss_backward", _lib.ptr(net.params), _lib.ptr(xt), _lib.ptr(tt),
              _lib.ptr(wt), ctypes.c_float(wsum), 4096, _lib.ptr(y), _lib.ptr(net.grads),
              _lib.ptr(loss), _lib.ptr(ws), ws.numel())
    assert torch.equal(loss, loss_d) and torch.equal(y, y_d) and torch.equal(net.grads, g_d)


def test_adam_matches_oracle(ssc):
    from streamsculptor_amd.mlp import MLP
    rng = np.random.default_rng(5)
    p = OM.init_params(3)
    net = MLP(params=p)
    m = np.zeros_like(p, dtype=np.float64)
    v = np.zeros_like(p, dtype=np.float64)
    pr = p.astype(np.float64)
    for step in range(1, 4):
        g = rng.normal(size=p.shape).astype(np.float32)
        net.grads.copy_(torch.as_tensor(g, device="cuda"))
        net.adam(1e-3, weight_decay=1e-4)
        pr, m, v = OM.adam(pr, g.astype(np.float64), m, v, step, 1e-3, wd=1e-4)
    np.testing.assert_allclose(net.params.cpu().double().numpy(), pr, rtol=1e-5, atol=1e-7)


def test_training_reduces_loss(ssc):
    """A few hundred Adam steps fit a smooth target (the kernels train the network)."""
    from streamsculptor_amd.mlp import MLP
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, (8192, 3)).astype(np.float32)
    r = np.linalg.norm(x, axis=1, keepdims=True)
    t = np.hstack([-1 / np.sqrt(r ** 2 + 0.3), x / (r ** 2 + 0.3) ** 1.5]).astype(np.float32)
    net = MLP(seed=0)
    xt, tt = torch.as_tensor(x, device="cuda"), torch.as_tensor(t, device="cuda")
    first = None
    for it in range(300):
        idx = torch.randint(0, 8192, (1024,), device="cuda")
        loss, _ = net.loss_backward(xt[idx], tt[idx])
        net.adam(2e-3)
        if first is None:
            first = loss.item()
    y = net.infer(x)
    final = float(((y.cpu().numpy() - t) ** 2).sum(1).mean())
    print(f"training: first batch loss {first:.3e}, final full loss {final:.3e}")
    assert final < 0.05 * first


def test_input_vjp_matches_oracle(ssc):
    """Register-chained input VJP (packed W4^T, W3^T, W2^T) against the BF16-emulating oracle:
    random cotangents, each output alone (one Jacobian row per point: a transposed or permuted
    W4^T fragment would mix the rows), and a ragged batch."""
    from streamsculptor_amd.mlp import MLP
    p, x, _, _ = _case(n=2500, seed=4)
    rng = np.random.default_rng(5)
    net = MLP(params=p)
    dy = rng.normal(size=(len(x), 4)).astype(np.float32)
    ref = OM.input_vjp(p, x, dy, emulate=True)
    got = net.input_vjp(x, dy)
    e = _nrel(got, ref)
    print(f"input vjp: {e:.2e}")
    assert e < 2e-3
    for k in range(4):
        ek = np.zeros_like(dy)
        ek[:, k] = 1.0
        ek_ref = OM.input_vjp(p, x, ek, emulate=True)
        ek_got = net.input_vjp(x, ek)
        assert _nrel(ek_got, ek_ref) < 2e-3, k
    assert _nrel(net.input_vjp(x[:77], dy[:77]), ref[:77]) < 2e-3
