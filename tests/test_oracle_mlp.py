"""Pins of the oracle MLP (CPU only): forward/backward against torch CPU autograd in float64,
central finite differences of the loss, and the BF16 rounding helper against torch's
bfloat16 conversion."""
import numpy as np
import torch

import oracle.mlp as M


def _torch_loss(p, x, t, w):
    P = {}
    o = 0
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    for name, s in M.SHAPES:
        n = int(np.prod(s))
        P[name] = pt[o:o + n].reshape(s)
        o += n
    h = torch.tensor(x, dtype=torch.float64)
    for li in (1, 2, 3):
        h = torch.nn.functional.silu(h @ P[f"W{li}"].T + P[f"b{li}"])
    y = h @ P["W4"].T + P["b4"]
    wt = torch.tensor(w, dtype=torch.float64)
    loss = (wt[:, None] * (y - torch.tensor(t, dtype=torch.float64)) ** 2).sum() / wt.sum()
    loss.backward()
    return loss.item(), pt.grad.numpy(), y.detach().numpy()


def test_forward_backward_against_torch_autograd():
    rng = np.random.default_rng(0)
    p = M.init_params(1).astype(np.float64) + rng.normal(scale=0.01, size=M.NPARAM)
    x = rng.uniform(-1, 1, (64, 3))
    t = rng.normal(size=(64, 4))
    w = rng.uniform(0.5, 2.0, 64)
    lo, go, yo = M.loss_and_grad(p, x, t, w, emulate=False)
    lt, gt, yt = _torch_loss(p, x, t, w)
    np.testing.assert_allclose(yo, yt, rtol=1e-12, atol=1e-14)
    assert abs(lo - lt) <= 1e-12 * abs(lt)
    np.testing.assert_allclose(go, gt, rtol=1e-9, atol=1e-12 * np.abs(gt).max())


def test_loss_gradient_finite_differences():
    rng = np.random.default_rng(1)
    p = M.init_params(2).astype(np.float64)
    x = rng.uniform(-1, 1, (16, 3))
    t = rng.normal(size=(16, 4))
    _, g, _ = M.loss_and_grad(p, x, t, emulate=False)
    for k in rng.choice(M.NPARAM, 12, replace=False):
        e = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += e
        pm[k] -= e
        fd = (M.loss_and_grad(pp, x, t, emulate=False)[0] - M.loss_and_grad(pm, x, t, emulate=False)[0]) / (2 * e)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-3, abs(g[k]))


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(2)
    a = (rng.normal(size=10000) * 10 ** rng.uniform(-5, 5, 10000)).astype(np.float32)
    ref = torch.tensor(a).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(M.bf16(a), ref)


def test_input_vjp_against_torch_and_fd():
    """d(dy . MLP(x))/dx: plain float64 form = torch autograd; the BF16-emulating form stays
    within BF16 rounding of it."""
    rng = np.random.default_rng(4)
    p = M.init_params(2).astype(np.float64) + rng.normal(scale=0.01, size=M.NPARAM)
    x = rng.uniform(-1, 1, (48, 3))
    dy = rng.normal(size=(48, 4))
    P, o = {}, 0
    for name, s in M.SHAPES:
        n = int(np.prod(s))
        P[name] = torch.tensor(p[o:o + n].reshape(s), dtype=torch.float64)
        o += n
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    h = xt
    for li in (1, 2, 3):
        h = torch.nn.functional.silu(h @ P[f"W{li}"].T + P[f"b{li}"])
    y = h @ P["W4"].T + P["b4"]
    (y * torch.tensor(dy)).sum().backward()
    ref = xt.grad.numpy()
    got = M.input_vjp(p, x, dy, emulate=False)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    # central differences of dy . y along each input axis
    e = 1e-6
    for c in range(3):
        xp, xm = x.copy(), x.copy()
        xp[:, c] += e
        xm[:, c] -= e
        fd = ((M.forward(p, xp, emulate=False)[0] - M.forward(p, xm, emulate=False)[0]) * dy).sum(1) / (2 * e)
        np.testing.assert_allclose(got[:, c], fd, rtol=1e-6, atol=1e-8)
    em = M.input_vjp(p.astype(np.float32), x.astype(np.float32), dy, emulate=True)
    assert np.linalg.norm(em - ref) / np.linalg.norm(ref) < 2e-2
