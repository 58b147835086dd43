"""ORACLE -- test infrastructure only: numpy restatement of the clump-force MLP
(3 -> 256 -> 256 -> 256 -> 4, SiLU), its weighted squared-error loss, backward pass and Adam
step, with optional emulation of the BF16 rounding points of the GPU kernels (weights and
hidden activations rounded to nearest-even BF16; layer-1 inputs fp32).

Pinned by tests/test_oracle_mlp.py against torch CPU autograd in float64 (no rounding) and
central finite differences of the loss.  The reference has no network: this follows the
surrogate specification of the stream workload (MLP (x, y, z) -> (Phi, grad Phi) of a clump
sum).
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 3, 256, 4
SHAPES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)),
          ("b3", (H,)), ("W4", (OUT, H)), ("b4", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unflatten(p):
    out, o = {}, 0
    for name, s in SHAPES:
        n = int(np.prod(s))
        out[name] = p[o:o + n].reshape(s)
        o += n
    return out


def flatten(d):
    return np.concatenate([np.asarray(d[name]).reshape(-1) for name, _ in SHAPES])


def bf16(a):
    """Round float32 values to the nearest-even BF16 (kept as float32)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def init_params(seed=0):
    """He-style uniform init (the GPU trainer uses the same)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, s in SHAPES:
        if name.startswith("W"):
            fan_in = s[1]
            lim = np.sqrt(6.0 / fan_in) * (0.5 if name == "W4" else 1.0)
            d[name] = rng.uniform(-lim, lim, s)
        else:
            d[name] = np.zeros(s)
    return flatten(d).astype(np.float32)


def forward(p, x, emulate=True):
    """Returns (y, cache).  emulate=True rounds like the GPU (BF16 weights/activations, fp32
    layer-1 inputs); emulate=False is plain float64."""
    P = unflatten(np.asarray(p, dtype=np.float64 if not emulate else np.float32))
    rnd = bf16 if emulate else (lambda a: a)
    f = np.float64
    x = np.asarray(x, dtype=np.float32 if emulate else np.float64).astype(f)
    W1 = rnd(P["W1"]).astype(f)
    Z1 = x @ W1.T + P["b1"].astype(f)
    if emulate:
        Z1 = Z1.astype(np.float32).astype(f)
    A1 = rnd(silu(Z1).astype(np.float32) if emulate else silu(Z1)).astype(f)
    cache = {"x": x, "Z": [Z1], "A": [A1]}
    A = A1
    for Wn, bn in (("W2", "b2"), ("W3", "b3")):
        Z = A @ rnd(P[Wn]).astype(f).T + P[bn].astype(f)
        if emulate:
            Z = Z.astype(np.float32).astype(f)
        A = rnd(silu(Z).astype(np.float32) if emulate else silu(Z)).astype(f)
        cache["Z"].append(Z)
        cache["A"].append(A)
    y = A @ rnd(P["W4"]).astype(f).T + P["b4"].astype(f)
    return y, cache


def loss_and_grad(p, x, target, weight=None, emulate=True):
    """Weighted squared error sum_b w_b |y_b - t_b|^2 / sum_b w_b and its gradient (flat)."""
    y, c = forward(p, x, emulate)
    f = np.float64
    P = unflatten(np.asarray(p, dtype=np.float64 if not emulate else np.float32))
    rnd = bf16 if emulate else (lambda a: a)
    n = y.shape[0]
    w = np.ones(n) if weight is None else np.asarray(weight, dtype=f)
    ws = w.sum()
    d = y - np.asarray(target, dtype=f)
    loss = float((w[:, None] * d * d).sum() / ws)
    dY = 2.0 * w[:, None] * d / ws
    g = {}
    dYr = rnd(dY.astype(np.float32)).astype(f) if emulate else dY
    g["W4"] = dYr.T @ c["A"][2]
    g["b4"] = dY.sum(0)
    dA = dYr @ rnd(P["W4"]).astype(f)
    for li, (Wn, bn) in ((2, ("W3", "b3")), (1, ("W2", "b2"))):
        dZ = dA * silu_grad(c["Z"][li])
        dZr = rnd(dZ.astype(np.float32)).astype(f) if emulate else dZ
        g[Wn] = dZr.T @ c["A"][li - 1]
        g[bn] = dZ.sum(0)
        dA = dZr @ rnd(P[Wn]).astype(f)
    dZ1 = dA * silu_grad(c["Z"][0])
    g["W1"] = dZ1.T @ c["x"]
    g["b1"] = dZ1.sum(0)
    return loss, flatten(g), y


def input_vjp(p, x, dy, emulate=True, return_y=False):
    """dx = d(dy . y)/dx for y = MLP(x): the vector-Jacobian product with respect to the inputs
    (used by the adjoint through the surrogate).  emulate=True rounds where the GPU kernel does:
    BF16 weights, fp32 output cotangents (the W4^T product runs in fp32), SiLU derivatives
    stored in BF16, BF16 hidden-layer cotangents between the matrix products; the last product
    with the fp32 (BF16-valued) W1.  return_y=True also returns the forward output."""
    y, c = forward(p, x, emulate)
    f = np.float64
    P = unflatten(np.asarray(p, dtype=np.float64 if not emulate else np.float32))
    rnd = bf16 if emulate else (lambda a: a)
    dy = np.asarray(dy, dtype=np.float32 if emulate else f).astype(f)
    dA = dy @ rnd(P["W4"]).astype(f)
    for li, Wn in ((2, "W3"), (1, "W2")):
        sg = rnd(silu_grad(c["Z"][li]).astype(np.float32)).astype(f) if emulate else silu_grad(c["Z"][li])
        dZ = dA * sg
        if emulate:
            dZ = rnd(dZ.astype(np.float32)).astype(f)
        dA = dZ @ rnd(P[Wn]).astype(f)
    sg = rnd(silu_grad(c["Z"][0]).astype(np.float32)).astype(f) if emulate else silu_grad(c["Z"][0])
    dZ1 = dA * sg
    dx = dZ1 @ rnd(P["W1"]).astype(f)
    return (dx, y) if return_y else dx


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    p = p - lr * ((m / c1) / (np.sqrt(v / c2) + eps) + wd * p)
    return p, m, v
