"""Host side of the clump-force MLP kernels (3 -> 256 -> 256 -> 256 -> 4, SiLU; BF16 MFMA with
fp32 accumulation).  Parameters, Adam moments and the training workspace are torch tensors
on the GPU; every computation is a call into libssc_mi355x.so.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

IN, H, OUT = 3, 256, 4
SHAPES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)),
          ("b3", (H,)), ("W4", (OUT, H)), ("b4", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def init_params(seed=0):
    """He-uniform weights (last layer halved), zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, s in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / s[1]) * (0.5 if name == "W4" else 1.0)
            parts.append(rng.uniform(-lim, lim, s).reshape(-1))
        else:
            parts.append(np.zeros(int(np.prod(s))))
    return np.concatenate(parts).astype(np.float32)


def _f32(x):
    if isinstance(x, torch.Tensor):
        return x.detach().to(device=_lib.device(), dtype=torch.float32).contiguous()
    return torch.as_tensor(np.asarray(x), dtype=torch.float32, device=_lib.device()).contiguous()


class MLP:
    def __init__(self, params=None, seed=0):
        L = _lib.require_gpu()
        assert L.ssc_mlp_num_params() == NPARAM
        p = init_params(seed) if params is None else params
        self.params = _f32(p).reshape(-1).clone()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grads = torch.zeros_like(self.params)
        self.step = 0
        self._ws, self._ws_n = None, 0
        self.packed = None
        self.packed_vjp = None

    def _workspace(self, n):
        if self._ws is None or self._ws_n < n:
            nbytes = _lib.load().ssc_mlp_workspace_bytes(int(n))
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device=self.params.device)
            self._ws_n = n
        return self._ws

    def forward(self, x):
        """Training-form forward (generic BF16 GEMMs)."""
        x = _f32(x).reshape(-1, IN)
        n = x.shape[0]
        y = torch.empty(n, OUT, dtype=torch.float32, device=x.device)
        ws = self._workspace(n)
        _lib.call("ssc_mlp_forward", _lib.ptr(self.params), _lib.ptr(x), n, _lib.ptr(y),
                  _lib.ptr(ws), ws.numel())
        return y

    def loss_backward(self, x, target, weight=None):
        """Weighted squared error and its gradient (into self.grads).  Returns (loss, y).  The
        minibatch's weight sum stays on the device (no host round trip per iteration)."""
        x = _f32(x).reshape(-1, IN)
        t = _f32(target).reshape(-1, OUT)
        n = x.shape[0]
        y = torch.empty(n, OUT, dtype=torch.float32, device=x.device)
        loss = torch.empty(1, dtype=torch.float32, device=x.device)
        ws = self._workspace(n)
        if weight is None:
            _lib.call("ssc_mlp_loss_backward", _lib.ptr(self.params), _lib.ptr(x), _lib.ptr(t),
                      None, ctypes.c_float(float(n)), n, _lib.ptr(y), _lib.ptr(self.grads),
                      _lib.ptr(loss), _lib.ptr(ws), ws.numel())
        else:
            w = _f32(weight).reshape(-1)
            wsum = w.sum().reshape(1)
            _lib.call("ssc_mlp_loss_backward_dev", _lib.ptr(self.params), _lib.ptr(x),
                      _lib.ptr(t), _lib.ptr(w), _lib.ptr(wsum), n, _lib.ptr(y),
                      _lib.ptr(self.grads), _lib.ptr(loss), _lib.ptr(ws), ws.numel())
        return loss, y

    def vjp(self, x, dy):
        """Gradient of sum(dy * y(x)) with respect to the parameters."""
        x = _f32(x).reshape(-1, IN)
        dy = _f32(dy).reshape(-1, OUT)
        n = x.shape[0]
        g = torch.empty_like(self.params)
        ws = self._workspace(n)
        _lib.call("ssc_mlp_backward", _lib.ptr(self.params), _lib.ptr(x), _lib.ptr(dy), n,
                  _lib.ptr(g), _lib.ptr(ws), ws.numel())
        return g

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, weight_decay=0.0, grads=None):
        self.step += 1
        g = self.grads if grads is None else grads
        _lib.call("ssc_mlp_adam", _lib.ptr(self.params), _lib.ptr(g), _lib.ptr(self.m),
                  _lib.ptr(self.v), NPARAM, ctypes.c_float(lr), ctypes.c_float(beta1),
                  ctypes.c_float(beta2), ctypes.c_float(eps), self.step,
                  ctypes.c_float(weight_decay))
        self.packed = None
        self.packed_vjp = None

    def pack(self):
        if self.packed is None:
            nbytes = _lib.load().ssc_mlp_packed_bytes()
            self.packed = torch.empty(nbytes, dtype=torch.uint8, device=self.params.device)
            _lib.call("ssc_mlp_pack", _lib.ptr(self.params), _lib.ptr(self.packed))
        return self.packed

    def pack_vjp(self):
        if self.packed_vjp is None:
            nbytes = _lib.load().ssc_mlp_packed_vjp_bytes()
            self.packed_vjp = torch.empty(nbytes, dtype=torch.uint8, device=self.params.device)
            _lib.call("ssc_mlp_pack_vjp", _lib.ptr(self.params), _lib.ptr(self.packed_vjp))
        return self.packed_vjp

    def input_vjp(self, x, dy):
        """d(sum(dy * y(x))) / dx with the packed, register-chained kernel (fp32 [n, 3])."""
        x = _f32(x).reshape(-1, IN)
        dy = _f32(dy).reshape(-1, OUT)
        n = x.shape[0]
        dx = torch.empty(n, IN, dtype=torch.float32, device=x.device)
        _lib.call("ssc_mlp_input_vjp", _lib.ptr(self.pack()), _lib.ptr(self.pack_vjp()),
                  _lib.ptr(x), _lib.ptr(dy), n, _lib.ptr(dx))
        return dx

    def infer(self, x):
        """Inference with the packed, register-chained kernel."""
        x = _f32(x).reshape(-1, IN)
        n = x.shape[0]
        y = torch.empty(n, OUT, dtype=torch.float32, device=x.device)
        _lib.call("ssc_mlp_infer", _lib.ptr(self.pack()), _lib.ptr(x), n, _lib.ptr(y))
        return y
