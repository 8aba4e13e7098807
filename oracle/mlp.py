"""numpy FP64 restatement of the learned proposal's MLP -- TEST INFRASTRUCTURE.

Only ``tests/`` and ``bench.py``'s checks may import this module.  It mirrors
the device network (cuda_pspin_amd/csrc/mlp.hip, api_learned.cpp): flat FP32
parameters [W1 (hid x in_pad), b1, W2 (hid x hid), b2, W3 (out x hid), b3],
SiLU hidden layers, a Gaussian negative log-likelihood head (mean = u +
out[:d], log-scale = clamp(out[d:] + ls0)), and the BF16 rounding points of
the device passes when ``emulate`` is set:

  forward   X0 = bf16(x); Z1 = X0 bf16(W1)^T + b1; H1 = bf16(silu(Z1)); ...
  backward  dO, dZ2, dZ1 rounded to BF16 before each product; weight and
            bias gradients accumulated from the rounded operands.

With ``emulate=False`` everything is plain FP64, which the finite-difference
and autograd checks in tests/test_oracle_learned.py pin.  This is this
build's own design (the reference has no learned proposal).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def bf16(x) -> np.ndarray:
    """Round to BF16 (nearest even) via FP32, returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


@dataclass
class Dims:
    N: int
    hid: int

    @property
    def dim(self):
        return 2 * self.N

    @property
    def in_dim(self):
        return 2 * self.N + 1

    @property
    def in_pad(self):
        return (self.in_dim + 63) // 64 * 64

    @property
    def out_dim(self):
        return 4 * self.N

    def offsets(self):
        w1 = 0
        b1 = w1 + self.hid * self.in_pad
        w2 = b1 + self.hid
        b2 = w2 + self.hid * self.hid
        w3 = b2 + self.hid
        b3 = w3 + self.out_dim * self.hid
        return dict(w1=w1, b1=b1, w2=w2, b2=b2, w3=w3, b3=b3, n=b3 + self.out_dim)

    @property
    def n_params(self):
        return self.offsets()["n"]


def unpack(p, d: Dims):
    o = d.offsets()
    p = np.asarray(p, dtype=np.float64)
    return dict(W1=p[o["w1"]:o["b1"]].reshape(d.hid, d.in_pad), b1=p[o["b1"]:o["w2"]],
                W2=p[o["w2"]:o["b2"]].reshape(d.hid, d.hid), b2=p[o["b2"]:o["w3"]],
                W3=p[o["w3"]:o["b3"]].reshape(d.out_dim, d.hid), b3=p[o["b3"]:o["n"]])


def pack(g, d: Dims):
    o = d.offsets()
    out = np.zeros(o["n"])
    out[o["w1"]:o["b1"]] = g["W1"].reshape(-1)
    out[o["b1"]:o["w2"]] = g["b1"]
    out[o["w2"]:o["b2"]] = g["W2"].reshape(-1)
    out[o["b2"]:o["w3"]] = g["b2"]
    out[o["w3"]:o["b3"]] = g["W3"].reshape(-1)
    out[o["b3"]:o["n"]] = g["b3"]
    return out


def forward(p, d: Dims, x, emulate=True):
    """x [M][in_dim] -> (O [M][out_dim], cache)."""
    P = unpack(p, d)
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    x = np.asarray(x, dtype=np.float64)
    X0 = np.zeros((x.shape[0], d.in_pad))
    X0[:, :d.in_dim] = r(x)
    W1, W2, W3 = r(P["W1"]), r(P["W2"]), r(P["W3"])
    Z1 = X0 @ W1.T + P["b1"]
    H1 = r(silu(Z1))
    Z2 = H1 @ W2.T + P["b2"]
    H2 = r(silu(Z2))
    O = H2 @ W3.T + P["b3"]
    return O, dict(X0=X0, Z1=Z1, H1=H1, Z2=Z2, H2=H2, W1=W1, W2=W2, W3=W3)


def nll(O, U, Y, d: Dims, ls0, ls_min, ls_max):
    """Mean Gaussian NLL per row and its gradient w.r.t. O (scaled by 1/M)."""
    M = O.shape[0]
    D = d.dim
    mu = np.asarray(U, dtype=np.float64) + O[:, :D]
    raw = O[:, D:] + ls0
    ls = np.clip(raw, ls_min, ls_max)
    isg = np.exp(-ls)
    rr = (np.asarray(Y, dtype=np.float64) - mu) * isg
    loss = np.sum(0.5 * rr * rr + ls) / M
    dO = np.zeros_like(O)
    dO[:, :D] = -rr * isg / M
    inside = (raw > ls_min) & (raw < ls_max)
    dO[:, D:] = np.where(inside, (1.0 - rr * rr) / M, 0.0)
    return loss, dO


def backward(d: Dims, cache, dO, emulate=True):
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    dOb = r(dO)
    g = {}
    g["W3"] = dOb.T @ cache["H2"]
    g["b3"] = dOb.sum(axis=0)
    dZ2 = r((dOb @ cache["W3"]) * dsilu(cache["Z2"]))
    g["W2"] = dZ2.T @ cache["H1"]
    g["b2"] = dZ2.sum(axis=0)
    dZ1 = r((dZ2 @ cache["W2"]) * dsilu(cache["Z1"]))
    g["W1"] = dZ1.T @ cache["X0"]
    g["b1"] = dZ1.sum(axis=0)
    return pack(g, d)


def loss_grad(p, d: Dims, x, U, Y, ls0, ls_min, ls_max, emulate=True):
    O, cache = forward(p, d, x, emulate)
    if emulate:
        O = O.astype(np.float32).astype(np.float64)  # the device keeps O in FP32
    loss, dO = nll(O, U, Y, d, ls0, ls_min, ls_max)
    return loss, backward(d, cache, dO, emulate)


def adam(p, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) in FP64; returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v


def init_params(d: Dims, rng: np.random.Generator, gain=1.0):
    """Random parameters in the device layout (padding columns of W1 zero)."""
    P = dict(W1=np.zeros((d.hid, d.in_pad)), b1=rng.uniform(-1, 1, d.hid) / np.sqrt(d.in_dim),
             W2=rng.uniform(-1, 1, (d.hid, d.hid)) / np.sqrt(d.hid), b2=rng.uniform(-1, 1, d.hid) / np.sqrt(d.hid),
             W3=gain * rng.uniform(-1, 1, (d.out_dim, d.hid)) / np.sqrt(d.hid),
             b3=gain * rng.uniform(-1, 1, d.out_dim) / np.sqrt(d.hid))
    P["W1"][:, :d.in_dim] = rng.uniform(-1, 1, (d.hid, d.in_dim)) / np.sqrt(d.in_dim)
    return pack(P, d).astype(np.float32)
