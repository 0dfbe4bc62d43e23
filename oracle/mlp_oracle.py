"""FP64 numpy restatement of the learned filter network and its loss.

TEST INFRASTRUCTURE ONLY (imported by tests/ and bench.py's checks, never by
the product).  The network has no counterpart in the reference; this file is
the independent definition the device kernels (mdet_amd/csrc/tnf_mlp.hip) are
checked against:
  x' = (tnf - mean) * istd              (NaN inputs -> 0 before scaling)
  h1 = gelu(W1 x' + b1), h2 = gelu(W2 h1 + b2), e = W3 h2 + b3
  gelu(z) = z * Phi(z)   (exact, erf-based)
  D = |e_i - e_j|^2,  L = mean_b  y*D + (1-y)*relu(margin - D)
Weights enter the forward pass rounded to BF16 (round to nearest even), as
on the device; biases stay FP32.  The analytic gradient is pinned by central
finite differences on unrounded weights (tests/test_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np

D_IN, D_H, D_E = 136, 256, 32
SHAPES = [("W1", (D_H, D_IN)), ("b1", (D_H,)), ("W2", (D_H, D_H)), ("b2", (D_H,)), ("W3", (D_E, D_H)),
          ("b3", (D_E,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)

_erf = np.vectorize(math.erf)


def bf16_round(x: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(p: np.ndarray, round_weights: bool = True):
    out, off = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        v = p[off:off + k]
        if round_weights and name.startswith("W"):
            v = bf16_round(v)
        out[name] = np.asarray(v, np.float64).reshape(shp)
        off += k
    return out


def gelu(z):
    return z * 0.5 * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def standardise(tnf: np.ndarray, mean: np.ndarray, istd: np.ndarray) -> np.ndarray:
    x = np.asarray(tnf, np.float64).copy()
    x[np.isnan(x)] = 0.0
    return (x - mean.astype(np.float64)) * istd.astype(np.float64)


def embed(p, mean, istd, tnf, round_weights=True, round_acts=False):
    """Embeddings; round_acts also rounds the standardised inputs and both
    hidden activations to BF16 (the inference kernel's arithmetic)."""
    w = unpack(p, round_weights)
    act = (lambda a: bf16_round(a.astype(np.float32)).astype(np.float64)) if round_acts else (lambda a: a)
    x = act(standardise(tnf, mean, istd))
    h1 = act(gelu(x @ w["W1"].T + w["b1"]))
    h2 = act(gelu(h1 @ w["W2"].T + w["b2"]))
    return h2 @ w["W3"].T + w["b3"]


def loss_and_grad(p, mean, istd, tnf, pi, pj, y, margin, round_weights=True):
    """Loss and gradient (layout of p) of one batch of pairs."""
    w = unpack(p, round_weights)
    rows = np.empty(2 * len(pi), np.int64)
    rows[0::2], rows[1::2] = pi, pj
    x = standardise(tnf[rows], mean, istd)
    z1 = x @ w["W1"].T + w["b1"]
    a1 = gelu(z1)
    z2 = a1 @ w["W2"].T + w["b2"]
    a2 = gelu(z2)
    e = a2 @ w["W3"].T + w["b3"]
    B = len(pi)
    diff = e[0::2] - e[1::2]
    D = (diff * diff).sum(1)
    y = np.asarray(y, np.float64)
    L = float(np.mean(y * D + (1 - y) * np.maximum(margin - D, 0.0)))
    dD = (y + (1 - y) * np.where(D < margin, -1.0, 0.0)) / B
    de = np.zeros_like(e)
    de[0::2] = dD[:, None] * 2 * diff
    de[1::2] = -dD[:, None] * 2 * diff
    g = {}
    g["W3"] = de.T @ a2
    g["b3"] = de.sum(0)
    dz2 = (de @ w["W3"]) * gelu_grad(z2)
    g["W2"] = dz2.T @ a1
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ w["W2"]) * gelu_grad(z1)
    g["W1"] = dz1.T @ x
    g["b1"] = dz1.sum(0)
    return L, np.concatenate([g[name].ravel() for name, _ in SHAPES])
