"""CPU restatement of the learned delta policy's network (test infrastructure
only: loaded by tests, never by the product).

The network the ADDS scheduler evaluates (32 -> 64 ReLU -> 4) has no
counterpart in the reference; it replaces the fixed per-phase heuristics of
JAS/hip_adds_sssp.cpp:683-727 and ONE/sssp_hip.cpp:564-571.  This module
restates its forward pass and the gradient of the training loss in float64
with the weights rounded to bf16 exactly as the device does (round to nearest
even), so the HIP kernels can be checked against it.  The restatement itself
is pinned by central finite differences and by a bf16 rounding table in
tests/test_mlp_oracle.py.
"""
from __future__ import annotations

import numpy as np


def bf16_round(a) -> np.ndarray:
    """fp32 -> bf16 -> fp32, round to nearest, ties to even (finite inputs)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


N_OUT = 4
N_PARAMS = 64 * 32 + 64 + N_OUT * 64 + N_OUT  # 2372


def unpack(flat):
    flat = np.asarray(flat, np.float64)
    w1 = flat[:2048].reshape(64, 32)
    b1 = flat[2048:2112]
    w2 = flat[2112:2112 + 64 * N_OUT].reshape(N_OUT, 64)
    b2 = flat[2112 + 64 * N_OUT:N_PARAMS]
    return w1, b1, w2, b2


def forward(flat, mean, inv_std, x, round_weights=True):
    """Outputs (n x 4) and hidden activations (n x 64)."""
    w1, b1, w2, b2 = unpack(flat)
    if round_weights:
        w1 = bf16_round(w1).astype(np.float64)
        w2 = bf16_round(w2).astype(np.float64)
    xn = (np.asarray(x, np.float64) - np.asarray(mean, np.float64)) * np.asarray(inv_std, np.float64)
    h = np.maximum(xn @ w1.T + b1, 0.0)
    return h @ w2.T + b2, h, xn


def loss(flat, mean, inv_std, x, y, w=None, round_weights=True) -> float:
    o, _, _ = forward(flat, mean, inv_std, x, round_weights)
    e = o - np.asarray(y, np.float64)
    sw = np.ones(len(e)) if w is None else np.asarray(w, np.float64)
    return float((sw * (e * e).sum(axis=1)).sum() / len(e))


def loss_grad(flat, mean, inv_std, x, y, w=None, round_weights=True):
    """(loss, gradient w.r.t. the 2372 trained parameters).  With rounded
    weights the gradient is taken through the rounding unchanged
    (straight-through), as the device does."""
    w1, b1, w2, b2 = unpack(flat)
    w2r = bf16_round(w2).astype(np.float64) if round_weights else w2
    o, h, xn = forward(flat, mean, inv_std, x, round_weights)
    n = len(o)
    sw = np.ones(n) if w is None else np.asarray(w, np.float64)
    e = o - np.asarray(y, np.float64)
    d = 2.0 * sw[:, None] * e / n            # dL/do
    gw2 = d.T @ h
    gb2 = d.sum(axis=0)
    dh = (d @ w2r) * (h > 0)
    gw1 = dh.T @ xn
    gb1 = dh.sum(axis=0)
    L = float((sw * (e * e).sum(axis=1)).sum() / n)
    return L, np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2])
