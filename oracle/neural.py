"""CPU reference of the neural-field sampler (test infrastructure only).

Multiresolution hash-grid encoding (16 levels x 2 features, 2^19 entries per
level, binary16 parameters) followed by a 32-64-64-1 ReLU MLP without biases
(binary16 weights and activations, binary32 accumulation), trained with MSE.
Forward and backward here are the checker for the HIP kernels'
forward / backward (openvkl_amd/csrc/neural.hip): same hashing, same cast
points (encoding and hidden activations rounded to binary16 before each
matrix product, accumulation in binary32), gradients in binary32.

Hash: ((x * 1) ^ (y * 2654435761) ^ (z * 805459861)) mod 2^19 on uint32 (the
Instant-NGP spatial hash, Mueller et al. 2022); levels whose (res+1)^3 grid fits
in the table are indexed densely instead.
"""
from __future__ import annotations

import numpy as np

L_LEVELS = 16
F_FEAT = 2
LOG2_T = 19
T_SIZE = 1 << LOG2_T
P1, P2 = np.uint32(2654435761), np.uint32(805459861)


def level_resolutions(n_min=16, n_max=2048, levels=L_LEVELS):
    b = np.exp((np.log(n_max) - np.log(n_min)) / (levels - 1))
    return np.array([int(np.floor(n_min * b ** l)) for l in range(levels)], np.uint32)


def corner_index(res, cx, cy, cz):
    """Table index of integer corners (uint32 arrays) at a level of resolution res."""
    side = np.uint64(res) + 1
    if side ** 3 <= T_SIZE:
        return (cx.astype(np.uint64) + side * (cy.astype(np.uint64) + side * cz.astype(np.uint64))).astype(np.uint32)
    h = (cx.astype(np.uint32) * np.uint32(1)) ^ (cy.astype(np.uint32) * P1) ^ (cz.astype(np.uint32) * P2)
    return (h & np.uint32(T_SIZE - 1)).astype(np.uint32)


def encode(x, table, res):
    """x: [S,3] float32 in [0,1]; table: [L, T, 2] float16. Returns enc [S, 32]
    float32 and the per-level corner indices / weights for the backward."""
    S = x.shape[0]
    enc = np.zeros((S, L_LEVELS * F_FEAT), np.float32)
    cache = []
    for l in range(L_LEVELS):
        pos = x.astype(np.float32) * np.float32(res[l])
        c = np.clip(np.floor(pos).astype(np.int64), 0, int(res[l]) - 1)
        w = (pos - c.astype(np.float32)).astype(np.float32)
        c = c.astype(np.uint32)
        idxs, wts = [], []
        f = np.zeros((S, 2), np.float32)
        for k in range(8):
            dx, dy, dz = (k >> 2) & 1, (k >> 1) & 1, k & 1
            wx = w[:, 0] if dx else np.float32(1) - w[:, 0]
            wy = w[:, 1] if dy else np.float32(1) - w[:, 1]
            wz = w[:, 2] if dz else np.float32(1) - w[:, 2]
            wk = (wx * wy * wz).astype(np.float32)
            idx = corner_index(res[l], c[:, 0] + dx, c[:, 1] + dy, c[:, 2] + dz)
            f += wk[:, None] * table[l, idx].astype(np.float32)
            idxs.append(idx)
            wts.append(wk)
        enc[:, 2 * l:2 * l + 2] = f
        cache.append((np.stack(idxs, 1), np.stack(wts, 1)))
    return enc, cache


def _round(a, half):
    """binary16 rounding of activations (the kernels' cast points), or none."""
    return a.astype(np.float16).astype(np.float32) if half else a


def forward(x, table, W1, W2, W3, res, half=True):
    """half=False evaluates every step in binary64 without binary16 rounding
    (the exact function the kernels approximate; used to pin the backward
    against finite differences)."""
    ft = np.float32 if half else np.float64
    enc, cache = encode(x, table, res) if half else encode64(x, table, res)
    X = _round(enc, half)
    h1 = np.maximum(X @ W1.astype(ft), 0).astype(ft)
    H1 = _round(h1, half)
    h2 = np.maximum(H1 @ W2.astype(ft), 0).astype(ft)
    H2 = _round(h2, half)
    y = (H2 @ W3.astype(ft)).reshape(-1)
    return y, dict(X=X, H1=H1, H2=H2, cache=cache)


def encode64(x, table, res):
    """encode() in binary64 (no rounding anywhere)."""
    S = x.shape[0]
    enc = np.zeros((S, L_LEVELS * F_FEAT), np.float64)
    cache = []
    for l in range(L_LEVELS):
        pos = x.astype(np.float64) * float(res[l])
        c = np.clip(np.floor(pos).astype(np.int64), 0, int(res[l]) - 1)
        w = pos - c
        c = c.astype(np.uint32)
        idxs, wts = [], []
        for k in range(8):
            dx, dy, dz = (k >> 2) & 1, (k >> 1) & 1, k & 1
            wk = (w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) * (w[:, 2] if dz else 1 - w[:, 2])
            idx = corner_index(res[l], c[:, 0] + dx, c[:, 1] + dy, c[:, 2] + dz)
            enc[:, 2 * l:2 * l + 2] += wk[:, None] * table[l, idx].astype(np.float64)
            idxs.append(idx)
            wts.append(wk)
        cache.append((np.stack(idxs, 1), np.stack(wts, 1)))
    return enc, cache


def backward(x, target, table, W1, W2, W3, res, half=True):
    """Gradients of sum((y - target)^2) (per-sample dy = 2 (y - t), not divided
    by the batch size), the GPU's unscaled convention. half=False: binary64
    without binary16 rounding."""
    ft = np.float32 if half else np.float64
    y, st = forward(x, table, W1, W2, W3, res, half)
    dy = (ft(2) * (y - target)).astype(ft)
    H1, H2, X = st["H1"], st["H2"], st["X"]
    dW3 = H2.T @ dy
    G2 = (dy[:, None] * W3.astype(ft).reshape(1, -1)) * (H2 > 0)
    G2h = _round(G2, half)
    dW2 = H1.T @ G2h
    G1 = (G2h @ W2.astype(ft).T) * (H1 > 0)
    G1h = _round(G1, half)
    dW1 = X.T @ G1h
    dX = G1h @ W1.astype(ft).T
    dtable = np.zeros(table.shape, ft)
    for l in range(L_LEVELS):
        idx, wts = st["cache"][l]
        for k in range(8):
            np.add.at(dtable[l, :, 0], idx[:, k], wts[:, k] * dX[:, 2 * l])
            np.add.at(dtable[l, :, 1], idx[:, k], wts[:, k] * dX[:, 2 * l + 1])
    return y, dict(W1=dW1, W2=dW2, W3=dW3, table=dtable)
