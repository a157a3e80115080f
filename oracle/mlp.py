"""Oracle of the per-node warm-start MLP (numpy) -- TEST INFRASTRUCTURE ONLY.

Restates the device network exactly as specified, not the reference (the reference has no
learned warm start; it restarts each load / continuation step from the previous state,
core/petsc/load_step_driver.py:58-110): y = W_L(...tanh(W_1 x + b_1)...) + b_L.  Every
matrix product takes BF16-rounded operands (round to nearest even) and accumulates in FP64;
activations, tanh and its derivative stay unrounded, as on the device (BF16 operands, FP32
accumulation).  ``backward`` returns the gradients of sum_i dy_i . y_i.
"""

from __future__ import annotations

import numpy as np


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round FP32 values to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    out = (u.astype(np.uint32) << 16).view(np.float32)
    nan = np.isnan(f)
    if np.any(nan):
        out = np.where(nan, np.float32(np.nan), out)
    return out.astype(np.float64)


def forward(ws, bs, x, rnd=bf16_round):
    """ws[l]: (out, in) float weights (already BF16 values), bs[l]: (out,).  Returns (y, acts).

    ``rnd`` rounds the product operands (identity for the exact-arithmetic network)."""
    h = np.asarray(x, dtype=np.float64)
    acts = []
    for l, (W, b) in enumerate(zip(ws, bs)):
        z = rnd(h) @ np.asarray(W, np.float64).T + np.asarray(b, np.float64)
        if l < len(ws) - 1:
            h = np.tanh(z)
            acts.append(h)
        else:
            h = z
    return h, acts


def backward(ws, bs, x, acts, dy, rnd=bf16_round):
    dz = np.asarray(dy, dtype=np.float64)
    dws = [None] * len(ws)
    dbs = [None] * len(ws)
    for l in range(len(ws) - 1, -1, -1):
        hin = np.asarray(x, np.float64) if l == 0 else acts[l - 1]
        dws[l] = rnd(dz).T @ rnd(hin)
        dbs[l] = dz.sum(axis=0)
        if l == 0:
            break
        dz = (rnd(dz) @ np.asarray(ws[l], np.float64)) * (1.0 - hin * hin)
    return dws, dbs
