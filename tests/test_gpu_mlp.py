"""GPU parity of the warm-start MLP kernels (BF16 MFMA) against the numpy oracle.

The oracle rounds the same product operands to BF16 but accumulates in FP64 and evaluates
tanh in FP64, so an operand can round to the neighbouring BF16 value now and then; the
bound is on the relative Frobenius error: 5e-3 for outputs and gradients.
"""

import numpy as np
import pytest
import torch

from fenics_nonlinear_energies_amd.warmstart import NodeMLP
from oracle import mlp as omlp

pytestmark = pytest.mark.gpu

DIMS = [10, 128, 128, 128, 128, 3]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _net_and_data(n=3000, seed=0):
    net = NodeMLP(DIMS, seed=seed)
    rng = np.random.default_rng(seed + 1)
    x = rng.standard_normal((n, DIMS[0])).astype(np.float32)
    ws = [w.cpu().numpy().astype(np.float64) for w in net.weights_bf16_values()]
    for b in net.b32:
        b.copy_(torch.as_tensor(rng.standard_normal(b.numel()).astype(np.float32) * 0.1))
    bs = [b.cpu().numpy().astype(np.float64) for b in net.b32]
    return net, x, ws, bs


def test_forward_matches_oracle(require_gpu):
    net, x, ws, bs = _net_and_data()
    y = net.forward(torch.as_tensor(x)).cpu().numpy()
    yo, _ = omlp.forward(ws, bs, x)
    assert _rel(y, yo) < 5e-3
    # bf16 weights are the rounded FP32 master weights
    w32 = net.w32[1].cpu().numpy()
    assert np.array_equal(net.weights_bf16_values()[1].cpu().numpy(), omlp.bf16_round(w32).astype(np.float32))


def test_backward_matches_oracle_and_is_deterministic(require_gpu):
    net, x, ws, bs = _net_and_data(n=2500, seed=3)
    dy = np.random.default_rng(7).standard_normal((x.shape[0], DIMS[-1])).astype(np.float32)
    net.forward(torch.as_tensor(x), keep=True)
    dw, db = net.backward(torch.as_tensor(dy))
    dw2, db2 = net.backward(torch.as_tensor(dy))
    assert all(torch.equal(a, b) for a, b in zip(dw, dw2)) and all(torch.equal(a, b) for a, b in zip(db, db2))
    _, acts = omlp.forward(ws, bs, x)
    dwo, dbo = omlp.backward(ws, bs, x, acts, dy)
    for l in range(len(ws)):
        assert _rel(dw[l].cpu().numpy(), dwo[l]) < 5e-3, l
        assert _rel(db[l].cpu().numpy(), dbo[l]) < 5e-3, l


def test_training_reduces_loss(require_gpu):
    rng = np.random.default_rng(5)
    n = 20000
    x = rng.uniform(-1, 1, (n, DIMS[0])).astype(np.float32)
    A = rng.standard_normal((DIMS[0], DIMS[-1])).astype(np.float32) * 0.5
    t = np.sin(x @ A).astype(np.float32)
    net = NodeMLP(DIMS, seed=11)
    losses = net.train_mse(torch.as_tensor(x), torch.as_tensor(t), epochs=40, batch=4096, lr=2e-3)
    assert losses[-1] < 0.05 * losses[0], losses[::8]
