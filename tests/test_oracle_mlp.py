"""Pins of the warm-start MLP oracle: BF16 rounding and the backward pass.

* ``bf16_round`` against hand-made bit patterns (ties to even, exact values kept);
* ``backward`` (exact arithmetic) against central finite differences of ``forward``.
"""

import numpy as np

from oracle import mlp


def test_bf16_round_bit_patterns():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # BF16 spacing at 1
    vals = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9, -2.5, 3.0e38, 0.0], np.float32)
    r = mlp.bf16_round(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie -> even (mantissa 0)
    assert r[2] == 1.0 + 2 * ulp  # tie -> even (mantissa 2)
    assert r[3] == 1.0
    assert r[4] == -2.5 and r[6] == 0.0
    assert np.isfinite(r[5])
    x = np.random.default_rng(0).standard_normal(1000).astype(np.float32)
    assert np.all(np.abs(mlp.bf16_round(x) - x) <= np.abs(x) * 2.0 ** -8 * (1 + 1e-7))
    assert one == 1.0


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    dims = [5, 7, 6, 3]
    ws = [rng.standard_normal((o, i)) * 0.5 for i, o in zip(dims[:-1], dims[1:])]
    bs = [rng.standard_normal(o) * 0.1 for o in dims[1:]]
    x = rng.standard_normal((9, dims[0]))
    dy = rng.standard_normal((9, dims[-1]))
    ident = lambda a: np.asarray(a, np.float64)
    y, acts = mlp.forward(ws, bs, x, rnd=ident)
    dws, dbs = mlp.backward(ws, bs, x, acts, dy, rnd=ident)
    f = lambda W, B: float(np.sum(dy * mlp.forward(W, B, x, rnd=ident)[0]))
    h = 1e-6
    for l in range(len(ws)):
        for _ in range(5):
            i, j = rng.integers(ws[l].shape[0]), rng.integers(ws[l].shape[1])
            wp = [w.copy() for w in ws]
            wm = [w.copy() for w in ws]
            wp[l][i, j] += h
            wm[l][i, j] -= h
            assert abs((f(wp, bs) - f(wm, bs)) / (2 * h) - dws[l][i, j]) < 1e-6
        k = rng.integers(bs[l].size)
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][k] += h
        bm[l][k] -= h
        assert abs((f(ws, bp) - f(ws, bm)) / (2 * h) - dbs[l][k]) < 1e-6
