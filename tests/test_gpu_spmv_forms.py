"""Both forms of the CSR row kernels -- lanes per row and CSR-stream -- against scipy.

The CSR-stream form streams a workgroup's rows through LDS in chunks of 2048 entries; the
matrices cover 7-entry rows (one chunk per workgroup), 47-entry rows (several chunks), ragged
rows with empty ones and rows longer than a chunk, row counts that leave the last workgroup
partly empty, and unaligned arrays (where the library falls back to lanes per row).  Results are
compared with scipy's product of the same (FP32-rounded) values, and repeated calls must agree
bit for bit.
"""

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu


def _matrix(kind, rng):
    if kind == "p1":
        n, rl = 5003, np.full(5003, 7)
    elif kind == "vector":
        n, rl = 3001, rng.integers(40, 55, 3001)
    elif kind == "ragged":
        n = 2500
        rl = rng.integers(0, 120, n)
        rl[rng.choice(n, 40, replace=False)] = 0
        rl[[3, 700, 701, 2499]] = [2500, 3100, 2047, 2049]  # longer than one chunk
    else:  # wide rectangular operator, few columns
        n, rl = 1777, rng.integers(1, 9, 1777)
    n_cols = 97 if kind == "rect" else n
    rows = []
    for i, k in enumerate(rl):
        k = min(int(k), n_cols)
        rows.append(np.sort(rng.choice(n_cols, k, replace=False)))
    indptr = np.concatenate([[0], np.cumsum([r.size for r in rows])]).astype(np.int32)
    indices = np.concatenate(rows).astype(np.int32) if indptr[-1] else np.zeros(0, np.int32)
    data = rng.standard_normal(indices.size)
    return sp.csr_matrix((data, indices, indptr), shape=(n, n_cols))


@pytest.mark.parametrize("kind", ["p1", "vector", "ragged", "rect"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_spmv_forms_match_scipy(require_gpu, kind, prec):
    import torch

    from fenics_nonlinear_energies_amd import amg, native
    from fenics_nonlinear_energies_amd.linalg import SPMV_FORM_STREAM, lanes_per_row

    rng = np.random.default_rng({"p1": 1, "vector": 2, "ragged": 3, "rect": 4}[kind])
    A = _matrix(kind, rng)
    if prec == "fp32":
        A.data = A.data.astype(np.float32).astype(np.float64)
    dev = torch.device("cuda")
    M = amg._DevMat(A.indptr, A.indices, A.shape[1], dev, prec)
    M.data.copy_(torch.as_tensor(A.data))
    x = rng.standard_normal(A.shape[1])
    ref = A @ x
    xt = torch.as_tensor(x).to(dev)
    avg = A.nnz / A.shape[0]
    codes = {"rows": M.prec | (lanes_per_row(avg) << 8)}
    for lanes in (1, 2, 4, 16, 64):
        codes[f"stream{lanes}"] = M.prec | (lanes << 8) | SPMV_FORM_STREAM
    for name, code in codes.items():
        M.code = code
        y1 = torch.full((A.shape[0],), np.nan, dtype=torch.float64, device=dev)
        y2 = torch.empty_like(y1)
        M.mult(xt, y1)
        M.mult(xt, y2)
        print(kind, prec, name, flush=True)
        np.testing.assert_allclose(y1.cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        assert torch.equal(y1, y2)
    # unaligned arrays: the stream form falls back to lanes per row
    ind = torch.zeros(A.nnz + 1, dtype=torch.int32, device=dev)
    ind[1:] = M.indices
    val = torch.zeros(A.nnz + 1, dtype=M.data.dtype, device=dev)
    val[1:] = M.data
    y = torch.empty(A.shape[0], dtype=torch.float64, device=dev)
    native.check(native.lib().fne_spmv(A.shape[0], native.ptr(M.indptr), native.ptr(ind[1:]), native.ptr(val[1:]),
                                       codes["stream4"], native.ptr(xt), native.ptr(y), native.stream_ptr()),
                 "fne_spmv")
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_vcycle_kernels_stream_equals_rows_form(require_gpu, prec):
    """fne_vcycle_op and fne_vcycle_cheb give the same values in both forms (ragged rows)."""
    import torch

    from fenics_nonlinear_energies_amd import amg
    from fenics_nonlinear_energies_amd.linalg import lanes_per_row

    rng = np.random.default_rng(5)
    A = _matrix("ragged", rng)
    if prec == "fp32":
        A.data = A.data.astype(np.float32).astype(np.float64)
    dev = torch.device("cuda")
    M = amg._DevMat(A.indptr, A.indices, A.shape[1], dev, prec)
    M.data.copy_(torch.as_tensor(A.data))
    n = A.shape[0]
    x, xp, r, w = (torch.as_tensor(rng.standard_normal(n)).to(dev) for _ in range(4))
    stream_code = M.code
    rows_code = M.prec | (lanes_per_row(A.nnz / n) << 8)
    out = {}
    for name, code in (("stream", stream_code), ("rows", rows_code)):
        M.code = code
        ys = []
        for op in (amg._SWEEP_FIRST, amg._SWEEP, amg._RESIDUAL):
            y = torch.empty(n, dtype=torch.float64, device=dev)
            M.op(op, None if op == amg._SWEEP_FIRST else x, r, w, None, y)
            ys.append(y)
        yy = x.clone()
        M.op(amg._ADD_PRODUCT, r, None, None, yy, yy)
        ys.append(yy)
        y = torch.empty(n, dtype=torch.float64, device=dev)
        M.cheb(x, xp, r, w, 0.4, y)
        ys.append(y)
        out[name] = ys
    Ax = A @ x.cpu().numpy()
    rn, wn, xn, xpn = r.cpu().numpy(), w.cpu().numpy(), x.cpu().numpy(), xp.cpu().numpy()
    x1 = wn * rn
    refs = [x1 + wn * (rn - A @ x1), xn + wn * (rn - Ax), rn - Ax, xn + A @ rn, xn + 0.4 * (xn - xpn) + wn * (rn - Ax)]
    for a, b, ref in zip(out["stream"], out["rows"], refs):
        scale = np.abs(ref).max()
        np.testing.assert_allclose(a.cpu().numpy(), ref, rtol=1e-11, atol=1e-12 * scale)
        np.testing.assert_allclose(b.cpu().numpy(), ref, rtol=1e-11, atol=1e-12 * scale)
