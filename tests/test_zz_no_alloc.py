"""fne_pcg_noalloc (the CG solve in the library's static device scratch) against fne_pcg with a
caller workspace: two calls in a row, the second on a smaller system, each bitwise equal to
fne_pcg on the same input (same kernels, same reduction order), plus the size limit."""

import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from fenics_nonlinear_energies_amd import native

pytestmark = pytest.mark.gpu


def _laplacian_2d(m, seed):
    """SPD 5-point Laplacian plus a seeded positive diagonal shift, n = m*m."""
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(m, m))
    A = sp.kron(sp.eye(m), T) + sp.kron(T, sp.eye(m))
    A = (A + sp.diags(np.random.default_rng(seed).uniform(0.01, 0.1, m * m))).tocsr()
    A.sort_indices()
    return A


def _device_csr(A):
    return (torch.as_tensor(A.indptr.astype(np.int32)).cuda(), torch.as_tensor(A.indices.astype(np.int32)).cuda(),
            torch.as_tensor(A.data.astype(np.float64)).cuda())


def _solve(lib, A, b, inv_diag, noalloc):
    ip, ix, vals = _device_csr(A)
    n = A.shape[0]
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    its, reason, rnorm = C.c_int32(0), C.c_int32(0), C.c_double(0.0)
    s = native.stream_ptr()
    if noalloc:
        rc = lib.fne_pcg_noalloc(n, native.ptr(ip), native.ptr(ix), native.ptr(vals), native.FNE_FP64,
                                 native.ptr(inv_diag), native.ptr(b), native.ptr(x), 1e-10, 1e-50, 1e5, 2000, 8,
                                 C.byref(its), C.byref(reason), C.byref(rnorm), s)
    else:
        nbytes = int(lib.fne_pcg_workspace_bytes(n, A.nnz))
        work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        rc = lib.fne_pcg(n, native.ptr(ip), native.ptr(ix), native.ptr(vals), native.FNE_FP64, native.ptr(inv_diag),
                         native.ptr(b), native.ptr(x), 1e-10, 1e-50, 1e5, 2000, 8, native.ptr(work), nbytes,
                         C.byref(its), C.byref(reason), C.byref(rnorm), s)
    native.check(rc, "fne_pcg_noalloc" if noalloc else "fne_pcg")
    torch.cuda.synchronize()
    return x.cpu().numpy(), its.value, reason.value, rnorm.value


def test_noalloc_cg_matches_workspace_cg(require_gpu):
    lib = native.lib()
    for m, seed in ((300, 1), (97, 2)):  # 90,000 rows, then 9,409 rows in the same scratch
        print(f"fne_pcg_noalloc case m={m} seed={seed}", flush=True)
        A = _laplacian_2d(m, seed)
        b = torch.as_tensor(np.random.default_rng(seed + 10).standard_normal(m * m)).cuda()
        inv_diag = torch.as_tensor(1.0 / A.diagonal()).cuda()
        x0, its0, r0, n0 = _solve(lib, A, b, inv_diag, noalloc=False)
        x1, its1, r1, n1 = _solve(lib, A, b, inv_diag, noalloc=True)
        assert r0 == 2 and its0 > 10
        assert (its1, r1) == (its0, r0) and n1 == n0
        assert np.array_equal(x1, x0)
        res = np.linalg.norm(A @ x1 - b.cpu().numpy()) / np.linalg.norm(b.cpu().numpy())
        assert res < 1e-8


def test_noalloc_cg_rejects_a_system_beyond_its_scratch(require_gpu):
    lib = native.lib()
    n = 600_000  # 5 vectors of 8 bytes per row alone exceed 16 MiB
    assert 40 * n > native.FNE_PCG_NOALLOC_MAX_BYTES
    ip = torch.zeros(n + 1, dtype=torch.int32, device="cuda")  # empty rows: nothing is launched
    v = torch.zeros(1, dtype=torch.float64, device="cuda")
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    its, reason, rnorm = C.c_int32(0), C.c_int32(0), C.c_double(0.0)
    rc = lib.fne_pcg_noalloc(n, native.ptr(ip), native.ptr(ip), native.ptr(v), native.FNE_FP64, None, native.ptr(x),
                             native.ptr(x), 1e-8, 1e-50, 1e5, 10, 4, C.byref(its), C.byref(reason), C.byref(rnorm),
                             native.stream_ptr())
    assert rc == native.FNE_EINVAL
    assert b"static scratch" in lib.fne_last_error()
