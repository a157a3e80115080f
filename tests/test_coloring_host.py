"""Distance-2 colouring of the SFD Hessian mode (fne_color_distance2, host C++).

Pinned by the property the SFD extraction needs -- within one colour no two columns meet in a
row, so one Hessian-vector product per colour recovers each entry -- on random symmetric
patterns and on the P4 slope's free-DOF pattern, and by equality with a direct numpy
restatement of the same greedy rule (rows in order, smallest colour unused in the distance-2
neighbourhood).  Argument errors return FNE_EINVAL.
"""

import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

from fenics_nonlinear_energies_amd import native

_i32 = C.POINTER(C.c_int32)


def _color(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    ip = np.ascontiguousarray(A.indptr, dtype=np.int32)
    ix = np.ascontiguousarray(A.indices, dtype=np.int32)
    col = np.empty(A.shape[0], dtype=np.int32)
    k = native.lib().fne_color_distance2(A.shape[0], ip.ctypes.data_as(_i32), ix.ctypes.data_as(_i32),
                                         col.ctypes.data_as(_i32))
    return int(k), col


def _greedy_numpy(A):
    A = sp.csr_matrix(A)
    n = A.shape[0]
    col = -np.ones(n, dtype=np.int64)
    for v in range(n):
        nb = A.indices[A.indptr[v]:A.indptr[v + 1]]
        d2 = np.unique(np.concatenate([A.indices[A.indptr[u]:A.indptr[u + 1]] for u in nb] + [nb]))
        used = set(col[d2][col[d2] >= 0].tolist())
        c = 0
        while c in used:
            c += 1
        col[v] = c
    return col


def _check_orthogonal(A, col):
    A = sp.csr_matrix(A)
    for i in range(A.shape[0]):
        c = col[A.indices[A.indptr[i]:A.indptr[i + 1]]]
        assert np.unique(c).size == c.size, f"row {i} has two columns of one colour"


@pytest.mark.parametrize("n,density,seed", [(60, 0.05, 0), (200, 0.02, 1), (1, 1.0, 2)])
def test_random_symmetric_patterns(n, density, seed):
    R = sp.random(n, n, density=density, random_state=seed, format="csr")
    A = (R + R.T + sp.identity(n)).tocsr()
    k, col = _color(A)
    assert k == col.max() + 1
    _check_orthogonal(A, col)
    np.testing.assert_array_equal(col, _greedy_numpy(A))


def test_slope_p4_free_dof_pattern():
    from fenics_nonlinear_energies_amd.problems import slope_stability as S

    p, adj, _ = S.build_slope_problem(4.0, degree=4)
    A = sp.csr_matrix(adj)
    A = ((A + A.T) != 0).astype(float) + sp.identity(A.shape[0])
    k, col = _color(A)
    _check_orthogonal(A, col)
    # every colour class is needed at least as many times as the largest row has entries
    assert k >= int(np.diff(sp.csr_matrix(A).indptr).max())
    np.testing.assert_array_equal(col, _greedy_numpy(A))


def test_argument_errors():
    lib = native.lib()
    ip = np.array([0, 1, 2], dtype=np.int32)
    bad = np.array([0, 5], dtype=np.int32)  # column 5 of a 2-row pattern
    col = np.empty(2, dtype=np.int32)
    assert lib.fne_color_distance2(2, ip.ctypes.data_as(_i32), bad.ctypes.data_as(_i32),
                                   col.ctypes.data_as(_i32)) == native.FNE_EINVAL
    assert lib.fne_color_distance2(-1, None, None, None) == native.FNE_EINVAL
    assert lib.fne_color_distance2(0, None, None, None) == 0
