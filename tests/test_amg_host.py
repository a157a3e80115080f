"""Host parts of the multigrid set-up: tentative prolongator, patterns, permutations.

P_tent is block-diagonal with orthonormal columns and reproduces the near-nullspace exactly
(P_tent B_c = B); the symbolic product pattern equals the structure of the scipy product;
the transpose permutation and position lookup agree with scipy's transpose and indexing.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from fenics_nonlinear_energies_amd import amg


@pytest.mark.parametrize("n,k,agg", [(100, 6, 48), (13, 3, 8), (50, 1, 8), (5, 6, 48), (96, 6, 96)])
def test_tentative_prolongator_reproduces_near_nullspace(n, k, agg):
    B = np.random.default_rng(n + k).standard_normal((n, k))
    Pt, Bc = amg.tentative_prolongator(B, agg)
    assert Pt.has_sorted_indices
    np.testing.assert_allclose(Pt @ Bc, B, atol=1e-13)
    np.testing.assert_allclose((Pt.T @ Pt).toarray(), np.eye(Pt.shape[1]), atol=1e-13)
    # each row only touches its own aggregate's coarse columns
    rows = np.repeat(np.arange(n), np.diff(Pt.indptr))
    cols_per_agg = np.minimum(amg._aggregate_sizes(n, agg), k)
    col_agg = np.repeat(np.arange(cols_per_agg.size), cols_per_agg)
    assert np.array_equal(col_agg[Pt.indices], rows // agg)


def test_pattern_product_and_permutations():
    rng = np.random.default_rng(0)
    A = sp.random(60, 40, density=0.1, random_state=1, format="csr")
    B = sp.random(40, 30, density=0.15, random_state=2, format="csr")
    A.data -= 0.5  # signed values: the numeric product may cancel, the pattern must not
    C = amg._pattern_product(A, B)
    dense = (np.abs(A.toarray()) > 0).astype(float) @ (np.abs(B.toarray()) > 0).astype(float)
    assert np.array_equal(C.toarray() > 0, dense > 0)
    perm = amg._transpose_perm(C)
    Cv = sp.csr_matrix((rng.standard_normal(C.nnz), C.indices, C.indptr), shape=C.shape)
    T = sp.csr_matrix((Cv.data[perm], *_t_pattern(C)), shape=(C.shape[1], C.shape[0]))
    np.testing.assert_array_equal(T.toarray(), Cv.toarray().T)
    r, c = np.nonzero(C.toarray())
    pos = amg._positions(C, r, c)
    np.testing.assert_array_equal(Cv.data[pos], Cv.toarray()[r, c])
    with pytest.raises(ValueError):
        amg._positions(C, np.array([0]), np.array([np.flatnonzero(C.toarray()[0] == 0)[0]]))


def _t_pattern(C):
    T = amg._pattern(C.T.tocsr())
    return T.indices, T.indptr


def test_coarse_inverse_cholesky_and_indefinite_fallback():
    """SPD coarse operator: the exact inverse (Cholesky).  Indefinite: the pseudo-inverse over
    the positive eigenvalues (closed form from numpy's eigendecomposition), symmetric and
    positive semi-definite, which the fallback flag reports."""
    from fenics_nonlinear_energies_amd.amg import coarse_inverse

    rng = np.random.default_rng(0)
    n = 60
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.geomspace(1e-3, 10.0, n)
    A = (q * lam) @ q.T
    A = 0.5 * (A + A.T)
    inv, fb = coarse_inverse(A)
    assert not fb and np.allclose(inv @ A, np.eye(n), atol=1e-8)
    lam2 = lam.copy()
    lam2[:3] = [-2.0, -0.5, 0.0]
    B = (q * lam2) @ q.T
    B = 0.5 * (B + B.T)
    inv2, fb2 = coarse_inverse(B)
    keep = lam2 > 1e-12 * lam2.sum() / n
    expect = (q[:, keep] / lam2[keep]) @ q[:, keep].T
    assert fb2 and np.allclose(inv2, expect, rtol=1e-8, atol=1e-8 * np.abs(expect).max())
    assert np.allclose(inv2, inv2.T, atol=1e-10) and np.linalg.eigvalsh(inv2).min() > -1e-8
