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


def test_coarse_inverse_negative_trace_keeps_a_positive_cutoff():
    """An indefinite coarse operator with a negative trace: the cutoff stays positive, so no
    negative eigenvalue is inverted and the result is positive semi-definite."""
    from fenics_nonlinear_energies_amd.amg import coarse_inverse

    rng = np.random.default_rng(1)
    n = 40
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.concatenate((-np.geomspace(1.0, 50.0, n - 5), np.geomspace(1e-2, 1.0, 5)))
    A = (q * lam) @ q.T
    A = 0.5 * (A + A.T)
    assert np.trace(A) < 0
    inv, fb = coarse_inverse(A)
    keep = lam > 1e-12 * np.abs(lam).max()
    expect = (q[:, keep] / lam[keep]) @ q[:, keep].T
    assert fb and np.allclose(inv, expect, rtol=1e-8, atol=1e-8 * np.abs(expect).max())
    assert np.linalg.eigvalsh(inv).min() > -1e-8


def test_lanczos_estimate_is_finite_with_a_negative_diagonal():
    """rho(D^-1 A) by Lanczos on |D|^-1/2 A |D|^-1/2: finite when an indefinite operator has a
    negative diagonal entry, and equal to the exact spectral radius on a small SPD operator."""
    import torch

    class _HostOp:
        def __init__(self, M):
            self.M = M

        def mult_tensor(self, x, y):
            y.copy_(torch.as_tensor(self.M @ x.numpy()))

    rng = np.random.default_rng(2)
    n = 12  # Lanczos with 12 steps on 12 unknowns is exact
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    S = (q * np.linspace(0.5, 3.0, n)) @ q.T
    pc = amg.SmoothedAggregationPC()
    d = np.diag(S).copy()
    lam = pc._lanczos_lambda_max(0, _HostOp(S), torch.as_tensor(1.0 / d), torch.as_tensor(rng.standard_normal(n)))
    exact = np.max(np.abs(np.linalg.eigvals(np.diag(1.0 / d) @ S)))
    assert abs(lam - exact) <= 1e-8 * exact
    T = S.copy()
    T[3, 3] = -1.0
    lam2 = pc._lanczos_lambda_max(0, _HostOp(T), torch.as_tensor(1.0 / np.diag(T)),
                                  torch.as_tensor(rng.standard_normal(n)))
    assert np.isfinite(lam2) and lam2 > 0


def test_fixed_prolongator_levels_symbolic_phase():
    """A fixed transfer heads the hierarchy unchanged (pattern and values), aggregation goes on
    below it, and the injection rows pick each coarse DOF's coinciding fine DOF."""
    import torch

    from fenics_nonlinear_energies_amd import pmg
    from fenics_nonlinear_energies_amd.problems import slope_stability as S

    vn, vt = S.slope_macro_mesh(2.0, **S.REFERENCE_GEOMETRY)
    n2, e2 = S.lagrange_mesh(vn, vt, 2)
    P21 = pmg.scalar_transfer(e2, 2, vt, 1, len(n2), len(vn))
    inj = amg._injection_rows(P21)
    np.testing.assert_allclose(n2[inj], vn, atol=1e-12)
    # a scalar Laplacian-like pattern on the P2 nodes (elements' node cliques)
    rows = np.repeat(e2, e2.shape[1], axis=1).ravel()
    cols = np.tile(e2, (1, e2.shape[1])).ravel()
    A = sp.csr_matrix((np.ones(rows.size), (rows, cols)), shape=(len(n2), len(n2)))
    A.sum_duplicates()
    A.sort_indices()

    class _Pat:
        n, nnz = A.shape[0], A.nnz
        indptr, indices = torch.as_tensor(A.indptr.astype(np.int32)), torch.as_tensor(A.indices.astype(np.int32))

    pc = amg.SmoothedAggregationPC(fixed_prolongators=[P21], max_coarse=20)
    pc._symbolic(_Pat, "cpu", 1)
    L0 = pc.levels[0]
    assert L0["fixed"] and L0["n"] == len(n2) and L0["nc"] == len(vn)
    np.testing.assert_array_equal(L0["P64"].indices.numpy(), P21.indices)
    np.testing.assert_array_equal(L0["Pfix"].numpy(), P21.data)
    assert len(pc.levels) >= 2 and not pc.levels[1]["fixed"] and pc.levels[1]["n"] == len(vn)


@pytest.mark.parametrize("steps", [1, 2, 3, 4])
def test_chebyshev_coefficients_give_the_chebyshev_error_polynomial(steps):
    """The smoother's recurrence x_{k+1} = x_k + c_k (x_k - x_{k-1}) + w_k D^-1 (r - A x_k)
    (x_0 = 0) with the multigrid's coefficients on a diagonal operator with D = I: the error of
    each eigencomponent lambda in [a, b] = [0.1, 1.1] rho shrinks by
    T_m((theta - lambda) / delta) / T_m(theta / delta), the Chebyshev closed form."""
    pc = amg.SmoothedAggregationPC(sweeps=steps, smoother="chebyshev")
    rho = 2.5
    lam = np.linspace(0.1 * rho, 1.1 * rho, 11)
    w = pc._cheb_weights(rho)
    c = pc._cheb_coefs()
    r = np.ones_like(lam)  # A x = r, exact x = r / lam
    x_prev, x = np.zeros_like(lam), np.zeros_like(lam)
    for k in range(steps):
        y = x + c[k] * (x - x_prev) + w[k] * (r - lam * x)
        x_prev, x = x, y
    err = (r / lam - x) / (r / lam)
    theta, delta = 0.6 * rho, 0.5 * rho
    def T(m, t):  # Chebyshev polynomial of the first kind, any real argument
        if abs(t) <= 1.0:
            return np.cos(m * np.arccos(t))
        return np.sign(t) ** m * np.cosh(m * np.arccosh(abs(t)))

    expect = np.array([T(steps, (theta - l) / delta) for l in lam]) / T(steps, theta / delta)
    np.testing.assert_allclose(err, expect, atol=1e-12)
    with pytest.raises(ValueError):
        amg.SmoothedAggregationPC(smoother="sor")


@pytest.mark.parametrize("case", ["spd", "indefinite", "negative_trace"])
def test_coarse_inverse_device_form_equals_the_host_form(case):
    """The tensor form of the coarse inverse (``coarse_factor="device"``: Cholesky and its
    inverse, or the positive-part pseudo-inverse) equals the LAPACK form on the same operator and
    takes the fallback on the same operators; run here on CPU tensors, on the GPU in set-ups."""
    import torch

    from fenics_nonlinear_energies_amd.amg import coarse_inverse, coarse_inverse_device

    rng = np.random.default_rng(2)
    n = 50
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.geomspace(1e-3, 10.0, n)
    if case == "indefinite":
        lam[:3] = [-2.0, -0.5, 0.0]
    elif case == "negative_trace":
        lam = np.concatenate((-np.geomspace(1.0, 50.0, n - 5), np.geomspace(1e-2, 1.0, 5)))
    A = (q * lam) @ q.T
    A = 0.5 * (A + A.T)
    ref, fb_ref = coarse_inverse(A)
    got, fb = coarse_inverse_device(torch.as_tensor(A))
    assert fb == fb_ref == (case != "spd")
    assert np.allclose(got.numpy(), ref, rtol=1e-8, atol=1e-9 * np.abs(ref).max())
