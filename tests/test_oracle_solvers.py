"""Pins of the oracle's Krylov solvers and line search by what the mathematics fixes (CPU).

* cg: the solution of an SPD system; one iteration for a diagonal matrix under Jacobi; at
  most k iterations for k distinct eigenvalues; the first iterate is the steepest-descent
  step; a sign change of p.Ap is reported as an indefinite matrix.
* gmres: the solution of a nonsymmetric system, also with restarts; at most k iterations for
  a diagonalisable matrix with k distinct eigenvalues; the returned residual meets rtol.
* stcg: with no radius (and no preconditioner) it is cg; it stops on the unpreconditioned
  residual; a radius below the first step gives the scaled first
  direction; the iterates grow in norm while the quadratic model falls (Steihaug 1983,
  Thm 2.1), so a radius between two iterates gives a boundary point better than every
  interior one; negative curvature ends on the boundary along p.
* golden_section: the minimiser of a parabola within tol, with the evaluation count the
  bracket's 1/gamma contraction fixes.
* the trust-region model solvers of the Newton driver (minimizers.solve_trust_1d/2d):
  closed forms.
"""

import math

import numpy as np
import pytest
import scipy.sparse as sp

from fenics_nonlinear_energies_amd.minimizers import solve_trust_1d, solve_trust_2d
from oracle import solvers


def _spd(n, cond, seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.geomspace(1.0, cond, n)
    return sp.csr_matrix((q * lam) @ q.T), rng.standard_normal(n)


def _q(A, b, s):
    return float(-b @ s + 0.5 * s @ (A @ s))


def test_cg_solves_spd_system():
    A, b = _spd(30, 100.0, 0)
    x, its, reason = solvers.cg(A, b, rtol=1e-14)
    assert reason == 2 and its <= 60
    xs = np.linalg.solve(A.toarray(), b)
    assert np.abs(x - xs).max() <= 1e-10 * np.abs(xs).max()


def test_cg_jacobi_on_diagonal_matrix_is_one_step():
    d = np.linspace(1.0, 50.0, 17)
    b = np.random.default_rng(1).standard_normal(17)
    x, its, reason = solvers.cg(sp.diags(d).tocsr(), b, rtol=1e-12)
    assert its == 1 and reason == 2
    assert np.allclose(x, b / d, rtol=1e-14, atol=0)


def test_cg_k_distinct_eigenvalues_k_iterations():
    rng = np.random.default_rng(2)
    q, _ = np.linalg.qr(rng.standard_normal((20, 20)))
    lam = np.repeat([1.0, 4.0, 9.0], [7, 7, 6])
    A = sp.csr_matrix((q * lam) @ q.T)
    b = rng.standard_normal(20)
    x, its, reason = solvers.cg(A, b, rtol=1e-10, inv_diag=np.ones(20))
    assert reason == 2 and its <= 3
    assert np.allclose(A @ x, b, atol=1e-8)


def test_cg_first_iterate_is_steepest_descent_step():
    A, b = _spd(12, 10.0, 3)
    x, its, reason = solvers.cg(A, b, rtol=1e-14, max_it=1, inv_diag=np.ones(12))
    assert its == 1 and reason == -3
    assert np.allclose(x, (b @ b) / (b @ (A @ b)) * b, rtol=1e-13, atol=0)


def test_cg_reports_indefinite_matrix():
    A = sp.diags([1.0, -1.0]).tocsr()
    _, its, reason = solvers.cg(A, np.ones(2), inv_diag=np.ones(2))
    assert reason == -10 and its == 0


def _nonsym(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n)) / math.sqrt(n) + 3.0 * np.eye(n)
    return sp.csr_matrix(M), rng.standard_normal(n)


@pytest.mark.parametrize("restart", [30, 5])
def test_gmres_solves_nonsymmetric_system(restart):
    A, b = _nonsym(25, 4)
    x, its, reason = solvers.gmres(A, b, rtol=1e-12, restart=restart)
    assert reason == 2
    if restart >= 25:
        assert its <= 25
    xs = np.linalg.solve(A.toarray(), b)
    assert np.abs(x - xs).max() <= 1e-9 * np.abs(xs).max()
    inv = solvers.jacobi_inverse(A)
    assert np.linalg.norm(inv * (b - A @ x)) <= 1e-12 * np.linalg.norm(inv * b) * (1 + 1e-6)


def test_gmres_k_distinct_eigenvalues_k_iterations():
    rng = np.random.default_rng(5)
    V = rng.standard_normal((18, 18)) + 4.0 * np.eye(18)
    lam = np.repeat([1.0, 2.0, 5.0, 7.0], [5, 5, 4, 4])
    A = sp.csr_matrix(V @ np.diag(lam) @ np.linalg.inv(V))
    b = rng.standard_normal(18)
    x, its, reason = solvers.gmres(A, b, rtol=1e-9, inv_diag=np.ones(18))
    assert reason == 2 and its <= 4
    assert np.linalg.norm(b - A @ x) <= 1e-8 * np.linalg.norm(b)


def test_stcg_without_radius_is_cg():
    # without a preconditioner both test the same residual norm
    A, b = _spd(40, 1e3, 6)
    one = np.ones(40)
    for rtol in (1e-3, 1e-10):
        xc, itc, rc = solvers.cg(A, b, rtol=rtol, inv_diag=one)
        xs, its, rs = solvers.stcg(A, b, radius=math.inf, rtol=rtol, inv_diag=one)
        assert (its, rs) == (itc, rc)
        assert np.abs(xs - xc).max() <= 1e-12 * np.abs(xc).max()


def test_stcg_stops_on_the_unpreconditioned_residual():
    # Jacobi with a badly scaled diagonal separates ||r|| from ||D^-1 r||
    A, b = _spd(40, 1e2, 9)
    s = np.geomspace(1.0, 1e3, 40)
    A = sp.csr_matrix(sp.diags(s) @ A @ sp.diags(s))
    x, its, reason = solvers.stcg(A, b, rtol=1e-6)
    assert reason == 2 and its >= 2
    assert np.linalg.norm(b - A @ x) <= 1e-6 * np.linalg.norm(b) * (1 + 1e-8)
    x1, _, _ = solvers.stcg(A, b, rtol=1e-6, max_it=its - 1)
    assert np.linalg.norm(b - A @ x1) > 1e-6 * np.linalg.norm(b)


def test_stcg_small_radius_is_scaled_first_direction():
    A, b = _spd(20, 50.0, 7)
    inv = solvers.jacobi_inverse(A)
    z0 = inv * b
    first = (b @ z0) / (z0 @ (A @ z0)) * z0
    delta = 1e-3 * np.linalg.norm(first)
    s, its, reason = solvers.stcg(A, b, radius=delta, rtol=1e-12)
    assert its == 1 and reason == 6
    assert np.allclose(s, delta * z0 / np.linalg.norm(z0), rtol=1e-13, atol=0)


def test_stcg_path_grows_in_norm_and_decreases_the_model():
    # the theorem holds in the norm the preconditioner induces: Euclidean without one
    # (well conditioned, so that rounding stays far below the increments)
    A, b = _spd(40, 30.0, 8)
    one = np.ones(40)
    _, n_its, _ = solvers.stcg(A, b, rtol=1e-6, inv_diag=one)
    path = [solvers.stcg(A, b, rtol=1e-6, max_it=k, inv_diag=one)[0] for k in range(1, n_its + 1)]
    norms = [np.linalg.norm(s) for s in path]
    qs = [_q(A, b, s) for s in path]
    assert n_its >= 6
    assert all(n1 > n0 for n0, n1 in zip(norms, norms[1:]))
    assert all(q1 < q0 for q0, q1 in zip(qs, qs[1:]))
    # a radius between the 3rd and 4th iterates: boundary point, better than every interior one
    delta = 0.5 * (norms[2] + norms[3])
    s, its, reason = solvers.stcg(A, b, radius=delta, rtol=1e-6, inv_diag=one)
    assert reason == 6 and its == 4
    assert abs(np.linalg.norm(s) - delta) <= 1e-12 * delta
    assert _q(A, b, s) < qs[2]


def test_stcg_negative_curvature_ends_on_the_boundary():
    A = sp.diags([2.0, -1.0, 3.0]).tocsr()
    b = np.array([0.1, 1.0, 0.1])
    s, its, reason = solvers.stcg(A, b, radius=0.7, inv_diag=np.ones(3))
    assert reason == 5 and its == 1
    assert np.allclose(s, 0.7 * b / np.linalg.norm(b), rtol=1e-14, atol=0)
    s_inf, _, reason_inf = solvers.stcg(A, b, inv_diag=np.ones(3))
    assert reason_inf == 5 and not s_inf.any()


def test_golden_section_parabola_and_evaluation_count():
    gam = 0.5 + math.sqrt(5.0) / 2.0
    for a, b, tol, xm in ((-0.5, 2.0, 1e-3, 0.3), (-0.5, 2.0, 1e-6, 1.7), (0.0, 1.0, 1e-2, 0.05)):
        seen = []

        def f(t):
            seen.append(t)
            return (t - xm) ** 2

        best, n = solvers.golden_section(f, a, b, tol)
        assert abs(best - xm) <= tol
        m = math.ceil(math.log((b - a) / tol) / math.log(gam))
        assert n == len(seen) and n in (4 + m, 5 + m)


def test_trust_1d_closed_form():
    assert solve_trust_1d(1.0, 2.0, 10.0) == -0.5
    assert solve_trust_1d(1.0, 2.0, 0.2) == -0.2
    assert solve_trust_1d(-1.0, -3.0, 0.4) == 0.4  # negative curvature: to the boundary downhill
    assert solve_trust_1d(1.0, 0.0, 0.4) == -0.4
    assert solve_trust_1d(0.0, -1.0, 0.4) == 0.0
    assert solve_trust_1d(1.0, 2.0, 0.0) == 0.0


def test_trust_2d_interior_and_boundary():
    mm = np.array([[3.0, 1.0], [1.0, 2.0]])
    rhs = np.array([0.5, -0.25])
    s, q = solve_trust_2d(rhs, mm, 10.0)
    su = np.linalg.solve(mm, -rhs)
    assert np.allclose(s, su, rtol=1e-14, atol=0)
    assert q == pytest.approx(rhs @ su + 0.5 * su @ mm @ su, rel=1e-14)
    # boundary: the exact constrained minimum lies within the 181-point scan's resolution
    delta = 0.05
    s, q = solve_trust_2d(rhs, mm, delta)
    assert abs(np.linalg.norm(s) - delta) <= 1e-14
    th = np.linspace(0.0, 2.0 * np.pi, 20001)
    pts = delta * np.stack([np.cos(th), np.sin(th)], axis=1)
    q_exact = min(float(rhs @ p + 0.5 * p @ mm @ p) for p in pts)
    assert q_exact <= q <= q_exact + 4.0 * delta**2 * (math.pi / 181) ** 2 + 2.0 * delta * np.linalg.norm(rhs) * (math.pi / 181) ** 2
