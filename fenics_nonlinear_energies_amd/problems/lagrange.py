"""Equispaced Lagrange elements and symmetric quadrature on the reference triangle.

The reference builds its P4 tabulation with basix (equispaced Lagrange variant,
``slope_stability/support/mesh.py:276-283``) and integrates with basix's default rule of
degree ``2k+1`` (``mesh.py:199-203``, degree 9 for P4).  basix is not a dependency here, so
both are regenerated from their definitions:

* the basis is the unique polynomial of total degree k interpolating the equispaced
  lattice points, obtained by inverting the monomial Vandermonde matrix;
* the degree-9 rule is the fully symmetric 19-point rule (centroid, four 3-point orbits,
  one 6-point orbit).  Published 15-digit values are refined by Gauss-Newton on the
  monomial moment equations until every monomial of degree <= 9 is integrated to
  rounding error.

Reference triangle: vertices (0,0), (1,0), (0,1); area 1/2, weights sum to 1/2.
Point order follows the usual convention for simplicial Lagrange elements: vertices
v0, v1, v2, then edge interior points of edge 0 (v1->v2), edge 1 (v0->v2), edge 2
(v0->v1), then interior points.
"""

from __future__ import annotations

from functools import lru_cache
from math import factorial

import numpy as np


def lattice_points(degree: int) -> np.ndarray:
    """Equispaced interpolation points, (n_p, 2)."""
    k = int(degree)
    if k < 1:
        raise ValueError("degree must be >= 1")
    pts = [(0.0, 0.0), (1.0, 0.0), (0.0, 1.0)]
    for a, b in (((1.0, 0.0), (0.0, 1.0)), ((0.0, 0.0), (0.0, 1.0)), ((0.0, 0.0), (1.0, 0.0))):
        for i in range(1, k):
            t = i / k
            pts.append((a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1])))
    for j in range(1, k):
        for i in range(1, k - j):
            pts.append((i / k, j / k))
    return np.asarray(pts, dtype=np.float64)


def _monomials(degree: int) -> list[tuple[int, int]]:
    return [(i, d - i) for d in range(degree + 1) for i in range(d, -1, -1)]


@lru_cache(maxsize=None)
def _coefficients(degree: int) -> np.ndarray:
    pts = lattice_points(degree)
    mon = _monomials(degree)
    V = np.stack([pts[:, 0] ** i * pts[:, 1] ** j for i, j in mon], axis=1)
    return np.linalg.inv(V)  # column a = coefficients of basis function a


def tabulate(degree: int, xi: np.ndarray):
    """Basis values and reference gradients at points xi (n, 2).

    Returns (phi (n_p, n), dphi_dxi1 (n_p, n), dphi_dxi2 (n_p, n)), the layout of the
    reference's ``hatp, dhat1, dhat2`` (``slope_stability/support/mesh.py:294-298``).
    """
    xi = np.asarray(xi, dtype=np.float64).reshape(-1, 2)
    x, y = xi[:, 0], xi[:, 1]
    mon = _monomials(degree)
    C = _coefficients(degree)
    M = np.stack([x**i * y**j for i, j in mon], axis=0)
    Mx = np.stack([i * x ** max(i - 1, 0) * y**j if i > 0 else np.zeros_like(x) for i, j in mon], axis=0)
    My = np.stack([j * x**i * y ** max(j - 1, 0) if j > 0 else np.zeros_like(x) for i, j in mon], axis=0)
    return C.T @ M, C.T @ Mx, C.T @ My


def monomial_integral(i: int, j: int) -> float:
    """Exact integral of x^i y^j over the reference triangle."""
    return factorial(i) * factorial(j) / factorial(i + j + 2)


# Fully symmetric degree-9 rule, 19 points (Dunavant 1985, Table degree 9), barycentric
# orbit parameters and weights normalised to area 1; refined below.
_D9_CENTROID_W = 0.097135796282799
_D9_ORBIT3 = (
    (0.489682519198738, 0.031334700227139),
    (0.437089591492937, 0.077827541004774),
    (0.188203535619033, 0.079647738927210),
    (0.044729513394453, 0.025577675658698),
)
_D9_ORBIT6 = ((0.036838412054736, 0.221962989160766, 0.043283539377289),)


def _expand(params: np.ndarray):
    """Parameter vector -> (points (19,2), weights (19,)) with weights summing to 1."""
    w0 = params[0]
    pts, wts = [(1.0 / 3.0, 1.0 / 3.0)], [w0]
    k = 1
    for _ in range(len(_D9_ORBIT3)):
        a, w = params[k], params[k + 1]
        k += 2
        b = 1.0 - 2.0 * a
        for lam in ((a, a, b), (a, b, a), (b, a, a)):
            pts.append((lam[1], lam[2]))
            wts.append(w)
    for _ in range(len(_D9_ORBIT6)):
        a, b, w = params[k], params[k + 1], params[k + 2]
        k += 3
        c = 1.0 - a - b
        for lam in ((a, b, c), (a, c, b), (b, a, c), (b, c, a), (c, a, b), (c, b, a)):
            pts.append((lam[1], lam[2]))
            wts.append(w)
    return np.asarray(pts), np.asarray(wts)


def _moment_residual(params: np.ndarray, degree: int) -> np.ndarray:
    pts, wts = _expand(params)
    res = []
    for i, j in _monomials(degree):
        res.append(0.5 * np.dot(wts, pts[:, 0] ** i * pts[:, 1] ** j) - monomial_integral(i, j))
    return np.asarray(res)


@lru_cache(maxsize=None)
def _degree9_rule() -> tuple[np.ndarray, np.ndarray]:
    p = [_D9_CENTROID_W]
    for a, w in _D9_ORBIT3:
        p += [a, w]
    for a, b, w in _D9_ORBIT6:
        p += [a, b, w]
    p = np.asarray(p, dtype=np.float64)
    for _ in range(20):  # Gauss-Newton on the moment equations (central-difference Jacobian)
        r = _moment_residual(p, 9)
        if np.max(np.abs(r)) < 1e-17:
            break
        J = np.empty((r.size, p.size))
        for k in range(p.size):
            h = 1e-7 * max(1.0, abs(p[k]))
            dp = np.zeros_like(p)
            dp[k] = h
            J[:, k] = (_moment_residual(p + dp, 9) - _moment_residual(p - dp, 9)) / (2 * h)
        step = np.linalg.lstsq(J, -r, rcond=None)[0]
        p = p + step
    pts, wts = _expand(p)
    return pts, 0.5 * wts


def quadrature(degree: int) -> tuple[np.ndarray, np.ndarray]:
    """(points (n_q, 2), weights (n_q,)) exact for polynomials of total degree <= ``degree``.

    Degree 1: centroid rule (``mesh.py:193-196``); degree <= 9: the 19-point rule above.
    """
    degree = int(degree)
    if degree <= 1:
        return np.array([[1.0 / 3.0, 1.0 / 3.0]]), np.array([0.5])
    if degree <= 9:
        return _degree9_rule()
    raise ValueError("quadrature degree > 9 is not provided")


def quadrature_degree_for_lagrange(degree: int) -> int:
    """Rule degree used for Lagrange degree k: 1 for k = 1, else 2k + 1 (``mesh.py:199-203``)."""
    return 1 if int(degree) <= 1 else 2 * int(degree) + 1
