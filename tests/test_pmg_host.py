"""Same-mesh p-multigrid transfers (host): P2 -> P4 and P1 -> P2 interpolation on the slope mesh.

Pinned by what interpolation fixes: a coarse-space polynomial is reproduced exactly at every
fine node (quadratics through P2 -> P4, linears through P1 -> P2 -> P4), rows sum to one, and
on the free DOFs the vector transfer maps the coarse rigid-body modes to the fine ones (the
slope's constraints act on whole edges, so no fine free DOF loses a coarse neighbour it needs).
"""

import numpy as np
import pytest

from fenics_nonlinear_energies_amd import pmg
from fenics_nonlinear_energies_amd.problems import slope_stability as S


@pytest.fixture(scope="module")
def meshes():
    vn, vt = S.slope_macro_mesh(2.0, **S.REFERENCE_GEOMETRY)
    n4, e4 = S.lagrange_mesh(vn, vt, 4)
    n2, e2 = S.lagrange_mesh(vn, vt, 2)
    return (vn, vt), (n2, e2), (n4, e4)


def test_scalar_transfers_reproduce_coarse_polynomials(meshes):
    (n1, e1), (n2, e2), (n4, e4) = meshes
    P42 = pmg.scalar_transfer(e4, 4, e2, 2, len(n4), len(n2))
    P21 = pmg.scalar_transfer(e2, 2, e1, 1, len(n2), len(n1))
    quad = lambda X: 1 + 2 * X[:, 0] - 3 * X[:, 1] + 0.5 * X[:, 0] ** 2 - X[:, 0] * X[:, 1] + 0.25 * X[:, 1] ** 2
    lin = lambda X: 0.3 - 0.7 * X[:, 0] + 1.1 * X[:, 1]
    assert np.abs(P42 @ quad(n2) - quad(n4)).max() <= 1e-11 * np.abs(quad(n4)).max()
    assert np.abs(P21 @ lin(n1) - lin(n2)).max() <= 1e-12 * np.abs(lin(n2)).max()
    assert np.abs((P42 @ P21) @ lin(n1) - lin(n4)).max() <= 1e-12 * np.abs(lin(n4)).max()
    np.testing.assert_allclose(np.asarray(P42.sum(axis=1)).ravel(), 1.0, atol=1e-13)
    # a coarse node coincides with a fine node: its row is the unit vector (injection)
    key4 = {tuple(np.round(x, 9)): i for i, x in enumerate(n4)}
    for j in range(0, len(n2), 37):
        row = P42.getrow(key4[tuple(np.round(n2[j], 9))])
        assert row.nnz == 1 and row.indices[0] == j and abs(row.data[0] - 1.0) < 1e-14


def test_free_vector_transfer_reproduces_fields_that_vanish_on_the_constraints(meshes):
    """A coarse field that is zero on the constrained DOFs loses nothing when the constrained
    columns are dropped, so the free-DOF transfer reproduces it like the full one (in the
    levels' own, here random, free-DOF orderings)."""
    (n1, e1), (n2, e2), (n4, e4) = meshes
    P42 = pmg.scalar_transfer(e4, 4, e2, 2, len(n4), len(n2))
    q4, q2 = S.q_mask_from_nodes(n4), S.q_mask_from_nodes(n2)
    free4 = np.flatnonzero(q4.ravel())
    free2 = np.flatnonzero(q2.ravel())
    rng = np.random.default_rng(0)
    perm4, perm2 = rng.permutation(free4.size), rng.permutation(free2.size)
    o4, o2 = pmg.free_order(free4, perm4), pmg.free_order(free2, perm2)
    P = pmg.vector_free_transfer(P42, 2, o4, o2)
    assert P.shape == (free4.size, free2.size)
    # y-displacement zero on y = 0 (its only constraint) and quadratic, so P2 holds it exactly;
    # the x-component is left zero
    def field(X):
        return np.stack((0.0 * X[:, 0], X[:, 1] * (1 + 0.1 * X[:, 0]) - 0.05 * X[:, 1] ** 2), axis=1)

    uc = field(n2).ravel()[o2]
    uf = field(n4).ravel()[o4]
    fine_y = o4 % 2 == 1
    np.testing.assert_allclose((P @ uc)[fine_y], uf[fine_y], atol=1e-12 * np.abs(uf).max())
    # full vector transfer (no constraints): quadratic fields in both components are exact
    Pfull = pmg.vector_free_transfer(P42, 2, np.arange(2 * len(n4)), np.arange(2 * len(n2)))
    g = lambda X: np.stack((0.5 * X[:, 0] ** 2 - X[:, 1], X[:, 0] * X[:, 1] + 2.0), axis=1).ravel()
    np.testing.assert_allclose(Pfull @ g(n2), g(n4), atol=1e-11 * np.abs(g(n4)).max())
