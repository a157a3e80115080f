"""Same-mesh p-multigrid transfers for P_k vector problems (the slope's P4 -> P2 -> P1 levels).

The reference preconditions its P4 Mohr-Coulomb solves with a PETSc PCMG hierarchy on the same
triangles, P4 -> P2 -> P1, followed by a tail of coarser P1 meshes
(``slope_stability/jax_petsc/multigrid.py:1291-1835``); its prolongations evaluate each coarse
Lagrange basis at the fine element's nodes (``multigrid.py:759-846``, ``:895-1022``) and its
coarse operators are Galerkin products.  Here the same two same-mesh transfers are built on the
host once per mesh, as CSR matrices between the free-DOF numberings of the levels; the level
operators are their Galerkin products on the device, and the levels below P1 are the
smoothed-aggregation hierarchy of ``amg.py`` (in place of the reference's geometric P1 tail).

Interpolation: fine node i of element e sits at the reference point xi_i of the fine lattice,
and its value is sum_a phi^coarse_a(xi_i) u_coarse[elems_coarse[e, a]] -- exact for every
coarse field, so a fine node shared by several elements gets the same row from each; the row
of the first element (in element order) is kept.  Vector problems interpolate each component
alike.  Rows and columns are then restricted to the free DOFs: a constrained coarse DOF carries
no correction, and the boundary conditions of the slope act on whole edges, so a fine DOF on a
constrained edge only sees coarse DOFs of that edge.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .problems import lagrange

_DROP = 1e-13  # basis values at lattice points below this are exact zeros in rounding


def scalar_transfer(fine_elems: np.ndarray, fine_degree: int, coarse_elems: np.ndarray, coarse_degree: int,
                    n_fine: int, n_coarse: int) -> sp.csr_matrix:
    """Nodal interpolation from the coarse to the fine scalar Lagrange space on the same
    triangles: (n_fine x n_coarse) CSR, one row per fine node."""
    fine_elems = np.asarray(fine_elems, dtype=np.int64)
    coarse_elems = np.asarray(coarse_elems, dtype=np.int64)
    if fine_elems.shape[0] != coarse_elems.shape[0]:
        raise ValueError("fine and coarse meshes must have the same triangles")
    xi = lagrange.lattice_points(fine_degree)  # (n_pf, 2)
    phi, _, _ = lagrange.tabulate(coarse_degree, xi)  # (n_pc, n_pf)
    W = phi.T  # W[i, a] = coarse basis a at fine lattice point i
    W = np.where(np.abs(W) < _DROP, 0.0, W)
    m, n_pf = fine_elems.shape
    # first slot (in element order) that holds each fine node
    nodes, first = np.unique(fine_elems.ravel(), return_index=True)
    if nodes.size != n_fine or nodes[0] != 0 or nodes[-1] != n_fine - 1:
        raise ValueError("every fine node must belong to an element")
    e = first // n_pf
    i = first % n_pf
    rows = np.repeat(np.arange(n_fine), W.shape[1])
    cols = coarse_elems[e]  # (n_fine, n_pc)
    vals = W[i]  # (n_fine, n_pc)
    keep = vals.ravel() != 0.0
    P = sp.csr_matrix((vals.ravel()[keep], (rows[keep], cols.ravel()[keep])), shape=(n_fine, n_coarse))
    P.sum_duplicates()
    P.sort_indices()
    return P


def vector_free_transfer(P_scalar: sp.csr_matrix, block_size: int, fine_free_order: np.ndarray,
                         coarse_free_order: np.ndarray) -> sp.csr_matrix:
    """Component-wise interpolation restricted to free DOFs, in the levels' own orderings.

    ``fine_free_order[i]`` is the global DOF (node * block_size + component) of fine free DOF i
    in the fine level's numbering; likewise for the coarse level.  Returns the
    (n_free_fine x n_free_coarse) CSR with sorted columns."""
    bs = int(block_size)
    Pv = sp.kron(P_scalar, sp.identity(bs, format="csr"), format="csr")  # DOF d = node * bs + comp
    n_coarse_dofs = Pv.shape[1]
    col_map = np.full(n_coarse_dofs, -1, dtype=np.int64)
    col_map[np.asarray(coarse_free_order, dtype=np.int64)] = np.arange(len(coarse_free_order))
    R = Pv[np.asarray(fine_free_order, dtype=np.int64)].tocoo()
    keep = col_map[R.col] >= 0
    out = sp.csr_matrix((R.data[keep], (R.row[keep], col_map[R.col[keep]])),
                        shape=(len(fine_free_order), len(coarse_free_order)))
    out.sum_duplicates()
    out.sort_indices()
    return out


def free_order(freedofs: np.ndarray, perm: np.ndarray) -> np.ndarray:
    """Global DOF of each free DOF in a level's reordered numbering (reordered i = freedofs[perm[i]])."""
    return np.asarray(freedofs, dtype=np.int64)[np.asarray(perm, dtype=np.int64)]
