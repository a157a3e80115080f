"""Free-DOF permutations, the global layout and the CSR pattern (host set-up, numpy).

Semantics follow core/petsc/reordered_element_base.py of the reference:
``select_permutation`` (:183-200), ``inverse_permutation`` (:203-206),
``local_vec_from_full`` (:224-235) and ``build_global_layout`` (:380-506), reduced to a
single device (lo = 0, hi = n_free).  ``block_morton`` is an additional ordering that
keeps row patches spatially compact, which the patch assembly kernels rely on for a low
halo fraction.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy import sparse
from scipy.sparse.csgraph import reverse_cuthill_mckee


# ------------------------------------------------------------------------------------------
# permutations
# ------------------------------------------------------------------------------------------
def _build_block_graph(adjacency: sparse.spmatrix, block_size: int) -> sparse.csr_matrix:
    rows, cols = adjacency.nonzero()
    n_blocks = adjacency.shape[0] // block_size
    block = sparse.coo_matrix(
        (np.ones(rows.size, dtype=np.int8), (rows // block_size, cols // block_size)),
        shape=(n_blocks, n_blocks),
    ).tocsr()
    block.data[:] = 1
    return block


def _expand_block_perm(block_perm: np.ndarray, block_size: int) -> np.ndarray:
    if int(block_size) <= 1:
        return np.asarray(block_perm, dtype=np.int64)
    perm = np.empty(block_perm.size * block_size, dtype=np.int64)
    for comp in range(block_size):
        perm[comp::block_size] = block_perm * block_size + comp
    return perm


def perm_identity(n_free: int) -> np.ndarray:
    return np.arange(n_free, dtype=np.int64)


def perm_block_rcm(adjacency: sparse.spmatrix, block_size: int) -> np.ndarray:
    graph = adjacency.tocsr() if int(block_size) <= 1 else _build_block_graph(adjacency, int(block_size))
    block_perm = np.asarray(reverse_cuthill_mckee(graph, symmetric_mode=True), dtype=np.int64)
    return _expand_block_perm(block_perm, int(block_size))


def perm_block_xyz(coords_all: np.ndarray, freedofs: np.ndarray, block_size: int) -> np.ndarray:
    """Lexicographic coordinate sort of free nodes, x slowest-varying key last (reference :163-172)."""
    block_size = int(block_size)
    node_ids = np.asarray(freedofs, dtype=np.int64)[::block_size] // block_size
    coords = np.asarray(coords_all[node_ids], dtype=np.float64)
    keys = tuple(coords[:, dim] for dim in reversed(range(coords.shape[1])))
    return _expand_block_perm(np.lexsort(keys).astype(np.int64), block_size)


def _morton_codes(coords: np.ndarray, bits: int = 20) -> np.ndarray:
    """Z-order codes on an isotropic grid: every axis is quantised with the largest extent,
    so an aligned block of codes is a cube in physical space.  (Scaling each axis by its own
    extent would turn the blocks of a 40:1:1 beam into lines along its axis.)"""
    lo = coords.min(axis=0)
    span = max(float(np.max(coords.max(axis=0) - lo)), 1e-300)
    q = np.floor((coords - lo) / span * ((1 << bits) - 1) + 0.5).astype(np.uint64)
    dim = coords.shape[1]
    code = np.zeros(coords.shape[0], dtype=np.uint64)
    for b in range(bits):
        for d in range(dim):
            code |= ((q[:, d] >> np.uint64(b)) & np.uint64(1)) << np.uint64(b * dim + d)
    return code


def perm_block_morton(coords_all: np.ndarray, freedofs: np.ndarray, block_size: int) -> np.ndarray:
    """Z-order (Morton) sort of free nodes: contiguous rows form compact spatial patches.

    Works per node: all free components of a node stay adjacent, in component order, so the
    rows of one node are consecutive even when only some of its components are free.
    """
    block_size = int(block_size)
    freedofs = np.asarray(freedofs, dtype=np.int64)
    node_of = freedofs // block_size
    nodes, first = np.unique(node_of, return_index=True)
    coords = np.asarray(coords_all[nodes], dtype=np.float64)
    code = _morton_codes(coords)
    node_order = np.lexsort((nodes, code))  # stable tie-break on node id
    rank = np.empty(nodes.size, dtype=np.int64)
    rank[node_order] = np.arange(nodes.size)
    node_rank = rank[np.searchsorted(nodes, node_of)]
    comp = freedofs % block_size
    return np.lexsort((comp, node_rank)).astype(np.int64)


def select_permutation(
    reorder_mode: str,
    *,
    adjacency: sparse.spmatrix | None,
    coords_all: np.ndarray,
    freedofs: np.ndarray,
    n_parts: int = 1,
    block_size: int = 1,
) -> np.ndarray:
    if reorder_mode == "none":
        return perm_identity(len(freedofs))
    if reorder_mode == "block_rcm":
        if adjacency is None:
            raise ValueError("block_rcm needs an adjacency matrix")
        return perm_block_rcm(adjacency, block_size)
    if reorder_mode == "block_xyz":
        return perm_block_xyz(coords_all, freedofs, block_size)
    if reorder_mode == "block_morton":
        return perm_block_morton(coords_all, freedofs, block_size)
    if reorder_mode == "block_metis":
        raise ValueError("block_metis needs pymetis, which is not available; use block_morton")
    raise ValueError(f"Unsupported element reorder mode: {reorder_mode!r}")


def inverse_permutation(perm: np.ndarray) -> np.ndarray:
    iperm = np.empty_like(perm)
    iperm[perm] = np.arange(perm.size, dtype=np.int64)
    return iperm


def validate_permutation(perm: np.ndarray, n_free: int) -> np.ndarray:
    perm = np.asarray(perm, dtype=np.int64).ravel()
    if perm.size != int(n_free):
        raise ValueError(f"Permutation size {perm.size} does not match number of free DOFs {n_free}")
    if perm.size and (
        int(perm.min()) != 0 or int(perm.max()) != int(n_free) - 1 or np.unique(perm).size != perm.size
    ):
        raise ValueError("Permutation override must contain each free DOF exactly once")
    return perm


def local_vec_from_full(full_reordered, total_to_free_reord, local_total_nodes, dirichlet_full):
    """Merge free (reordered) values and Dirichlet values into the local nodal vector."""
    local_reord = total_to_free_reord[local_total_nodes]
    v_local = np.asarray(dirichlet_full[local_total_nodes], dtype=np.float64).copy()
    free_mask = local_reord >= 0
    if np.any(free_mask):
        v_local[free_mask] = full_reordered[local_reord[free_mask]]
    return v_local


# ------------------------------------------------------------------------------------------
# global layout + CSR pattern
# ------------------------------------------------------------------------------------------
@dataclass
class GlobalLayout:
    perm: np.ndarray
    iperm: np.ndarray
    lo: int
    hi: int
    n_free: int
    total_to_free_reord: np.ndarray
    indptr: np.ndarray  # CSR of the owned rows (all rows on one device), int32/int64
    indices: np.ndarray
    elem_owner: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.indices.size)


def csr_from_keys(keys: np.ndarray, n: int) -> tuple[np.ndarray, np.ndarray]:
    keys = np.unique(np.asarray(keys, dtype=np.int64))
    rows = keys // n
    cols = keys % n
    indptr = np.zeros(n + 1, dtype=np.int64)
    indptr[1:] = np.cumsum(np.bincount(rows, minlength=n))
    return indptr, cols


def element_pattern_keys(elem_free: np.ndarray, n_free: int, chunk: int | None = None) -> np.ndarray:
    """Sorted unique row*n+col keys of all (free, free) pairs of every element."""
    if chunk is None:  # about 16M candidate pairs per block
        nl = max(1, int(elem_free.shape[1]) if elem_free.ndim == 2 else 1)
        chunk = max(1024, (1 << 24) // (nl * nl))
    parts = []
    n = np.int64(n_free)
    for s in range(0, elem_free.shape[0], chunk):
        blk = elem_free[s : s + chunk]
        r = blk[:, :, None].astype(np.int64)
        c = blk[:, None, :].astype(np.int64)
        valid = (r >= 0) & (c >= 0)
        k = (np.broadcast_to(r, valid.shape)[valid] * n + np.broadcast_to(c, valid.shape)[valid])
        parts.append(np.unique(k))
    if not parts:
        return np.zeros(0, dtype=np.int64)
    return np.unique(np.concatenate(parts))


def build_global_layout(
    params: dict,
    adjacency: sparse.spmatrix | None,
    perm: np.ndarray,
    *,
    block_size: int,
    dirichlet_key: str,
) -> GlobalLayout:
    freedofs = np.asarray(params["freedofs"], dtype=np.int64).ravel()
    elems = np.asarray(params["elems"], dtype=np.int64)
    n_total = int(np.asarray(params[dirichlet_key]).size)
    n_free = int(freedofs.size)
    iperm = inverse_permutation(perm)
    total_to_free_reord = np.full(n_total, -1, dtype=np.int64)
    total_to_free_reord[freedofs] = iperm
    elem_free = total_to_free_reord[elems]
    keys = element_pattern_keys(elem_free, n_free)
    if adjacency is not None:
        adj = adjacency.tocoo()
        r = iperm[np.asarray(adj.row, dtype=np.int64)]
        c = iperm[np.asarray(adj.col, dtype=np.int64)]
        adj_keys = np.unique(r * np.int64(n_free) + c)
        missing = np.setdiff1d(keys, adj_keys, assume_unique=True)
        if missing.size:
            raise ValueError(
                f"adjacency pattern lacks {missing.size} element couplings; it must contain the element pattern"
            )
        keys = adj_keys
    indptr, indices = csr_from_keys(keys, n_free)
    masked = np.where(elem_free >= 0, elem_free, n_free)
    elem_min = masked.min(axis=1) if elems.size else np.zeros(0, dtype=np.int64)
    elem_owner = np.where(elem_min < n_free, 0, -1).astype(np.int64)
    idx_dtype = np.int32 if max(n_free, keys.size) < np.iinfo(np.int32).max else np.int64
    return GlobalLayout(
        perm=np.asarray(perm, dtype=np.int64),
        iperm=iperm,
        lo=0,
        hi=n_free,
        n_free=n_free,
        total_to_free_reord=total_to_free_reord,
        indptr=indptr.astype(idx_dtype),
        indices=indices.astype(idx_dtype),
        elem_owner=elem_owner,
    )


def element_dof_codes(elems: np.ndarray, total_to_free_reord: np.ndarray, n_free: int) -> np.ndarray:
    """edof code per element DOF: reordered free index, or n_free + total index if constrained."""
    elems = np.asarray(elems, dtype=np.int64)
    fr = total_to_free_reord[elems]
    return np.where(fr >= 0, fr, n_free + elems)
