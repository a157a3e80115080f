"""Owner-computes patch schedule for deterministic, atomic-free gradient/Hessian assembly.

Replaces the reference's scatter plan (``_build_scatter_data``,
core/petsc/reordered_element_base.py:1464-1534) whose ``np.add.at`` segmented sum is
replaced on the GPU by an owner-computes reduction:

* the free DOFs (CSR rows, reordered) are cut into patches of ``rows_per_patch``
  consecutive rows -- one workgroup each, one row per thread;
* a patch evaluates every element that touches one of its rows (owned elements plus a
  halo, as the reference's overlap domains do per MPI rank, :509-597) into LDS;
* each row thread then sums, in a fixed order, the contributions listed in its incidence
  list -- so each CSR value and gradient entry is written by exactly one thread, once.

Elements are renumbered so that the owned elements of a patch are contiguous (the device
element arrays use the new order; ``elem_order[new] = old``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

P1_SLOT_BITS = 13
P1_POS_BITS = 5
P1_POS_NONE = (1 << P1_POS_BITS) - 1


@dataclass
class PatchPlan:
    n_patch: int
    rows_per_patch: int
    chunk: int
    prow: np.ndarray  # int32 [n_patch+1]
    pelem_ptr: np.ndarray  # int32 [n_patch+1]
    pelem: np.ndarray  # int32, new element ids
    elem_order: np.ndarray  # int64 [n_elem], new -> old
    inc_ptr: np.ndarray  # int32 [n_rows+1]
    inc: np.ndarray  # uint32 items
    inc_aux: np.ndarray | None
    max_patch_nnz: int
    max_row_len: int
    pown: np.ndarray | None = None  # int32 [n_patch] owned elements per patch (listed first)

    @property
    def max_patch_elems(self) -> int:
        return int(np.diff(self.pelem_ptr).max(initial=0))

    @property
    def halo_factor(self) -> float:
        owned = int(np.count_nonzero(self.elem_order >= 0))
        return float(self.pelem.size) / max(1, owned)


def _patch_rows(n_rows: int, rows_per_patch: int) -> np.ndarray:
    n_patch = max(1, -(-n_rows // rows_per_patch)) if n_rows else 0
    prow = np.minimum(np.arange(n_patch + 1, dtype=np.int64) * rows_per_patch, n_rows)
    return prow


def _elements_by_patch(row_of: np.ndarray, rows_per_patch: int, n_patch: int):
    """Owner patch per element, new element order, and the (patch, element) incidence pairs.

    row_of: (n_elem, k) assembly rows touched per element (-1 = none).
    """
    n_elem = row_of.shape[0]
    valid = row_of >= 0
    masked = np.where(valid, row_of, np.iinfo(np.int64).max)
    min_row = masked.min(axis=1) if n_elem else np.zeros(0, dtype=np.int64)
    free_any = min_row < np.iinfo(np.int64).max
    owner = np.where(free_any, min_row // rows_per_patch, n_patch)
    # elements with only constrained DOFs still carry energy: spread them over the patches as
    # owned elements without incidences (they add to the fused energy only)
    dir_only = np.nonzero(~free_any)[0]
    if n_patch > 0 and dir_only.size:
        owner[dir_only] = np.arange(dir_only.size) % n_patch
    elem_order = np.argsort(owner, kind="stable").astype(np.int64)
    new_of_old = np.empty(n_elem, dtype=np.int64)
    new_of_old[elem_order] = np.arange(n_elem, dtype=np.int64)
    owner_new = owner[elem_order]
    row_new = row_of[elem_order]
    valid_new = row_new >= 0
    e_idx = np.broadcast_to(np.arange(n_elem, dtype=np.int64)[:, None], row_new.shape)[valid_new]
    p_idx = row_new[valid_new] // rows_per_patch
    lone = np.nonzero(~valid_new.any(axis=1) & (owner_new < n_patch))[0]
    pair_key = np.unique(np.concatenate((p_idx * np.int64(n_elem) + e_idx, owner_new[lone] * np.int64(n_elem) + lone)))
    pp = pair_key // n_elem
    pe = pair_key % n_elem
    halo = (owner_new[pe] != pp).astype(np.int64)
    order = np.lexsort((pe, halo, pp))
    pelem = pe[order]
    pelem_ptr = np.zeros(n_patch + 1, dtype=np.int64)
    pelem_ptr[1:] = np.cumsum(np.bincount(pp, minlength=n_patch))
    slot_sorted = np.empty(pair_key.size, dtype=np.int64)
    slot_sorted[order] = np.arange(pair_key.size, dtype=np.int64) - pelem_ptr[pp[order]]
    pown = np.bincount(pp[halo == 0], minlength=n_patch)
    return elem_order, new_of_old, row_new, pelem, pelem_ptr, pair_key, slot_sorted, pown


def build_p1_plan(
    elem_free: np.ndarray,
    indptr: np.ndarray,
    indices: np.ndarray,
    n_free: int,
    *,
    rows_per_patch: int = 256,
    chunk: int = 512,
) -> PatchPlan:
    """Patch plan for 3-DOF scalar elements (P1 triangles)."""
    elem_free = np.asarray(elem_free, dtype=np.int64)
    n_elem, n_loc = elem_free.shape
    if n_loc != 3:
        raise ValueError("build_p1_plan expects 3 DOFs per element")
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    prow = _patch_rows(n_free, rows_per_patch)
    n_patch = prow.size - 1
    elem_order, _, ef, pelem, pelem_ptr, pair_key, slot_sorted, pown = _elements_by_patch(
        elem_free, rows_per_patch, n_patch
    )
    counts = np.diff(pelem_ptr)
    if counts.size and counts.max() >= (1 << P1_SLOT_BITS):
        raise ValueError("too many elements in one patch for 13-bit slots; reduce rows_per_patch")
    row_len = np.diff(indptr)
    max_row_len = int(row_len.max()) if row_len.size else 0
    if max_row_len >= P1_POS_NONE:
        raise ValueError(f"CSR row of length {max_row_len} exceeds the P1 kernel limit {P1_POS_NONE - 1}")

    # incidences (row r of element e at local position a)
    e_i, a_i = np.nonzero(ef >= 0)
    r_i = ef[e_i, a_i]
    p_i = r_i // rows_per_patch
    slot = slot_sorted[np.searchsorted(pair_key, p_i * np.int64(n_elem) + e_i)]
    csr_keys = np.repeat(np.arange(n_free, dtype=np.int64), row_len) * np.int64(n_free) + indices
    item = slot.astype(np.uint64) | (a_i.astype(np.uint64) << np.uint64(P1_SLOT_BITS))
    for b in range(3):
        c = ef[e_i, b]
        pos = np.full(e_i.size, P1_POS_NONE, dtype=np.int64)
        ok = c >= 0
        k = np.searchsorted(csr_keys, r_i[ok] * np.int64(n_free) + c[ok])
        if not np.array_equal(csr_keys[k], r_i[ok] * np.int64(n_free) + c[ok]):
            raise RuntimeError("element coupling missing from the CSR pattern")
        pos[ok] = k - indptr[r_i[ok]]
        item |= pos.astype(np.uint64) << np.uint64(P1_SLOT_BITS + 2 + P1_POS_BITS * b)
    order = np.lexsort((slot, r_i))
    inc = item[order].astype(np.uint32)
    inc_ptr = np.zeros(n_free + 1, dtype=np.int64)
    inc_ptr[1:] = np.cumsum(np.bincount(r_i, minlength=n_free))
    patch_nnz = indptr[prow[1:]] - indptr[prow[:-1]]
    return PatchPlan(
        n_patch=n_patch,
        rows_per_patch=rows_per_patch,
        chunk=int(chunk),
        prow=prow.astype(np.int32),
        pelem_ptr=pelem_ptr.astype(np.int32),
        pelem=pelem.astype(np.int32),
        elem_order=elem_order,
        inc_ptr=inc_ptr.astype(np.int32),
        inc=inc,
        inc_aux=None,
        max_patch_nnz=int(patch_nnz.max()) if patch_nnz.size else 0,
        max_row_len=max_row_len,
        pown=pown.astype(np.int32),
    )


BLOCK_SLOT_BITS = 16
BLOCK_POS_NONE = 255


def build_block_plan(
    elem_free: np.ndarray,
    indptr: np.ndarray,
    indices: np.ndarray,
    n_free: int,
    *,
    block_size: int,
    nodes_per_patch: int = 64,
    chunk: int = 128,
) -> PatchPlan:
    """Patch plan for vector elements whose nodes are fully free or fully constrained.

    Rows come in node blocks of ``block_size`` consecutive free DOFs.  Incidence lists are
    per node row; an item (uint64) is slot[0:16) | a[16:18) | bpos_b[24+8b : 32+8b), where
    bpos_b is the block column of the element's node b inside the node's block row (255 for
    a constrained node).  Supports up to 4 nodes per element (P1 tetrahedra).
    """
    bs = int(block_size)
    elem_free = np.asarray(elem_free, dtype=np.int64)
    n_elem, n_loc = elem_free.shape
    n_nodes_e = n_loc // bs
    if n_nodes_e > 4 or n_nodes_e * bs != n_loc:
        raise ValueError("build_block_plan supports elements with <= 4 nodes")
    if n_free % bs:
        raise ValueError("free DOFs must come in whole node blocks")
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    node = elem_free[:, 0::bs]
    free_node = node >= 0
    # node rows: all components free and consecutive, or all constrained
    for i in range(bs):
        comp = elem_free[:, i::bs]
        if not np.array_equal(comp >= 0, free_node):
            raise ValueError("mixed free/constrained node components are not supported by the block plan")
        if np.any(comp[free_node] != node[free_node] + i):
            raise ValueError("node components are not consecutive in the free-DOF order")
        if np.any(node[free_node] % bs):
            raise ValueError("node blocks are not aligned to block_size")
    nrow = np.where(free_node, node // bs, -1)
    n_nrows = n_free // bs
    # block pattern from the scalar CSR (first row of each node block)
    brow_len = (indptr[bs * np.arange(n_nrows) + 1] - indptr[bs * np.arange(n_nrows)]) // bs
    bptr = np.zeros(n_nrows + 1, dtype=np.int64)
    bptr[1:] = np.cumsum(brow_len)
    first = indptr[bs * np.arange(n_nrows)]
    sel = np.repeat(first, brow_len) + bs * (np.arange(bptr[-1]) - np.repeat(bptr[:-1], brow_len))
    bcol = indices[sel] // bs
    if int(brow_len.max(initial=0)) >= BLOCK_POS_NONE:
        raise ValueError("node degree too large for 8-bit block positions")

    prow_nodes = _patch_rows(n_nrows, nodes_per_patch)
    n_patch = prow_nodes.size - 1
    elem_order, _, nr, pelem, pelem_ptr, pair_key, slot_sorted, pown = _elements_by_patch(nrow, nodes_per_patch, n_patch)
    if np.diff(pelem_ptr).max(initial=0) >= (1 << BLOCK_SLOT_BITS):
        raise ValueError("too many elements in one patch")
    e_i, a_i = np.nonzero(nr >= 0)
    na = nr[e_i, a_i]
    p_i = na // nodes_per_patch
    slot = slot_sorted[np.searchsorted(pair_key, p_i * np.int64(n_elem) + e_i)]
    bkeys = np.repeat(np.arange(n_nrows, dtype=np.int64), brow_len) * np.int64(n_nrows) + bcol
    item = slot.astype(np.uint64) | (a_i.astype(np.uint64) << np.uint64(16))
    for b in range(n_nodes_e):
        nb = nr[e_i, b]
        pos = np.full(e_i.size, BLOCK_POS_NONE, dtype=np.int64)
        ok = nb >= 0
        key = na[ok] * np.int64(n_nrows) + nb[ok]
        k = np.searchsorted(bkeys, key)
        if not np.array_equal(bkeys[k], key):
            raise RuntimeError("element node coupling missing from the block pattern")
        pos[ok] = k - bptr[na[ok]]
        item |= pos.astype(np.uint64) << np.uint64(24 + 8 * b)
    for b in range(n_nodes_e, 4):
        item |= np.uint64(BLOCK_POS_NONE) << np.uint64(24 + 8 * b)
    order = np.lexsort((slot, na))
    inc = item[order]
    inc_ptr = np.zeros(n_nrows + 1, dtype=np.int64)
    inc_ptr[1:] = np.cumsum(np.bincount(na, minlength=n_nrows))
    prow = bs * prow_nodes
    patch_nnz = indptr[prow[1:]] - indptr[prow[:-1]]
    return PatchPlan(
        n_patch=n_patch,
        rows_per_patch=bs * nodes_per_patch,
        chunk=int(chunk),
        prow=prow.astype(np.int32),
        pelem_ptr=pelem_ptr.astype(np.int32),
        pelem=pelem.astype(np.int32),
        elem_order=elem_order,
        inc_ptr=inc_ptr.astype(np.int32),
        inc=inc,
        inc_aux=None,
        max_patch_nnz=int(patch_nnz.max()) if patch_nnz.size else 0,
        max_row_len=int(np.diff(indptr).max(initial=0)),
        pown=pown.astype(np.int32),
    )


def build_p1_rowgather_plan(
    elem_free: np.ndarray,
    indptr: np.ndarray,
    indices: np.ndarray,
    n_free: int,
    *,
    locality_rows: int = 256,
) -> PatchPlan:
    """Row-gather schedule for P1 scalar elements (one thread per CSR row, no staging).

    Items (uint64): e[0:32) | a[32:34) | pos_b[40+8b : 48+8b) with 255 for a constrained
    column; each row's items are sorted by element.  Elements are renumbered by the
    ``locality_rows`` block of their first row so that neighbouring rows share cache lines.
    """
    elem_free = np.asarray(elem_free, dtype=np.int64)
    n_elem = elem_free.shape[0]
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    n_blk = max(1, -(-n_free // locality_rows))
    elem_order, _, ef, _, _, _, _, _ = _elements_by_patch(elem_free, locality_rows, n_blk)
    row_len = np.diff(indptr)
    if int(row_len.max(initial=0)) >= 255:
        raise ValueError("row-gather items hold 8-bit column positions")
    e_i, a_i = np.nonzero(ef >= 0)
    r_i = ef[e_i, a_i]
    csr_keys = np.repeat(np.arange(n_free, dtype=np.int64), row_len) * np.int64(n_free) + indices
    item = e_i.astype(np.uint64) | (a_i.astype(np.uint64) << np.uint64(32))
    for b in range(3):
        c = ef[e_i, b]
        pos = np.full(e_i.size, 255, dtype=np.int64)
        ok = c >= 0
        key = r_i[ok] * np.int64(n_free) + c[ok]
        k = np.searchsorted(csr_keys, key)
        if not np.array_equal(csr_keys[k], key):
            raise RuntimeError("element coupling missing from the CSR pattern")
        pos[ok] = k - indptr[r_i[ok]]
        item |= pos.astype(np.uint64) << np.uint64(40 + 8 * b)
    order = np.lexsort((e_i, r_i))
    inc_ptr = np.zeros(n_free + 1, dtype=np.int64)
    inc_ptr[1:] = np.cumsum(np.bincount(r_i, minlength=n_free))
    return PatchPlan(
        n_patch=0,
        rows_per_patch=0,
        chunk=64,
        prow=np.zeros(1, dtype=np.int32),
        pelem_ptr=np.zeros(1, dtype=np.int32),
        pelem=np.zeros(1, dtype=np.int32),
        elem_order=elem_order,
        inc_ptr=inc_ptr.astype(np.int32),
        inc=item[order],
        inc_aux=None,
        max_patch_nnz=0,
        max_row_len=int(row_len.max(initial=0)),
    )


def color_patch_elements(pelem_ptr: np.ndarray, owned_rows: np.ndarray, n_rows: int) -> tuple[np.ndarray, np.ndarray]:
    """Greedy per-patch coloring (host C++ routine of libfne_hip): colors[i], colors-per-patch."""
    import ctypes

    from . import native

    n_patch = int(pelem_ptr.size - 1)
    ptr64 = np.ascontiguousarray(pelem_ptr, dtype=np.int64)
    rows = np.ascontiguousarray(owned_rows, dtype=np.int64)
    colors = np.zeros(rows.shape[0], dtype=np.int32)
    rc = native.lib().fne_color_patch_elements(
        n_patch, ptr64.ctypes.data_as(ctypes.c_void_p), rows.ctypes.data_as(ctypes.c_void_p),
        int(rows.shape[1]), int(n_rows), colors.ctypes.data_as(ctypes.c_void_p),
    )
    if rc < 0:
        native.check(rc, "fne_color_patch_elements")
    pid = np.repeat(np.arange(n_patch), np.diff(ptr64))
    ncol = np.zeros(n_patch, dtype=np.int64)
    if colors.size:
        np.maximum.at(ncol, pid, colors.astype(np.int64) + 1)
    return colors, ncol


def build_p1_colored_plan(
    elem_free: np.ndarray,
    indptr: np.ndarray,
    indices: np.ndarray,
    n_free: int,
    *,
    rows_per_patch: int = 192,
    elems_per_thread: int = 2,
    block: int = 256,
) -> PatchPlan:
    """Colored patch plan for P1 scalar elements (see p1_color_kernel).

    Item per patch element (uint64): pos_ab at bits 5(3a+b) (column position of vertex b inside
    the row of vertex a when that row belongs to the patch, else 31) | energy owner at bit 48 |
    color at bits 52..59.  ``chunk`` of the returned plan holds the elements per thread.
    """
    elem_free = np.asarray(elem_free, dtype=np.int64)
    n_elem, n_loc = elem_free.shape
    if n_loc != 3:
        raise ValueError("build_p1_colored_plan expects 3 DOFs per element")
    if rows_per_patch > block:
        raise ValueError("rows_per_patch must not exceed the workgroup size")
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    row_len = np.diff(indptr)
    if int(row_len.max(initial=0)) >= 31:
        raise ValueError("CSR rows longer than 30 entries do not fit the 5-bit positions")
    prow = _patch_rows(n_free, rows_per_patch)
    n_patch = prow.size - 1
    elem_order, _, ef, pelem, pelem_ptr, _, _, pown = _elements_by_patch(elem_free, rows_per_patch, n_patch)
    if np.diff(pelem_ptr).max(initial=0) > elems_per_thread * block:
        raise ValueError("patch has more elements than elems_per_thread * block")
    pid = np.repeat(np.arange(n_patch), np.diff(pelem_ptr))
    efp = ef[pelem]  # (n_pe, 3) rows of each patch element
    lo = prow[pid][:, None]
    hi = prow[pid + 1][:, None]
    owned = (efp >= lo) & (efp < hi)
    owned_rows = np.where(owned, efp, -1)
    colors, ncol = color_patch_elements(pelem_ptr, owned_rows, n_free)
    # order each patch's elements by color: the kernel's color steps then touch few waves
    # (a wave whose 64 slots mix many colors would execute the scatter once per color)
    slot0 = np.arange(pelem.size) - pelem_ptr[pid]
    owner_bit = slot0 < pown[pid]
    reorder = np.lexsort((pelem, colors, pid))
    pelem, efp, owned, colors, owner_bit = pelem[reorder], efp[reorder], owned[reorder], colors[reorder], owner_bit[reorder]
    csr_keys = np.repeat(np.arange(n_free, dtype=np.int64), row_len) * np.int64(n_free) + indices
    item = np.zeros(pelem.size, dtype=np.uint64)
    for a in range(3):
        ra = efp[:, a]
        for b in range(3):
            cb = efp[:, b]
            pos = np.full(pelem.size, 31, dtype=np.int64)
            ok = owned[:, a] & (cb >= 0)
            key = ra[ok] * np.int64(n_free) + cb[ok]
            k = np.searchsorted(csr_keys, key)
            if not np.array_equal(csr_keys[k], key):
                raise RuntimeError("element coupling missing from the CSR pattern")
            pos[ok] = k - indptr[ra[ok]]
            item |= pos.astype(np.uint64) << np.uint64(5 * (3 * a + b))
    item |= owner_bit.astype(np.uint64) << np.uint64(48)
    item |= colors.astype(np.uint64) << np.uint64(52)
    patch_nnz = indptr[prow[1:]] - indptr[prow[:-1]]
    plan = PatchPlan(
        n_patch=n_patch,
        rows_per_patch=rows_per_patch,
        chunk=int(elems_per_thread),
        prow=prow.astype(np.int32),
        pelem_ptr=pelem_ptr.astype(np.int32),
        pelem=pelem.astype(np.int32),
        elem_order=elem_order,
        inc_ptr=np.zeros(1, dtype=np.int32),
        inc=np.zeros(1, dtype=np.uint32),
        inc_aux=None,
        max_patch_nnz=int(patch_nnz.max(initial=0)),
        max_row_len=int(row_len.max(initial=0)),
        pown=pown.astype(np.int32),
    )
    plan.pitem = item
    plan.pcolors = ncol.astype(np.int32)
    return plan


ROW_POS_NONE = 255


@dataclass
class RowIncidencePlan(PatchPlan):
    """Per-row incidence lists for elements with up to 32 DOFs (row-gather assembly).

    ``inc[k]`` = element * 32 + local DOF of the k-th (row, element) incidence, rows in
    CSR order and elements ascending within a row; ``pitem`` holds, per incidence, the 32
    byte positions of the element's local DOFs inside the row (255 = constrained / none).
    """

    pitem: np.ndarray | None = None  # uint8 [n_items, 32]


def build_row_incidence_plan(elem_free: np.ndarray, indptr: np.ndarray, indices: np.ndarray, n_free: int,
                             *, item_chunk: int = 1 << 20) -> RowIncidencePlan:
    """Row-gather schedule: each CSR row sums the rows of its incident element matrices.

    Replaces the reference's scatter plan (``_build_scatter_data``,
    core/petsc/reordered_element_base.py:1464-1534): the same (element, i, j) -> CSR
    position map, grouped by destination row so that the device reduction is a fixed-order
    gather (no atomics).  Elements are renumbered by their smallest free row for locality.
    """
    elem_free = np.asarray(elem_free, dtype=np.int64)
    n_elem, n_loc = elem_free.shape
    if n_loc > 32:
        raise ValueError("row incidence plan supports at most 32 DOFs per element")
    if n_elem >= (1 << 27):
        raise ValueError("too many elements for 27-bit element ids")
    indptr = np.asarray(indptr, dtype=np.int64)
    indices = np.asarray(indices, dtype=np.int64)
    row_len = np.diff(indptr)
    max_row_len = int(row_len.max(initial=0))
    if max_row_len >= ROW_POS_NONE:
        raise ValueError("CSR rows longer than 254 entries are not supported by 8-bit positions")
    masked = np.where(elem_free >= 0, elem_free, n_free)
    key = masked.min(axis=1) if n_elem else np.zeros(0, dtype=np.int64)
    elem_order = np.argsort(key, kind="stable").astype(np.int64)
    ef = elem_free[elem_order]
    e_i, c_i = np.nonzero(ef >= 0)
    r_i = ef[e_i, c_i]
    order = np.lexsort((c_i, e_i, r_i))
    e_i, c_i, r_i = e_i[order], c_i[order], r_i[order]
    inc_ptr = np.zeros(n_free + 1, dtype=np.int64)
    inc_ptr[1:] = np.cumsum(np.bincount(r_i, minlength=n_free))
    inc = (e_i.astype(np.uint64) * np.uint64(32) + c_i.astype(np.uint64)).astype(np.uint32)
    n_items = e_i.size
    pos = np.full((n_items, 32), ROW_POS_NONE, dtype=np.uint8)
    nn = np.int64(n_free)
    csr_keys = np.repeat(np.arange(n_free, dtype=np.int64), row_len) * nn + indices  # sorted
    for s in range(0, n_items, item_chunk):
        sl = slice(s, min(n_items, s + item_chunk))
        cols = ef[e_i[sl]]  # (k, n_loc)
        rows = r_i[sl]
        valid = cols >= 0
        k = np.searchsorted(csr_keys, (rows[:, None] * nn + np.where(valid, cols, 0)).ravel()).reshape(cols.shape)
        k = np.minimum(k, max(csr_keys.size - 1, 0))
        rel = k - indptr[rows][:, None]
        ok = valid & (csr_keys[k] == rows[:, None] * nn + cols)
        if np.any(valid & ~ok):
            raise ValueError("element coupling missing from the CSR pattern")
        pos[sl, :n_loc] = np.where(ok, rel, ROW_POS_NONE).astype(np.uint8)
    empty = np.zeros(1, dtype=np.int32)
    return RowIncidencePlan(
        n_patch=0,
        rows_per_patch=0,
        chunk=0,
        prow=empty,
        pelem_ptr=empty,
        pelem=np.zeros(0, dtype=np.int32),
        elem_order=elem_order,
        inc_ptr=inc_ptr.astype(np.int32),
        inc=inc,
        inc_aux=None,
        max_patch_nnz=0,
        max_row_len=max_row_len,
        pown=None,
        pitem=pos,
    )

