"""Structured P1 triangle meshes and operators.

Follows the structured right-triangle conventions of the reference's
plaplace_u3/support/mesh.py:51-160 (every square cell split by the same diagonal into
(ll, lr, ul) and (ur, ul, lr); gradients ``dvx, dvy`` and ``vol`` as in
``_triangle_operators``), generalised to any rectangle and cell count.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def structured_triangle_mesh(nx: int, ny: int, x0: float, x1: float, y0: float, y1: float):
    """Nodes ``(ny+1)*(nx+1)`` (x fastest) and ``2*nx*ny`` triangles (lower ones first)."""
    xs = np.linspace(x0, x1, nx + 1, dtype=np.float64)
    ys = np.linspace(y0, y1, ny + 1, dtype=np.float64)
    gx, gy = np.meshgrid(xs, ys, indexing="xy")
    nodes = np.column_stack((gx.reshape(-1), gy.reshape(-1)))
    ids = np.arange((nx + 1) * (ny + 1), dtype=np.int64).reshape(ny + 1, nx + 1)
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    jj = jj.reshape(-1)
    ii = ii.reshape(-1)
    ll = ids[jj, ii]
    lr = ids[jj, ii + 1]
    ul = ids[jj + 1, ii]
    ur = ids[jj + 1, ii + 1]
    elems = np.vstack((np.column_stack((ll, lr, ul)), np.column_stack((ur, ul, lr)))).astype(np.int64)
    return nodes, elems


def triangle_operators(nodes: np.ndarray, elems: np.ndarray):
    """P1 basis gradients and triangle areas: (dvx, dvy, vol), each (n_elem, 3)/(n_elem,)."""
    v = np.asarray(nodes[np.asarray(elems, dtype=np.int64)], dtype=np.float64)
    x0, y0 = v[:, 0, 0], v[:, 0, 1]
    x1, y1 = v[:, 1, 0], v[:, 1, 1]
    x2, y2 = v[:, 2, 0], v[:, 2, 1]
    det = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0)
    if np.any(det == 0.0):
        raise ValueError("degenerate triangle in mesh")
    dvx = np.stack((y1 - y2, y2 - y0, y0 - y1), axis=1) / det[:, None]
    dvy = np.stack((x2 - x1, x0 - x2, x1 - x0), axis=1) / det[:, None]
    vol = 0.5 * np.abs(det)
    return dvx, dvy, vol


def triangle_geometry_rows(nodes: np.ndarray, elems: np.ndarray) -> np.ndarray | None:
    """(7, n_elem) rows dvx0..2, dvy0..2, vol from the vertex coordinates -- the operation order
    the P1 kernels use with FNE_FLAG_NODE_GEO; None when the mesh has a degenerate triangle or
    is not planar."""
    nodes = np.asarray(nodes, dtype=np.float64)
    if nodes.ndim != 2 or nodes.shape[1] != 2 or np.asarray(elems).shape[-1] != 3:
        return None
    try:
        dvx, dvy, vol = triangle_operators(nodes, elems)
    except ValueError:
        return None
    return np.concatenate([dvx.T, dvy.T, vol[None, :]], axis=0)


def rectangle_boundary_mask(nodes: np.ndarray, x0, x1, y0, y1, tol: float = 1e-12) -> np.ndarray:
    x = nodes[:, 0]
    y = nodes[:, 1]
    return (
        np.isclose(x, x0, atol=tol)
        | np.isclose(x, x1, atol=tol)
        | np.isclose(y, y0, atol=tol)
        | np.isclose(y, y1, atol=tol)
    )


def free_dof_adjacency(elems: np.ndarray, freedofs: np.ndarray, n_total: int) -> sp.coo_matrix:
    """Free-DOF coupling pattern (ones), as stored in the reference's ``adjacency`` group."""
    freedofs = np.asarray(freedofs, dtype=np.int64)
    full_to_free = np.full(n_total, -1, dtype=np.int64)
    full_to_free[freedofs] = np.arange(freedofs.size, dtype=np.int64)
    local = full_to_free[np.asarray(elems, dtype=np.int64)]
    k = local.shape[1]
    rr = np.repeat(local, k, axis=1).reshape(-1)
    cc = np.tile(local, (1, k)).reshape(-1)
    mask = (rr >= 0) & (cc >= 0)
    n = int(freedofs.size)
    keys = np.unique(rr[mask] * np.int64(n) + cc[mask])
    return sp.coo_matrix((np.ones(keys.size), (keys // n, keys % n)), shape=(n, n))


def lumped_nodal_load(elems: np.ndarray, vol: np.ndarray, f: float, n_nodes: int) -> np.ndarray:
    """Consistent P1 load of a constant source: f_i = f * sum_{e contains i} vol_e / 3."""
    w = np.repeat(np.asarray(vol, dtype=np.float64) / 3.0, elems.shape[1])
    return float(f) * np.bincount(np.asarray(elems, dtype=np.int64).reshape(-1), weights=w, minlength=n_nodes)
