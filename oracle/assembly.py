"""Oracle global assembly (numpy + scipy), in the ORIGINAL free-DOF numbering.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``local_state``: ``u_0.at[freedofs].set(u)`` of every serial J (e.g. plaplace/jax/jax_energy.py:8).
* ``scatter_plan`` / ``scatter_add``: the reference's prebuilt Hessian scatter
  (core/petsc/reordered_element_base.py:1464-1534) and its ``np.add.at`` segmented sum
  (:1753-1757) into a sorted owned COO pattern.
* ``assemble``: energy ``sum_e E_e - f.u`` (:1579-1611), gradient (:1613-1633) and the CSR
  Hessian (:1733-1775) for one problem dictionary.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import elements


def local_state(params: dict, u_free: np.ndarray, dirichlet_key: str = "u_0") -> np.ndarray:
    v = np.asarray(params[dirichlet_key], dtype=np.float64).ravel().copy()
    v[np.asarray(params["freedofs"], dtype=np.int64).ravel()] = np.asarray(u_free, dtype=np.float64)
    return v


def free_index(params: dict, n_total: int) -> np.ndarray:
    fd = np.asarray(params["freedofs"], dtype=np.int64).ravel()
    idx = np.full(n_total, -1, dtype=np.int64)
    idx[fd] = np.arange(fd.size)
    return idx


def scatter_plan(elem_free: np.ndarray, n_free: int):
    """(e, i, j, positions) with positions into the sorted unique (row, col) key list."""
    rows = elem_free[:, :, None]
    cols = elem_free[:, None, :]
    valid = (rows >= 0) & (cols >= 0)
    e, i, j = np.nonzero(valid)
    keys = elem_free[e, i].astype(np.int64) * np.int64(n_free) + elem_free[e, j]
    sorted_keys = np.unique(keys)
    positions = np.searchsorted(sorted_keys, keys)
    return e, i, j, positions, sorted_keys


def scatter_add(elem_h: np.ndarray, plan) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """np.add.at segmented sum -> (indptr, indices, data) of the CSR over the key list."""
    e, i, j, positions, sorted_keys = plan
    n_keys = sorted_keys.size
    vals = np.zeros(n_keys, dtype=np.float64)
    np.add.at(vals, positions, elem_h[e, i, j])
    return sorted_keys, vals


def element_eval(kind: str, params: dict, v_full: np.ndarray, want=("e", "g", "h"), **kw):
    """Per-element quantities for problem ``kind`` at the full nodal vector ``v_full``."""
    elems = np.asarray(params["elems"], dtype=np.int64)
    ve = v_full[elems]
    if kind == "plaplace":
        return elements.plaplace(ve, params["dvx"], params["dvy"], params["vol"], float(params["p"]), want)
    if kind == "gl":
        return elements.ginzburg_landau(
            ve, params["dvx"], params["dvy"], params["vol"], np.asarray(params["ip"]), np.asarray(params["w"]),
            float(params["eps"]), want,
        )
    if kind == "he":
        return elements.neohookean(
            ve, params["dphix"], params["dphiy"], params["dphiz"], params["vol"], float(params["C1"]),
            float(params["D1"]), bool(kw.get("use_abs_det", False)), want,
        )
    if kind == "mc":
        return elements.mohr_coulomb(
            ve, params["elem_B"], params["quad_weight"], params["eps_p_old"], float(params["E"]),
            float(params["nu"]), float(params["phi_deg"]), float(params["cohesion"]),
            float(params.get("reg", 1e-12)), bool(kw.get("elastic_only", False)), want,
        )
    raise ValueError(kind)


def load_vector(kind: str, params: dict) -> np.ndarray | None:
    """Nodal load on the free DOFs (original order) or None."""
    fd = np.asarray(params["freedofs"], dtype=np.int64).ravel()
    if kind == "plaplace":
        return np.asarray(params["f"], dtype=np.float64)[fd]
    if kind == "mc":
        return np.asarray(params["force"], dtype=np.float64)[fd]
    return None


def assemble(kind: str, params: dict, u_free: np.ndarray, dirichlet_key: str = "u_0", want=("e", "g", "h"), **kw):
    """Energy (float), gradient (n_free,) and Hessian (scipy CSR) in original free numbering."""
    v = local_state(params, u_free, dirichlet_key)
    elems = np.asarray(params["elems"], dtype=np.int64)
    n_free = int(np.asarray(params["freedofs"]).size)
    fi = free_index(params, v.size)
    ef = fi[elems]
    q = element_eval(kind, params, v, want, **kw)
    f = load_vector(kind, params)
    out = {}
    if "e" in want:
        E = float(np.sum(q["e"]))
        if f is not None:
            E -= float(np.dot(f, u_free))
        out["e"] = E
    if "g" in want:
        g = np.zeros(n_free, dtype=np.float64)
        m = ef >= 0
        np.add.at(g, ef[m], q["g"][m])
        if f is not None:
            g -= f
        out["g"] = g
    if "h" in want:
        keys, vals = scatter_add(q["h"], scatter_plan(ef, n_free))
        H = sp.csr_matrix((vals, (keys // n_free, keys % n_free)), shape=(n_free, n_free))
        H.sort_indices()
        out["h"] = H
    return out


def dense_hessian_bruteforce(kind: str, params: dict, u_free: np.ndarray, dirichlet_key: str = "u_0", **kw):
    """Dense double loop over elements and local pairs (tiny meshes only)."""
    v = local_state(params, u_free, dirichlet_key)
    elems = np.asarray(params["elems"], dtype=np.int64)
    n_free = int(np.asarray(params["freedofs"]).size)
    fi = free_index(params, v.size)
    q = element_eval(kind, params, v, ("h",), **kw)
    H = np.zeros((n_free, n_free))
    for e in range(elems.shape[0]):
        for a in range(elems.shape[1]):
            ra = fi[elems[e, a]]
            if ra < 0:
                continue
            for b in range(elems.shape[1]):
                cb = fi[elems[e, b]]
                if cb >= 0:
                    H[ra, cb] += q["h"][e, a, b]
    return H
