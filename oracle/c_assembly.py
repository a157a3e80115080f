"""ctypes front end of the C oracle (``fne_oracle.c``) -- TEST INFRASTRUCTURE ONLY.

``P1Assembly(kind, params)`` precomputes, on the host and untimed, the CSR pattern of the free
DOFs (the sorted unique (row, col) keys of ``assembly.scatter_plan``), each element entry's CSR
position and each free row's incidence list; ``__call__(u_free)`` runs the OpenMP assembly and
returns (energy, gradient, scipy CSR Hessian) in the original free numbering, like
``assembly.assemble``.  Used by tests/test_oracle_c.py (against ``assembly.assemble``) and as
the all-cores CPU baseline of bench.py.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import scipy.sparse as sp

from .assembly import free_index, load_vector

_LIB = Path(__file__).resolve().parent / "_lib" / "libfne_oracle.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            raise FileNotFoundError(f"{_LIB} is not built (run `make -C oracle` or __graft_entry__.build())")
        _lib = C.CDLL(str(_LIB))
        p = C.c_void_p
        _lib.fne_oracle_threads.restype = C.c_int
        _lib.fne_oracle_p1_assemble.restype = C.c_int
        _lib.fne_oracle_p1_assemble.argtypes = [C.c_int, C.c_double, C.c_int, p, p, p, p, p, p, p, C.c_int, C.c_int,
                                                p, p, p, p, p, C.c_int64, p, p, p, p]
    return _lib


def threads() -> int:
    return int(lib().fne_oracle_threads())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class P1Assembly:
    """Fused energy + gradient + CSR Hessian of a P1 scalar problem ("plaplace" or "gl")."""

    def __init__(self, kind: str, params: dict, dirichlet_key: str = "u_0"):
        if kind not in ("plaplace", "gl"):
            raise ValueError(kind)
        self.kind = kind
        self.params = params
        self.dirichlet_key = dirichlet_key
        elems = np.asarray(params["elems"], dtype=np.int64)
        n_total = int(np.asarray(params[dirichlet_key]).size)
        fd = np.asarray(params["freedofs"], dtype=np.int64).ravel()
        n_free = fd.size
        ef = free_index(params, n_total)[elems]  # (N, 3), -1 = constrained
        rows, cols = ef[:, :, None], ef[:, None, :]
        keys = np.where((rows >= 0) & (cols >= 0), rows * np.int64(n_free) + cols, -1).reshape(-1, 9)
        sorted_keys = np.unique(keys[keys >= 0])
        pos = np.where(keys >= 0, np.searchsorted(sorted_keys, np.maximum(keys, 0)), -1)
        self.indptr = np.concatenate(([0], np.cumsum(np.bincount(sorted_keys // n_free, minlength=n_free))))
        self.indices = (sorted_keys % n_free).astype(np.int64)
        self.nnz = int(sorted_keys.size)
        e_idx = np.repeat(np.arange(elems.shape[0], dtype=np.int64), 3)
        a_idx = np.tile(np.arange(3, dtype=np.int64), elems.shape[0])
        r = ef.ravel()
        ok = r >= 0
        order = np.lexsort((3 * e_idx[ok] + a_idx[ok], r[ok]))
        self.inc = (3 * e_idx[ok] + a_idx[ok])[order].astype(np.int64)
        self.inc_ptr = np.concatenate(([0], np.cumsum(np.bincount(r[ok], minlength=n_free)))).astype(np.int64)
        self.pos = np.ascontiguousarray(pos, dtype=np.int64)
        self.elems = np.ascontiguousarray(elems, dtype=np.int32)
        self.dvx = np.ascontiguousarray(params["dvx"], dtype=np.float64)
        self.dvy = np.ascontiguousarray(params["dvy"], dtype=np.float64)
        self.vol = np.ascontiguousarray(params["vol"], dtype=np.float64)
        self.fd = fd
        self.free_nodes = fd.astype(np.int32)
        f = load_vector(kind, params)
        self.f = None if f is None else np.ascontiguousarray(f, dtype=np.float64)
        if kind == "gl":
            self.family, self.param = 2, float(params["eps"])
            self.ip = np.ascontiguousarray(params["ip"], dtype=np.float64)
            self.w = np.ascontiguousarray(params["w"], dtype=np.float64)
            self.nq = int(self.w.size)
        else:
            self.family, self.param = 1, float(params["p"])
            self.ip = self.w = None
            self.nq = 0
        self.n_free = n_free
        self.u_full = np.asarray(params[dirichlet_key], dtype=np.float64).ravel().copy()
        self.scratch = np.empty(13 * elems.shape[0], dtype=np.float64)
        self.grad = np.empty(n_free, dtype=np.float64)
        self.vals = np.empty(self.nnz, dtype=np.float64)

    def run(self, u_free: np.ndarray) -> float:
        """Assemble into self.grad / self.vals; returns the energy."""
        self.u_full[self.fd] = np.asarray(u_free, dtype=np.float64)
        e = C.c_double(0.0)
        rc = lib().fne_oracle_p1_assemble(
            self.family, self.param, int(self.elems.shape[0]), _ptr(self.elems), _ptr(self.dvx), _ptr(self.dvy),
            _ptr(self.vol), _ptr(self.u_full), _ptr(self.ip), _ptr(self.w), self.nq, self.n_free,
            _ptr(self.free_nodes), _ptr(self.f), _ptr(self.inc_ptr), _ptr(self.inc), _ptr(self.pos), self.nnz,
            _ptr(self.scratch), C.byref(e), _ptr(self.grad), _ptr(self.vals))
        if rc != 0:
            raise RuntimeError(f"fne_oracle_p1_assemble failed ({rc})")
        return float(e.value)

    def __call__(self, u_free: np.ndarray):
        E = self.run(u_free)
        H = sp.csr_matrix((self.vals.copy(), self.indices, self.indptr), shape=(self.n_free, self.n_free))
        return E, self.grad.copy(), H
