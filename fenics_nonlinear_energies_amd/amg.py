"""Smoothed-aggregation algebraic multigrid preconditioner on the GPU.

Stands in for the reference's PETSc GAMG / hypre preconditioners with near-nullspace
(``use_near_nullspace``, ``hyperelasticity/jax_petsc/solver.py:140-217``;
``slope_stability/jax_petsc/solver.py:44-63``), which have no GPU counterpart here.

Set-up (host, numpy/scipy, once per operator; level operators and transfers are stored in
the operator's precision, FP32 in the mixed-precision mode): aggregates are consecutive runs of
``agg_dofs`` unknowns in the assembler's reordered numbering -- the Morton order makes them
compact spatial blobs, so no graph aggregation pass is needed; the tentative prolongator
is the per-aggregate QR of the near-nullspace vectors (rigid-body modes for elasticity,
constants for scalar problems); it is smoothed once with damped Jacobi,
P = (I - omega D^-1 A) P_tent with omega = 4 / (3 rho(D^-1 A)); coarse operators are
Galerkin products P^T A P, recursively, until the coarse problem is small enough for a dense
Cholesky factorisation.

Apply (device): one symmetric V-cycle -- two damped-Jacobi sweeps before and after the
coarse correction -- with every operator product, the coarsest dense inverse included, on
the CSR SpMV kernel (``fne_spmv``).  The cycle is recorded once per set-up as a HIP graph
and replayed.  It is a fixed symmetric positive-definite linear operator, so it
preconditions CG.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import torch

from . import native


class _DevMat:
    """CSR operator on the device (rows may differ from columns)."""

    def __init__(self, A: sp.csr_matrix, device, precision: str):
        A = A.tocsr()
        A.sort_indices()
        self.n_rows = int(A.shape[0])
        self.indptr = torch.as_tensor(A.indptr.astype(np.int32)).to(device)
        self.indices = torch.as_tensor(A.indices.astype(np.int32)).to(device)
        dt = torch.float32 if precision == "fp32" else torch.float64
        self.data = torch.as_tensor(A.data.astype(np.float64)).to(device=device, dtype=dt)
        self.prec = native.FNE_FP32 if precision == "fp32" else native.FNE_FP64
        avg = A.nnz / max(1, self.n_rows)
        lpr = 1
        while lpr < 64 and 3 * lpr < avg:
            lpr *= 2
        self.lpr = lpr

    def mult(self, x: torch.Tensor, y: torch.Tensor) -> None:
        native.check(
            native.lib().fne_spmv(self.n_rows, native.ptr(self.indptr), native.ptr(self.indices),
                                  native.ptr(self.data), self.prec | (self.lpr << 8), native.ptr(x), native.ptr(y),
                                  native.stream_ptr()),
            "fne_spmv",
        )


def _spectral_radius_dinv_a(A: sp.csr_matrix, dinv: np.ndarray, its: int = 15, seed: int = 0) -> float:
    x = np.random.default_rng(seed).standard_normal(A.shape[0])
    lam = 1.0
    for _ in range(its):
        y = dinv * (A @ x)
        lam = float(np.linalg.norm(y) / max(np.linalg.norm(x), 1e-300))
        x = y / max(np.linalg.norm(y), 1e-300)
    return lam * 1.05


def tentative_prolongator(B: np.ndarray, agg_dofs: int):
    """Block-diagonal P_tent (n x n_agg k) with orthonormal blocks and the coarse near-nullspace."""
    n, k = B.shape
    n_agg = -(-n // agg_dofs)
    rows, cols, vals = [], [], []
    Bc = np.zeros((n_agg * k, k))
    for a in range(n_agg):
        lo, hi = a * agg_dofs, min(n, (a + 1) * agg_dofs)
        Q, R = np.linalg.qr(B[lo:hi])
        kk = Q.shape[1]
        ii = np.arange(lo, hi)
        for m in range(kk):
            rows.append(ii)
            cols.append(np.full(ii.size, a * k + m))
            vals.append(Q[:, m])
        Bc[a * k:a * k + kk] = R
    P = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n_agg * k))
    # drop coarse columns of rank-deficient aggregates (all-zero columns)
    keep = np.flatnonzero(np.asarray(abs(P).sum(axis=0)).ravel() > 0)
    return P[:, keep].tocsr(), Bc[keep]


class SmoothedAggregationPC:
    """``pc_type="gamg"``: smoothed-aggregation AMG V-cycle on the device."""

    def __init__(self, near_nullspace: np.ndarray | None = None, *, agg_dofs: int | None = None, max_coarse: int = 1200,
                 max_levels: int = 8, sweeps: int = 2, precision: str = "fp64", use_graph: bool = True):
        self.B = None if near_nullspace is None else np.asarray(near_nullspace, dtype=np.float64)
        self.use_graph = bool(use_graph)
        self._graph = None
        self.agg_dofs = agg_dofs
        self.max_coarse = int(max_coarse)
        self.max_levels = int(max_levels)
        self.sweeps = int(sweeps)
        self.precision = precision
        self.levels = []
        self.setup_count = 0

    def setup(self, A_host: sp.csr_matrix, device) -> None:
        A = A_host.tocsr().astype(np.float64)
        n = A.shape[0]
        B = self.B if self.B is not None else np.ones((n, 1))
        k = B.shape[1]
        agg = int(self.agg_dofs or (8 * k if k > 1 else 8))
        levels = []
        cache = getattr(self, "_ptent", None)
        if cache is None:
            cache = self._ptent = []  # tentative prolongators depend on the near-nullspace only
        for lvl in range(self.max_levels):
            d = A.diagonal().copy()
            d[d == 0.0] = 1.0
            dinv = 1.0 / d
            if A.shape[0] <= self.max_coarse:
                break
            if lvl >= len(cache):
                cache.append(tentative_prolongator(B, max(agg, B.shape[1])))
            Pt, Bc = cache[lvl]
            rho = _spectral_radius_dinv_a(A, dinv)
            omega = 4.0 / (3.0 * rho)
            P = (Pt - omega * sp.diags(dinv) @ (A @ Pt)).tocsr()
            Ac = (P.T @ (A @ P)).tocsr()
            Ac = 0.5 * (Ac + Ac.T)
            levels.append({
                "A": _DevMat(A, device, self.precision),
                "P": _DevMat(P, device, self.precision),
                "R": _DevMat(P.T.tocsr(), device, self.precision),
                "dinv": torch.as_tensor(dinv).to(device),
                "omega": omega,
                "n": A.shape[0],
            })
            if Ac.shape[0] >= A.shape[0]:
                A = Ac
                break
            A, B = Ac, Bc
        # coarsest level: dense Cholesky (shifted by a tiny multiple of the trace if singular)
        Ad = A.toarray()
        Ad = 0.5 * (Ad + Ad.T)
        try:
            L = sla.cholesky(Ad, lower=True)
        except np.linalg.LinAlgError:
            shift = 1e-12 * np.trace(Ad) / max(1, Ad.shape[0])
            w, V = np.linalg.eigh(Ad)
            w = np.where(w > shift, w, np.inf)
            Linv = (V / np.sqrt(w)).T  # pseudo-inverse factor
            self.coarse = torch.as_tensor(np.ascontiguousarray(Linv)).to(device)
        else:
            Linv = sla.solve_triangular(L, np.eye(L.shape[0]), lower=True)
            self.coarse = torch.as_tensor(np.ascontiguousarray(Linv)).to(device)
        # the coarse inverse Linv^T Linv as a dense CSR operator: one SpMV kernel, no BLAS call
        Ainv = Linv.T @ Linv
        self.coarse_op = _DevMat(sp.csr_matrix(Ainv), device, "fp64")
        self.levels = levels
        self.setup_count += 1
        self._graph = None
        if self.use_graph and device is not None and torch.device(device).type == "cuda":
            self._capture(n, device)

    def _capture(self, n: int, device) -> None:
        """Record one V-cycle as a HIP graph (static input/output buffers): the cycle is
        launch-bound on the coarse levels, and a replay issues all its kernels at once."""
        self._g_in = torch.zeros(n, dtype=torch.float64, device=device)
        self._g_out = torch.zeros(n, dtype=torch.float64, device=device)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(2):
                self._g_out.copy_(self._cycle(0, self._g_in))
        torch.cuda.current_stream().wait_stream(side)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._g_out.copy_(self._cycle(0, self._g_in))
        self._graph = g

    def _coarse_solve(self, r: torch.Tensor) -> torch.Tensor:
        y = torch.empty(self.coarse_op.n_rows, dtype=r.dtype, device=r.device)
        self.coarse_op.mult(r, y)
        return y

    def _cycle(self, lvl: int, r: torch.Tensor) -> torch.Tensor:
        if lvl == len(self.levels):
            return self._coarse_solve(r)
        L = self.levels[lvl]
        A, dinv, om = L["A"], L["dinv"], L["omega"]
        x = om * dinv * r
        t = torch.empty_like(r)
        for _ in range(self.sweeps - 1):
            A.mult(x, t)
            x.add_(om * dinv * (r - t))
        A.mult(x, t)
        res = r - t
        rc = torch.empty(L["R"].n_rows, dtype=r.dtype, device=r.device)
        L["R"].mult(res, rc)
        xc = self._cycle(lvl + 1, rc)
        L["P"].mult(xc, t)
        x.add_(t)
        for _ in range(self.sweeps):
            A.mult(x, t)
            x.add_(om * dinv * (r - t))
        return x

    def apply(self, r: torch.Tensor, z: torch.Tensor) -> None:
        if self._graph is not None:
            self._g_in.copy_(r)
            self._graph.replay()
            z.copy_(self._g_out)
        else:
            z.copy_(self._cycle(0, r))

    def summary(self) -> dict:
        return {"levels": len(self.levels) + 1, "sizes": [L["n"] for L in self.levels] + [int(self.coarse.shape[0])]}
