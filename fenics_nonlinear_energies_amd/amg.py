"""Smoothed-aggregation algebraic multigrid preconditioner on the GPU.

Stands in for the reference's PETSc GAMG / hypre preconditioners with near-nullspace
(``use_near_nullspace``, ``hyperelasticity/jax_petsc/solver.py:140-217``;
``slope_stability/jax_petsc/solver.py:44-63``), which have no GPU counterpart here.  Like
PETSc's, the hierarchy is rebuilt from the current Hessian at every set-up.

Hierarchy: aggregates are consecutive runs of ``agg_dofs`` unknowns in the assembler's
reordered numbering -- the (isotropic) Morton order makes them compact spatial blobs, so no
graph aggregation pass is needed.  By default a fine-level aggregate holds 16 nodes' worth of
unknowns (16 k for k near-nullspace vectors) and a coarse-level one 8 coarse nodes (8 k): on
the Neo-Hookean beam this keeps the operator complexity near 1.4 at about 9 CG iterations
per decade-1 solve.  The tentative prolongator is the per-aggregate QR of the
near-nullspace vectors (rigid-body modes for elasticity, constants for scalar problems);
it is smoothed once with damped Jacobi, P = (I - omega D^-1 A) P_tent with
omega = 4 / (3 rho(D^-1 A)); coarse operators are Galerkin products P^T A P, recursively,
until the coarse problem is small enough for a dense inverse.

Set-up in two phases.  Symbolic (host, scipy, once per Hessian pattern, done by the
assembler at construction): the patterns of P, A P, P^T and P^T A P on every level, the
transpose permutations and the diagonal positions.  Numeric (device, every set-up): the
per-aggregate QR of the near-nullspace (``fne_aggregate_qr``; the near-nullspace may follow
the current state), the spectral radius by a short Lanczos run, and the three products per level
on the fixed-pattern product kernel (``fne_spgemm_numeric``); only the coarsest operator goes
to the host for its dense inverse.
Level operators and transfers are kept in FP64 for the products and copied to the
operator's precision (FP32 in the mixed-precision mode) for the cycle.

Apply (device): one symmetric V-cycle -- two smoothing steps before and after the coarse
correction, Chebyshev with Jacobi scaling on [0.1, 1.1] rho(D^-1 A) by default (PETSc's
KSPCHEBYSHEV bounds, the default PCGAMG smoother; ``fne_vcycle_cheb``) or damped Jacobi
(``fne_vcycle_op``) -- with every operator product, the coarsest dense inverse included, on
the CSR SpMV kernels; every level, the finest included, uses the operator copied
at set-up, so a kept preconditioner stays the same linear operator while the Hessian changes.
The cycle is recorded as a HIP graph after the first set-up and replayed; later set-ups
rewrite its buffers in place.  It is a fixed symmetric positive-definite linear operator between
set-ups, so it preconditions CG.
"""

from __future__ import annotations

import time

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp
import torch

from . import native


def _spmv_code(prec: int, avg: float) -> int:
    from .linalg import spmv_code

    return spmv_code(prec, avg)


# fne_vcycle_op operations
_SWEEP_FIRST, _SWEEP, _RESIDUAL, _ADD_PRODUCT = 0, 1, 2, 3


class _DevMat:
    """CSR operator on the device (rows may differ from columns); ``data`` is rewritten in place."""

    def __init__(self, indptr: np.ndarray, indices: np.ndarray, n_cols: int, device, precision: str):
        self.n_rows = int(np.asarray(indptr).size - 1)
        self.n_cols = int(n_cols)
        self.nnz = int(np.asarray(indices).size)
        self.indptr = torch.as_tensor(np.asarray(indptr, dtype=np.int32)).to(device)
        self.indices = torch.as_tensor(np.asarray(indices, dtype=np.int32)).to(device)
        dt = torch.float32 if precision == "fp32" else torch.float64
        self.data = torch.zeros(self.nnz, dtype=dt, device=device)
        self.prec = native.FNE_FP32 if precision == "fp32" else native.FNE_FP64
        self.code = _spmv_code(self.prec, self.nnz / max(1, self.n_rows))

    @classmethod
    def sharing_pattern(cls, indptr: torch.Tensor, indices: torch.Tensor, n_cols: int, precision: str) -> "_DevMat":
        """Operator with its own values on another matrix's (device) pattern."""
        m = cls.__new__(cls)
        m.n_rows, m.n_cols, m.nnz = int(indptr.numel() - 1), int(n_cols), int(indices.numel())
        m.indptr, m.indices = indptr, indices
        dt = torch.float32 if precision == "fp32" else torch.float64
        m.data = torch.zeros(m.nnz, dtype=dt, device=indices.device)
        m.prec = native.FNE_FP32 if precision == "fp32" else native.FNE_FP64
        m.code = _spmv_code(m.prec, m.nnz / max(1, m.n_rows))
        return m

    def op(self, op: int, x, r, w, y_in, y: torch.Tensor) -> None:
        """Fused V-cycle operator (fne_vcycle_op): sweep / residual / prolongation-add."""
        native.check(
            native.lib().fne_vcycle_op(op, self.n_rows, native.ptr(self.indptr), native.ptr(self.indices),
                                       native.ptr(self.data), self.code, native.ptr(x),
                                       native.ptr(r), native.ptr(w), native.ptr(y_in), native.ptr(y),
                                       native.stream_ptr()),
            "fne_vcycle_op",
        )

    def cheb(self, x, x_prev, r, w, c: float, y: torch.Tensor) -> None:
        """Chebyshev smoothing step y = x + c (x - x_prev) + w.(r - A x) (fne_vcycle_cheb)."""
        native.check(
            native.lib().fne_vcycle_cheb(self.n_rows, native.ptr(self.indptr), native.ptr(self.indices),
                                         native.ptr(self.data), self.code, native.ptr(x),
                                         native.ptr(x_prev), native.ptr(r), native.ptr(w), float(c), native.ptr(y),
                                         native.stream_ptr()),
            "fne_vcycle_cheb",
        )

    def mult(self, x: torch.Tensor, y: torch.Tensor) -> None:
        native.check(
            native.lib().fne_spmv(self.n_rows, native.ptr(self.indptr), native.ptr(self.indices),
                                  native.ptr(self.data), self.code, native.ptr(x), native.ptr(y),
                                  native.stream_ptr()),
            "fne_spmv",
        )


def _spgemm(a_indptr, a_indices, a_data, b: _DevMat, b_data: torch.Tensor, c: _DevMat, c_data: torch.Tensor):
    """c_data <- (A B) on C's pattern (fne_spgemm_numeric); A given by its CSR arrays."""
    n_rows = int(a_indptr.numel() - 1)
    prec = native.FNE_FP32 if a_data.dtype == torch.float32 else native.FNE_FP64
    native.check(
        native.lib().fne_spgemm_numeric(n_rows, native.ptr(a_indptr), native.ptr(a_indices), native.ptr(a_data), prec,
                                        native.ptr(b.indptr), native.ptr(b.indices), native.ptr(b_data),
                                        native.ptr(c.indptr), native.ptr(c.indices), native.ptr(c_data),
                                        native.stream_ptr()),
        "fne_spgemm_numeric",
    )


def _pattern(M: sp.csr_matrix) -> sp.csr_matrix:
    """Structure of M with all-ones values (sorted columns)."""
    M = sp.csr_matrix(M)
    M.sort_indices()
    return sp.csr_matrix((np.ones(M.nnz), M.indices.copy(), M.indptr.copy()), shape=M.shape)


def _pattern_product(A: sp.csr_matrix, B: sp.csr_matrix) -> sp.csr_matrix:
    # all-positive values: no entry of the product cancels, so this is the full structure
    return _pattern(_pattern(A) @ _pattern(B))


def _transpose_perm(M: sp.csr_matrix) -> np.ndarray:
    """perm with (M^T).data == M.data[perm] for M^T in sorted CSR."""
    T = sp.csr_matrix((np.arange(M.nnz, dtype=np.float64), M.indices, M.indptr), shape=M.shape).T.tocsr()
    T.sort_indices()
    return T.data.astype(np.int64)


def _positions(M: sp.csr_matrix, rows: np.ndarray, cols: np.ndarray) -> np.ndarray:
    """Index into M.data of each (row, col) entry (all must be in M's sorted pattern)."""
    key_m = np.repeat(np.arange(M.shape[0], dtype=np.int64), np.diff(M.indptr)) * M.shape[1] + M.indices
    key = np.asarray(rows, dtype=np.int64) * M.shape[1] + np.asarray(cols, dtype=np.int64)
    pos = np.searchsorted(key_m, key)
    if key.size and (np.any(pos >= key_m.size) or np.any(key_m[np.minimum(pos, key_m.size - 1)] != key)):
        raise ValueError("entry outside the pattern")
    return pos


def coarse_inverse(Ad: np.ndarray) -> tuple[np.ndarray, bool]:
    """Dense inverse of the (symmetrised) coarsest operator for the V-cycle's coarse solve.

    Cholesky (LAPACK potrf/potri) when the operator is positive definite; otherwise the
    pseudo-inverse over its eigenvalues above 1e-12 max|eigenvalue| (a positive cutoff whatever
    the sign of the trace), which keeps the coarse correction positive semi-definite when the
    Hessian is indefinite.  Returns (inverse, whether the fallback was taken).
    """
    nc = Ad.shape[0]
    c, info = sla.lapack.dpotrf(Ad, lower=1)
    if info == 0:
        Ainv, info = sla.lapack.dpotri(c, lower=1)  # lower triangle of the inverse
    if info == 0:
        return np.tril(Ainv) + np.tril(Ainv, -1).T, False
    # the same LAPACK as the Cholesky above: mixing it with torch's CPU thread pool right after a
    # failed factorisation oversubscribed the host and cost ~0.1 s per fallback
    w, V = sla.eigh(np.ascontiguousarray(Ad, dtype=np.float64))
    cut = 1e-12 * float(np.abs(w).max()) if w.size else 0.0
    w = np.where(w > cut, w, np.inf)
    return (V / w) @ V.T, True


def coarse_inverse_device(S: torch.Tensor) -> tuple[torch.Tensor, bool]:
    """``coarse_inverse`` in FP64 on the tensor's device: Cholesky and its inverse
    (``torch.linalg.cholesky_ex`` / ``torch.cholesky_inverse``) when the symmetric operator ``S``
    is positive definite, else the pseudo-inverse over its eigenvalues above 1e-12 of the largest
    |eigenvalue| (``torch.linalg.eigh``).  One host read (the factorisation's status) per call.
    Returns (inverse, whether the fallback was taken)."""
    L, info = torch.linalg.cholesky_ex(S)
    if int(info.item()) == 0:
        return torch.cholesky_inverse(L), False
    w, V = torch.linalg.eigh(S)
    cut = 1e-12 * w.abs().max()
    keep = w > cut
    winv = torch.where(keep, 1.0 / torch.where(keep, w, torch.ones_like(w)), torch.zeros_like(w))
    return (V * winv) @ V.T, True


def _injection_rows(P: sp.csr_matrix) -> np.ndarray:
    """For each column j of an interpolation P, a row whose only entry is P[i, j] = 1 (the fine
    DOF that coincides with coarse DOF j); raises if a column has none."""
    P = sp.csr_matrix(P)
    lens = np.diff(P.indptr)
    unit = np.flatnonzero(lens == 1)
    unit = unit[np.abs(P.data[P.indptr[unit]] - 1.0) <= 1e-12]
    cols = P.indices[P.indptr[unit]]
    first = np.full(P.shape[1], -1, dtype=np.int64)
    order = np.argsort(cols, kind="stable")
    c_sorted, u_sorted = cols[order], unit[order]
    keep = np.concatenate(([True], c_sorted[1:] != c_sorted[:-1])) if c_sorted.size else np.zeros(0, bool)
    first[c_sorted[keep]] = u_sorted[keep]
    if np.any(first < 0):
        raise ValueError("a coarse DOF coincides with no fine DOF; pass coarse_near_nullspace")
    return first


def _aggregate_sizes(n: int, agg: int) -> np.ndarray:
    na, rem = divmod(n, agg)
    return np.array([agg] * na + ([rem] if rem else []), dtype=np.int64)


def tentative_prolongator(B: np.ndarray, agg_dofs: int):
    """Block-diagonal P_tent (n x sum_a min(rows_a, k), CSR) and the coarse near-nullspace.

    Aggregate a holds rows [a*agg, (a+1)*agg); its block is the Q factor of those rows of B
    (batched numpy QR, signs chosen so that R has a non-negative diagonal) and its coarse
    near-nullspace rows are the R factor.  Host twin of ``fne_aggregate_qr``.
    """
    n, k = B.shape
    sizes = _aggregate_sizes(n, agg_dofs)
    na_full = n // agg_dofs
    vals, Bc = [], []
    def signed(Q, R):  # R with a non-negative diagonal (the convention of fne_aggregate_qr)
        kk = Q.shape[-1]
        d = np.diagonal(R[..., :kk], axis1=-2, axis2=-1)
        sg = np.where(d < 0.0, -1.0, 1.0)
        return Q * sg[..., None, :], R * sg[..., :, None]

    if na_full:
        Q, R = signed(*np.linalg.qr(B[: na_full * agg_dofs].reshape(na_full, agg_dofs, k)))
        vals.append(Q.reshape(-1))
        Bc.append(R.reshape(-1, k))
    if sizes.size > na_full:
        Q, R = signed(*np.linalg.qr(B[na_full * agg_dofs:]))
        vals.append(Q.reshape(-1))
        Bc.append(R)
    kk = np.minimum(sizes, k)
    col0 = np.concatenate(([0], np.cumsum(kk)[:-1]))
    row_k = np.repeat(kk, sizes)
    indptr = np.concatenate(([0], np.cumsum(row_k))).astype(np.int64)
    start = np.repeat(np.repeat(col0, sizes), row_k)  # first coarse column of each entry's aggregate
    indices = start + np.arange(indptr[-1]) - np.repeat(indptr[:-1], row_k)
    Pt = sp.csr_matrix((np.concatenate(vals), indices, indptr), shape=(n, int(kk.sum())))
    return Pt, np.concatenate(Bc, axis=0)


class SmoothedAggregationPC:
    """``pc_type="gamg"``: smoothed-aggregation AMG V-cycle on the device."""

    LANCZOS_STEPS = 12  # Lanczos steps for rho(D^-1 A) at every set-up
    RHO_SAFETY = 1.1  # the Ritz value from below is raised by this factor

    def __init__(self, near_nullspace: np.ndarray | None = None, *, agg_dofs: int | tuple | None = None,
                 max_coarse: int = 1200, max_levels: int = 8, sweeps: int = 2, precision: str = "fp64",
                 use_graph: bool = True, fixed_prolongators: list | None = None,
                 coarse_near_nullspace: np.ndarray | None = None, smoother: str = "chebyshev",
                 coarse_factor: str = "host"):
        """``fixed_prolongators``: host CSR transfers (n_l x n_{l+1}) of the finest levels, used as
        given (no aggregation, no smoothing; e.g. the same-mesh P4 -> P2 -> P1 interpolations of
        ``pmg.py``); aggregation continues below the last of them with ``coarse_near_nullspace``
        (rows of the coarsest fixed level; default: the near-nullspace injected through the
        transfers' unit rows).  ``coarse_factor``: where the coarsest operator's dense FP64 inverse
        is computed at set-up, "host" (LAPACK) or "device" (the GPU's dense solvers, no host
        round trip); the cycle applies it on the device either way.  The two agree to rounding;
        the inexact Newton paths built on them can differ by an iteration or two."""
        if smoother not in ("jacobi", "chebyshev"):
            raise ValueError(f"unsupported multigrid smoother {smoother!r} (jacobi | chebyshev)")
        if coarse_factor not in ("host", "device"):
            raise ValueError(f"coarse_factor must be 'host' or 'device', not {coarse_factor!r}")
        self.coarse_factor = coarse_factor
        # level smoother: damped Jacobi (omega = 4 / (3 rho)), or Chebyshev on [0.1, 1.1] rho with
        # Jacobi scaling (PETSc's KSPCHEBYSHEV eigenvalue bounds, the default PCGAMG smoother)
        self.smoother = smoother
        self.fixed = [sp.csr_matrix(P) for P in (fixed_prolongators or [])]
        for P in self.fixed:
            P.sort_indices()
        self.B_coarse_fixed = None if coarse_near_nullspace is None else np.asarray(coarse_near_nullspace,
                                                                                  dtype=np.float64)
        self.B = None if near_nullspace is None else np.asarray(near_nullspace, dtype=np.float64)
        self.use_graph = bool(use_graph)
        self._graph = None
        self.agg_dofs = agg_dofs
        self.max_coarse = int(max_coarse)
        self.max_levels = int(max_levels)
        self.sweeps = int(sweeps)
        self.precision = precision
        self.levels: list[dict] = []
        self.setup_count = 0
        self.setup_time = 0.0
        self.coarse_eig_fallbacks = 0  # set-ups whose coarsest operator was not positive definite
        self.symbolic_time = 0.0
        self._key = None
        self._A0 = None

    # ---------------------------------------------------------------------------------------
    def _aggs(self, k: int) -> tuple:
        if self.agg_dofs is None:
            return (16 * k, 8 * k) if k > 1 else (8,)
        return tuple(int(a) for a in np.atleast_1d(self.agg_dofs))

    def _symbolic(self, A0, device, k: int) -> None:
        """Patterns of every level's products (host, once per Hessian pattern)."""
        t0 = time.perf_counter()
        n = A0.n
        a_indices = A0.indices.cpu().numpy()
        A_pat = _pattern(sp.csr_matrix((np.ones(A0.nnz), a_indices, A0.indptr.cpu().numpy()), shape=(n, n)))
        if not np.array_equal(A_pat.indices, a_indices):
            raise ValueError("the Hessian's column indices must be sorted within each row")
        aggs = self._aggs(k)
        prec = self.precision
        levels = []
        nfix = len(self.fixed)
        for lvl in range(self.max_levels + nfix):
            nl = A_pat.shape[0]
            fixed = lvl < nfix
            if fixed:
                Pf = self.fixed[lvl]
                if Pf.shape[0] != nl:
                    raise ValueError(f"fixed prolongator {lvl} has {Pf.shape[0]} rows for a level of size {nl}")
                agg, nc = 0, Pf.shape[1]
                Pt = Pf
                P_pat = _pattern(Pf)
            else:
                if nl <= self.max_coarse:
                    break
                agg = max(aggs[min(lvl - nfix, len(aggs) - 1)], k)
                Pt, _ = tentative_prolongator(np.ones((nl, k)), agg)
                nc = Pt.shape[1]
                if nc >= nl:
                    break
                P_pat = _pattern_product(A_pat, Pt)
            AP_pat = _pattern_product(A_pat, P_pat)
            R_pat = _pattern(P_pat.T.tocsr())
            Ac_pat = _pattern_product(R_pat, AP_pat)
            if Ac_pat.shape != (nc, nc):
                raise RuntimeError("coarse pattern has the wrong shape")
            pt_rows = np.repeat(np.arange(nl), np.diff(Pt.indptr))
            diag = _positions(A_pat, np.arange(nl), np.arange(nl))
            levels.append({
                "n": nl,
                "nc": nc,
                "agg": agg,
                "fixed": fixed,
                "Pfix": torch.as_tensor(Pf.data.copy()).to(device) if fixed else None,
                # FP64 masters for the products
                "Pt": _DevMat(Pt.indptr, Pt.indices, nc, device, "fp64"),
                "P64": _DevMat(P_pat.indptr, P_pat.indices, nc, device, "fp64"),
                "AP64": _DevMat(AP_pat.indptr, AP_pat.indices, nc, device, "fp64"),
                "R64": _DevMat(R_pat.indptr, R_pat.indices, nl, device, "fp64"),
                "Ac64": _DevMat(Ac_pat.indptr, Ac_pat.indices, nc, device, "fp64"),
                # transfers of the cycle, in the operator's precision
                "P": _DevMat(P_pat.indptr, P_pat.indices, nc, device, prec),
                "R": _DevMat(R_pat.indptr, R_pat.indices, nl, device, prec),
                "ptpos": torch.as_tensor(_positions(P_pat, pt_rows, Pt.indices)).to(device),
                "p_rows": torch.as_tensor(np.repeat(np.arange(nl), np.diff(P_pat.indptr))).to(device),
                "tperm_p": torch.as_tensor(_transpose_perm(P_pat)).to(device),
                "tperm_ac": torch.as_tensor(_transpose_perm(Ac_pat)).to(device),
                "diag": torch.as_tensor(diag).to(device),
                "wd": torch.zeros(nl, dtype=torch.float64, device=device),
                "wcheb": torch.zeros((max(1, self.sweeps), nl), dtype=torch.float64, device=device),
                "Bc": torch.zeros((nc, k), dtype=torch.float64, device=device),
                # fixed levels: the coarse near-nullspace rows taken from the fine rows that the
                # transfer injects (a unit row: the fine DOF coincides with the coarse one)
                "inject": torch.as_tensor(_injection_rows(Pf)).to(device) if fixed else None,
                "v0": torch.as_tensor(np.random.default_rng(lvl).standard_normal(nl)).to(device),
            })
            A_pat = Ac_pat
        # the cycle's copy of every level's operator, in the operator's precision: the cycle
        # stays the fixed operator of its set-up while the assembler rewrites the Hessian
        for i, L in enumerate(levels):
            ip, ix = (A0.indptr, A0.indices) if i == 0 else (levels[i - 1]["Ac64"].indptr, levels[i - 1]["Ac64"].indices)
            L["A"] = _DevMat.sharing_pattern(ip, ix, L["n"], prec)
        nc = A_pat.shape[0]
        self._coarse_pat = A_pat
        cols = np.tile(np.arange(nc), nc)
        self.coarse_op = _DevMat(np.arange(0, nc * nc + 1, nc), cols, nc, device, "fp64")
        self._coarse_n = nc
        # first use of the dense solvers of the coarse inverse at the coarse size (library
        # loading, thread pools, device solver handles), here in the symbolic phase rather than in
        # whichever numeric set-up first needs them: a definite operator (Cholesky) and an
        # indefinite one (the eigendecomposition fallback)
        warm = np.diag(np.linspace(1.0, 2.0, nc))
        warm_indef = warm.copy()
        warm_indef[0, 0] = -1.0
        if self.coarse_factor == "device" and torch.device(device).type == "cuda":
            rows = np.repeat(np.arange(nc, dtype=np.int64), np.diff(A_pat.indptr))
            self._coarse_flat = torch.as_tensor(rows * nc + A_pat.indices.astype(np.int64)).to(device)
            coarse_inverse_device(torch.as_tensor(warm, device=device))
            coarse_inverse_device(torch.as_tensor(warm_indef, device=device))
        else:
            coarse_inverse(warm)
            coarse_inverse(warm_indef)
        self.levels = levels
        self.symbolic_time += time.perf_counter() - t0

    # ---------------------------------------------------------------------------------------
    def _tentative_values(self, B: np.ndarray) -> list:
        """P_tent values of every level (host QR) from the near-nullspace of the finest."""
        out = []
        for L in self.levels:
            if L["fixed"]:
                out.append(None)
                B = self._fixed_coarse_B(L, B)
                continue
            Pt, Bc = tentative_prolongator(B, L["agg"])
            if Pt.nnz != L["Pt"].nnz:
                raise RuntimeError("near-nullspace changed the aggregate structure")
            out.append(Pt.data)
            B = Bc
        return out

    def _fixed_coarse_B(self, L: dict, B):
        """Near-nullspace on the level below a fixed transfer: the given one below the last
        fixed level, else the rows of the coarse DOFs' coinciding fine DOFs."""
        if self.B_coarse_fixed is not None and L is [M for M in self.levels if M["fixed"]][-1]:
            Bc = self.B_coarse_fixed
            return torch.as_tensor(Bc).to(B.device) if isinstance(B, torch.Tensor) else Bc
        if isinstance(B, torch.Tensor):
            return B[L["inject"]].contiguous()
        return B[L["inject"].cpu().numpy()]

    @staticmethod
    def _key_of(A0, k: int) -> tuple:
        return (id(A0), A0.nnz, A0.indices.data_ptr(), int(k))

    def prepare(self, A0, k: int, device=None) -> None:
        """Symbolic phase only (host patterns for the Hessian pattern of ``A0`` and k
        near-nullspace vectors); ``setup`` then starts with the numeric phase."""
        device = A0.indices.device if device is None else device
        key = self._key_of(A0, k)
        if key != self._key:
            self._graph = None
            self._symbolic(A0, device, k)
            self._key = key

    def setup(self, A0, device=None, near_nullspace: np.ndarray | None = None) -> None:
        """(Re)build the hierarchy from the device Hessian ``A0`` (a ``DeviceCSR``).

        ``near_nullspace`` (n x k host array) replaces the one given at construction for this
        and later set-ups (e.g. rigid-body modes at the current configuration).
        """
        t_setup = time.perf_counter()
        device = A0.indices.device if device is None else device
        if near_nullspace is not None:
            self.B = near_nullspace if isinstance(near_nullspace, torch.Tensor) else np.asarray(
                near_nullspace, dtype=np.float64)
        n = A0.n
        B = self.B if self.B is not None else np.ones((n, 1))
        if B.shape[0] != n:
            raise ValueError(f"near-nullspace has {B.shape[0]} rows for an operator of size {n}")
        k = int(B.shape[1])
        self.prepare(A0, k, device)
        self._A0 = A0
        if k in (1, 2, 3, 6):
            # tentative prolongators on the device: QR of every aggregate, level by level
            if isinstance(B, torch.Tensor):
                Bl = B.to(device=device, dtype=torch.float64).contiguous()
            else:
                if getattr(self, "_B_host", None) is not B:
                    self._B_host, self._B_dev = B, torch.as_tensor(np.ascontiguousarray(B)).to(device)
                Bl = self._B_dev
            for L in self.levels:
                if L["fixed"]:
                    L["Bc"].copy_(self._fixed_coarse_B(L, Bl))
                else:
                    native.check(native.lib().fne_aggregate_qr(L["n"], k, L["agg"], native.ptr(Bl),
                                                               native.ptr(L["Pt"].data), native.ptr(L["Bc"]),
                                                               native.stream_ptr()), "fne_aggregate_qr")
                Bl = L["Bc"]
            pt_vals = None
        else:
            Bh = B.detach().cpu().numpy() if isinstance(B, torch.Tensor) else B
            pt_vals = self._tentative_values(Bh)
        a_ptr, a_idx, a_dat = A0.indptr, A0.indices, A0.data
        if self.levels:
            self.levels[0]["A"].data.copy_(A0.data)
        for lvl, L in enumerate(self.levels):
            d = A0.diagonal() if lvl == 0 else a_dat[L["diag"]]
            d = torch.where(d == 0.0, torch.ones_like(d), d)
            dinv = 1.0 / d
            # rho(D^-1 A): largest Ritz value of a Lanczos run on D^-1/2 A D^-1/2, times a safety
            # factor, so that the smoother's omega rho stays well below 2
            lam = self._lanczos_lambda_max(lvl, A0, dinv, L["v0"])
            omega = 4.0 / (3.0 * self.RHO_SAFETY * lam)
            L["wd"].copy_(omega * dinv)
            if self.smoother == "chebyshev":
                for k, coef in enumerate(self._cheb_weights(lam)):
                    L["wcheb"][k].copy_(coef * dinv)
            Pt, P64 = L["Pt"], L["P64"]
            if L["fixed"]:  # a given transfer, used as it is
                P64.data.copy_(L["Pfix"])
            else:  # P = P_tent - omega D^-1 (A P_tent)
                if pt_vals is not None:
                    Pt.data.copy_(torch.as_tensor(pt_vals[lvl]))
                _spgemm(a_ptr, a_idx, a_dat, Pt, Pt.data, P64, P64.data)
                P64.data.mul_(-omega * dinv[L["p_rows"]])
                P64.data.index_add_(0, L["ptpos"], Pt.data)
            # A P, R = P^T, A_c = R (A P), symmetrised
            AP64, R64, Ac64 = L["AP64"], L["R64"], L["Ac64"]
            _spgemm(a_ptr, a_idx, a_dat, P64, P64.data, AP64, AP64.data)
            torch.index_select(P64.data, 0, L["tperm_p"], out=R64.data)
            _spgemm(R64.indptr, R64.indices, R64.data, AP64, AP64.data, Ac64, Ac64.data)
            Ac64.data.copy_(0.5 * (Ac64.data + Ac64.data[L["tperm_ac"]]))
            L["P"].data.copy_(P64.data)
            L["R"].data.copy_(R64.data)
            if lvl + 1 < len(self.levels):
                self.levels[lvl + 1]["A"].data.copy_(Ac64.data)
            a_ptr, a_idx, a_dat = Ac64.indptr, Ac64.indices, Ac64.data
        # coarsest level: dense FP64 inverse (Cholesky; positive-part pseudo-inverse if the
        # operator is not positive definite), on the host or the device, applied on the device
        nc, cp = self._coarse_n, self._coarse_pat
        if getattr(self, "_coarse_flat", None) is not None:
            Ad = torch.zeros(nc * nc, dtype=torch.float64, device=self.coarse_op.data.device)
            Ad.index_copy_(0, self._coarse_flat, a_dat.double())
            Ad = Ad.view(nc, nc)
            Ainv, fallback = coarse_inverse_device(0.5 * (Ad + Ad.T))
            self.coarse_op.data.copy_(Ainv.reshape(-1))
        else:
            Ad = sp.csr_matrix((a_dat.double().cpu().numpy(), cp.indices, cp.indptr), shape=(nc, nc)).toarray()
            Ainv, fallback = coarse_inverse(0.5 * (Ad + Ad.T))
            self.coarse_op.data.copy_(torch.as_tensor(np.ascontiguousarray(Ainv).ravel()))
        self.coarse_eig_fallbacks += int(fallback)
        self.setup_count += 1
        self.setup_time += time.perf_counter() - t_setup
        if self.use_graph and self._graph is None and torch.device(device).type == "cuda":
            self._capture(n, device)

    CHEB_BOUNDS = (0.1, 1.1)  # [a, b] = these times the estimate of rho(D^-1 A)

    def _cheb_rhos(self) -> list:
        """rho_k of the Chebyshev recurrence; they depend only on b / a, so the coefficients c_k
        of the recorded cycle stay valid across set-ups."""
        lo, hi = self.CHEB_BOUNDS
        sigma = (hi + lo) / (hi - lo)
        rhos = [1.0 / sigma]
        for _ in range(1, max(1, self.sweeps)):
            rhos.append(1.0 / (2.0 * sigma - rhos[-1]))
        return rhos

    def _cheb_coefs(self) -> list:
        """c_k of step k (c_0 = 0): the weight of x_k - x_{k-1}."""
        r = self._cheb_rhos()
        return [0.0] + [r[k] * r[k - 1] for k in range(1, len(r))]

    def _cheb_weights(self, lam: float) -> list:
        """Scalar factors of D^-1 in step k: 1 / theta, then 2 rho_k / delta."""
        lo, hi = self.CHEB_BOUNDS
        theta, delta = 0.5 * (hi + lo) * lam, 0.5 * (hi - lo) * lam
        r = self._cheb_rhos()
        return [1.0 / theta] + [2.0 * r[k] / delta for k in range(1, len(r))]

    def _lanczos_lambda_max(self, lvl: int, A0, dinv: torch.Tensor, v0: torch.Tensor) -> float:
        """Largest eigenvalue of D^-1 A (similar to S = D^-1/2 A D^-1/2) from LANCZOS_STEPS
        Lanczos steps on S from the fixed start vector v0; the recurrence stays on the device and
        one transfer brings the tridiagonal coefficients to the host.  A Krylov estimate: a
        few warm-started power iterations, used before, under-estimated rho once plasticity
        changed the spectrum, and the smoother (and so the V-cycle) became indefinite."""
        s = torch.sqrt(torch.abs(dinv))  # |D|^-1/2: finite where an indefinite Hessian has a negative diagonal
        mult = A0.mult_tensor if lvl == 0 else self.levels[lvl - 1]["Ac64"].mult
        v = v0 / torch.linalg.vector_norm(v0)
        v_prev = torch.zeros_like(v)
        beta = torch.zeros((), dtype=v.dtype, device=v.device)
        t = torch.empty_like(v)
        y = torch.empty_like(v)
        coef = []
        for _ in range(self.LANCZOS_STEPS):
            torch.mul(s, v, out=t)
            mult(t, y)
            w = s * y - beta * v_prev
            alpha = torch.dot(w, v)
            w = w - alpha * v
            beta = torch.linalg.vector_norm(w)
            coef.append(alpha)
            coef.append(beta)
            v_prev, v = v, w / torch.clamp(beta, min=1e-300)
        c = torch.stack(coef).cpu().numpy()
        a, b = c[0::2], c[1::2]
        k = a.size
        if not np.all(np.isfinite(c)):
            raise FloatingPointError("non-finite Lanczos coefficients in the multigrid set-up")
        lam = float(sla.eigh_tridiagonal(a, b[: k - 1], eigvals_only=True, select="i",
                                         select_range=(k - 1, k - 1))[0])
        return max(lam, 1e-300)

    # ---------------------------------------------------------------------------------------
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

    def _cheb_smooth(self, L: dict, r: torch.Tensor, x: torch.Tensor | None) -> torch.Tensor:
        """``sweeps`` Chebyshev steps on A x = r from x (from zero when x is None): step k is
        x_{k+1} = x_k + c_k (x_k - x_{k-1}) + w_k (r - A x_k), one CSR pass (fne_vcycle_cheb);
        from zero the first step is x_1 = w_0 r."""
        A, W, cs = L["A"], L["wcheb"], self._cheb_coefs()
        x_prev = None
        start = 0
        if x is None:
            x = W[0] * r
            x_prev = torch.zeros_like(r)
            start = 1
        for k in range(start, W.shape[0]):
            y = x_prev if (x_prev is not None and k > 1) else torch.empty_like(r)
            A.cheb(x, x_prev if cs[k] != 0.0 else None, r, W[k], cs[k], y)
            x_prev, x = x, y
        return x

    def _cycle(self, lvl: int, r: torch.Tensor) -> torch.Tensor:
        """One V-cycle from level ``lvl``: a fused CSR pass per smoothing step, residual and
        prolongation.  Jacobi: x_{k+1} = x_k + w (r - A x_k), w = omega D^-1, from x = 0;
        Chebyshev: ``_cheb_smooth``."""
        if lvl == len(self.levels):
            return self._coarse_solve(r)
        L = self.levels[lvl]
        A, wd = L["A"], L["wd"]
        t = torch.empty_like(r)
        if self.smoother == "chebyshev":
            x = self._cheb_smooth(L, r, None)
        elif self.sweeps == 1:
            x = wd * r
        else:
            x = torch.empty_like(r)
            A.op(_SWEEP_FIRST, None, r, wd, None, x)  # the first two sweeps from x = 0
            for _ in range(self.sweeps - 2):
                A.op(_SWEEP, x, r, wd, None, t)
                x, t = t, x
        A.op(_RESIDUAL, x, r, None, None, t)
        rc = torch.empty(L["R"].n_rows, dtype=r.dtype, device=r.device)
        L["R"].mult(t, rc)
        xc = self._cycle(lvl + 1, rc)
        L["P"].op(_ADD_PRODUCT, xc, None, None, x, x)
        if self.smoother == "chebyshev":
            return self._cheb_smooth(L, r, x)
        for _ in range(self.sweeps):
            A.op(_SWEEP, x, r, wd, None, t)
            x, t = t, x
        return x

    def check_live(self) -> None:
        """Raise unless set up (the cycle reads only buffers the hierarchy owns or shares)."""
        if self._A0 is None:
            raise RuntimeError("multigrid apply before setup()")

    def apply(self, r: torch.Tensor, z: torch.Tensor) -> None:
        self.check_live()
        if self._graph is not None:
            self._g_in.copy_(r)
            self._graph.replay()
            z.copy_(self._g_out)
        else:
            z.copy_(self._cycle(0, r))

    def operator_complexity(self) -> float:
        nnz = [self._A0.nnz] + [L["Ac64"].nnz for L in self.levels]
        return float(sum(nnz) / max(1, nnz[0]))

    def summary(self) -> dict:
        return {"levels": len(self.levels) + 1, "sizes": [L["n"] for L in self.levels] + [int(self._coarse_n)],
                "operator_complexity": round(self.operator_complexity(), 3)}
