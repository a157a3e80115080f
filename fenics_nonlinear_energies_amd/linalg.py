"""CSR matrix object and Krylov solvers on the GPU (the ``A`` / ``ksp`` pair of the assembler).

* ``DeviceCSR.mult`` is the CSR SpMV hot loop (replaces PETSc ``MatMult``,
  ``hessian_matvec_fn = A.mult`` in core/petsc/scalar_problem_driver.py:294).
* ``KSP`` reproduces the PETSc KSP semantics the reference relies on -- zero initial guess,
  left preconditioning, convergence on the preconditioned residual norm
  ``||B r_k|| <= max(rtol*||B r_0||, abstol)``, divergence at ``dtol``, iteration cap -- for
  ``cg`` (KSPCG) and ``gmres`` (KSPGMRES, restart 30, classical Gram-Schmidt without
  refinement as PETSc's default orthogonalization), and ``stcg``
  (Steihaug-Toint CG with a radius, as set by KSPCGSetRadius in core/petsc/trust_ksp.py),
  which tests the unpreconditioned residual norm ``||r_k||`` instead, as PETSc's KSPSTCG does
  by default.
  The preconditioner is point Jacobi (``jacobi``; zero diagonal entries are replaced by 1
  as PETSc's PCJACOBI does), ``none``, ``gamg`` (alias ``hypre``): the smoothed-aggregation
  AMG V-cycle of ``amg.py`` with the problem's near-nullspace, ``mg``: the same V-cycle under
  given top-level transfers (the same-mesh p-multigrid of ``pmg.py``), or ``python``: a
  user context (``setPythonContext``).
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import native
from .vec import DeviceVec, dot_device


def lanes_per_row(avg_row_len: float) -> int:
    """Lanes per CSR row of the SpMV kernels (``csr_lanes_per_row`` of the library): about three
    entries per lane for short rows (4 lanes for the 7-entry P1 rows); from 8 lanes up the lanes
    read groups of four entries, and one pass covers an average row plus its unaligned start
    (16 lanes for the 40-entry tetrahedral and 47-entry P4 vector rows)."""
    lpr = 1
    while lpr < 64 and 3 * lpr < avg_row_len:
        lpr *= 2
    if lpr >= 8:
        while lpr < 64 and 4 * lpr < avg_row_len + 3:
            lpr *= 2
    return lpr


# precision-argument field (bits 16..19) of fne_spmv / fne_vcycle_op / fne_vcycle_cheb: the
# CSR-stream form (a workgroup streams 256 / lanes consecutive rows' entries through LDS with
# 16-byte loads, then lanes per row sum them)
SPMV_FORM_STREAM = 1 << 16


def stream_lanes_per_row(avg_row_len: float) -> int:
    """Lanes per row of the CSR-stream form (``csr_stream_lanes`` of the library): about eight
    entries per lane -- 1 for the 7-entry P1 rows (256 rows per workgroup), 4 for the 40-50-entry
    vector rows."""
    lpr = 1
    while lpr < 64 and 2 * lpr <= avg_row_len / 8.0:
        lpr *= 2
    return lpr


def spmv_code(fne_precision: int, avg_row_len: float) -> int:
    """The precision argument of the CSR row kernels: value type, CSR-stream form and its lanes
    per row (kernels given unaligned arrays fall back to lanes per row with the same count)."""
    return int(fne_precision) | (stream_lanes_per_row(avg_row_len) << 8) | SPMV_FORM_STREAM


class DeviceCSR:
    """Square CSR matrix on the GPU with a fixed pattern; values FP64 or FP32."""

    def __init__(self, indptr, indices, n: int, precision: str = "fp64", device="cuda"):
        self.n = int(n)
        self.indptr = torch.as_tensor(np.asarray(indptr, dtype=np.int32)).to(device)
        self.indices = torch.as_tensor(np.asarray(indices, dtype=np.int32)).to(device)
        self.nnz = int(self.indices.numel())
        self.set_precision(precision)
        row_len = np.diff(np.asarray(indptr, dtype=np.int64))
        self.avg_row_len = float(row_len.mean()) if row_len.size else 1.0
        self.lanes_per_row = stream_lanes_per_row(self.avg_row_len)

    def set_precision(self, precision: str) -> None:
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        self.precision = precision
        dtype = torch.float64 if precision == "fp64" else torch.float32
        self.data = torch.zeros(self.nnz, dtype=dtype, device=self.indices.device)

    @property
    def fne_precision(self) -> int:
        return native.FNE_FP64 if self.precision == "fp64" else native.FNE_FP32

    def getSize(self):
        return (self.n, self.n)

    def getType(self) -> str:
        return "seqaijhip"

    def mult(self, x: DeviceVec, y: DeviceVec) -> None:
        self.mult_tensor(x.t, y.t)

    def mult_tensor(self, x: torch.Tensor, y: torch.Tensor) -> None:
        prec = spmv_code(self.fne_precision, self.avg_row_len)
        native.check(
            native.lib().fne_spmv(
                self.n,
                native.ptr(self.indptr),
                native.ptr(self.indices),
                native.ptr(self.data),
                prec,
                native.ptr(x),
                native.ptr(y),
                native.stream_ptr(),
            ),
            "fne_spmv",
        )

    def diagonal(self) -> torch.Tensor:
        d = torch.empty(self.n, dtype=torch.float64, device=self.indices.device)
        native.check(
            native.lib().fne_csr_diag(
                self.n,
                native.ptr(self.indptr),
                native.ptr(self.indices),
                native.ptr(self.data),
                self.fne_precision,
                native.ptr(d),
                native.stream_ptr(),
            ),
            "fne_csr_diag",
        )
        return d

    def getDiagonal(self, d: DeviceVec) -> None:
        """PETSc ``MatGetDiagonal``: fill ``d`` with the diagonal."""
        d.t.copy_(self.diagonal())

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix(
            (self.data.double().cpu().numpy(), self.indices.cpu().numpy(), self.indptr.cpu().numpy()),
            shape=(self.n, self.n),
        )

    def destroy(self) -> None:
        self.data = None


class MatrixFreeOperator:
    """A shell matrix: products and diagonal come from callbacks instead of stored values.

    The call surface of the reference's PETSc Python matrix over ``_MatrixFreeHessianContext``
    (``slope_stability/jax_petsc/reordered_element_assembler.py:25-50``, created at ``:313-327``):
    ``mult`` / ``multTranspose`` (the operator is symmetric), ``getDiagonal``, ``createVecs``,
    ``duplicate`` and ``stateIncrease``.  ``mult_fn(x, y)`` and ``diag_fn(d)`` act on device
    FP64 tensors over the free DOFs.  A KSP given this operator runs its host-driven Krylov loops
    (the device CG kernels read CSR values); its preconditioner comes from the optional second
    matrix of ``KSP.setOperators`` or, for Jacobi, from ``getDiagonal``.
    """

    def __init__(self, n: int, mult_fn, diag_fn=None, *, block_size: int = 1, device="cuda", name: str = "shell"):
        self.n = int(n)
        self._mult = mult_fn
        self._diag = diag_fn
        self.block_size = int(block_size)
        self.device = device
        self.name = str(name)
        self.state = 0
        self.near_nullspace = None

    def getSize(self):
        return (self.n, self.n)

    def getType(self) -> str:
        return "python"

    def getBlockSize(self) -> int:
        return self.block_size

    def setNearNullSpace(self, modes) -> None:
        self.near_nullspace = modes

    def stateIncrease(self) -> None:
        """Mark the operator as changed (a new linearisation point)."""
        self.state += 1

    def mult_tensor(self, x: torch.Tensor, y: torch.Tensor) -> None:
        self._mult(x, y)

    def mult(self, x: DeviceVec, y: DeviceVec) -> None:
        self._mult(x.t, y.t)

    def multTranspose(self, x: DeviceVec, y: DeviceVec) -> None:
        self._mult(x.t, y.t)

    def diagonal(self) -> torch.Tensor:
        if self._diag is None:
            raise RuntimeError(f"matrix-free operator {self.name!r} has no diagonal")
        d = torch.empty(self.n, dtype=torch.float64, device=self.device)
        self._diag(d)
        return d

    def getDiagonal(self, d: DeviceVec) -> None:
        if self._diag is None:
            raise RuntimeError(f"matrix-free operator {self.name!r} has no diagonal")
        self._diag(d.t)

    def createVecs(self):
        return DeviceVec.zeros(self.n, device=self.device), DeviceVec.zeros(self.n, device=self.device)

    def duplicate(self, copy: bool = False) -> "MatrixFreeOperator":
        del copy
        op = MatrixFreeOperator(self.n, self._mult, self._diag, block_size=self.block_size, device=self.device,
                                name=self.name)
        op.near_nullspace = self.near_nullspace
        return op

    def destroy(self) -> None:
        self._mult = None
        self._diag = None


# preconditioners built by amg.SmoothedAggregationPC: smoothed aggregation ("gamg", alias
# "hypre") and the p-multigrid with given top-level transfers ("mg", PETSc PCMG)
MULTIGRID_PC_TYPES = ("gamg", "hypre", "mg")

# KSP converged reasons (PETSc numbering)
KSP_CONVERGED_RTOL = 2
KSP_CONVERGED_ATOL = 3
KSP_CONVERGED_CG_NEG_CURVE = 5
KSP_CONVERGED_STEP_LENGTH = 6
KSP_DIVERGED_ITS = -3
KSP_DIVERGED_DTOL = -4
KSP_DIVERGED_BREAKDOWN = -5
KSP_DIVERGED_INDEFINITE_PC = -8
KSP_DIVERGED_NANORINF = -9
KSP_DIVERGED_INDEFINITE_MAT = -10


class KSP:
    """Krylov solver with the PETSc KSP call surface used by the reference drivers."""

    def __init__(self, ksp_type="cg", pc_type="jacobi", rtol=1e-3, max_it=10000, atol=1e-50, dtol=1e5, restart=30):
        self.ksp_type = str(ksp_type)
        self.pc_type = str(pc_type)
        self.rtol = float(rtol)
        self.atol = float(atol)
        self.dtol = float(dtol)
        self.max_it = int(max_it)
        self.restart = int(restart)
        self.radius = None
        self.A: DeviceCSR | MatrixFreeOperator | None = None
        self.P: DeviceCSR | None = None  # preconditioner matrix (None: the operator itself)
        self._inv_diag: torch.Tensor | None = None
        self.its = 0
        self.reason = 0
        self.rnorm = float("nan")
        self.pc_setup_count = 0
        self.check_every = 16  # device CG: iterations enqueued between polls of the stop flag
        # the same with the multigrid graph: a cycle enqueued after the stop is wasted, so poll
        # more often (every 2 iterations measured 6 % faster than 1 or 4 on the 320x16x16 beam)
        self.check_every_graph_pc = 2
        # device GMRES / FGMRES: Arnoldi steps enqueued between polls of the cycle-end flag
        self.check_every_gmres = 8
        self.near_nullspace = None  # (n, k) host array for pc_type "gamg" (rigid modes, constants)
        self.near_nullspace_fn = None  # optional callable -> (n, k) array at the current state, or None
        self.pc_precision = None  # multigrid storage precision ("fp64" / "fp32"); None: the operator's
        self._amg = None
        self.python_pc = None  # pc_type "python": a context with setUp(pc) and apply(pc, b, x)
        # pc_type "mg": host CSR transfers of the top levels (finest first; e.g. the same-mesh
        # P4 -> P2 -> P1 interpolations), the near-nullspace below them (None: injected) and
        # the smoothing sweeps per level; aggregation continues below the last transfer
        self.mg_prolongators = None
        self.mg_coarse_near_nullspace = None
        self.mg_sweeps = 2  # smoothing steps per level, each side (-mg_levels_ksp_max_it)
        self.mg_smoother = "chebyshev"  # "chebyshev" | "jacobi" (-mg_levels_ksp_type chebyshev | richardson)
        self.mg_coarse_factor = "host"  # where the coarsest level is factorised: "host" | "device"

    # PETSc-style configuration -----------------------------------------------------------
    def setType(self, t: str) -> None:
        self.ksp_type = str(t)

    def setTolerances(self, rtol=None, atol=None, divtol=None, max_it=None) -> None:
        if rtol is not None:
            self.rtol = float(rtol)
        if atol is not None:
            self.atol = float(atol)
        if divtol is not None:
            self.dtol = float(divtol)
        if max_it is not None:
            self.max_it = int(max_it)

    def setOperators(self, A, P=None) -> None:
        """PETSc ``KSPSetOperators(A, P)``: Krylov products with ``A``, preconditioner built from
        ``P`` (default ``A``).  ``A`` may be a ``MatrixFreeOperator``."""
        self.A = A
        self.P = None if P is None or P is A else P

    def _pmat(self):
        return self.P if self.P is not None else self.A

    def setPythonContext(self, ctx) -> None:
        """PETSc ``PCSetPythonContext`` for pc_type "python" (``ctx.setUp(pc)``, ``ctx.apply(pc, b, x)``)."""
        self.python_pc = ctx

    def _device_loop(self) -> bool:
        """The device CG kernels need the operator's CSR values and a preconditioner built from
        that same matrix; anything else runs the host-driven loops."""
        return isinstance(self.A, DeviceCSR) and self.P is None and self.pc_type != "python"

    def setCGRadius(self, radius: float) -> None:
        self.radius = float(radius)

    def setUp(self) -> None:
        """Preconditioner set-up (Jacobi: inverse diagonal, zero entries -> 1)."""
        if self.A is None:
            raise RuntimeError("KSP.setUp called before setOperators")
        pmat = self._pmat()
        self._amg = None if self.pc_type not in MULTIGRID_PC_TYPES else getattr(self, "_amg", None)
        if self.pc_type == "jacobi":
            d = pmat.diagonal()
            d = torch.where(d == 0.0, torch.ones_like(d), d)
            self._inv_diag = 1.0 / d
        elif self.pc_type == "none":
            self._inv_diag = None
        elif self.pc_type == "python":
            if self.python_pc is None:
                raise RuntimeError("pc_type 'python' needs setPythonContext(ctx) first")
            self.python_pc.setUp(None)
            self._inv_diag = None
        elif self.pc_type in MULTIGRID_PC_TYPES:
            if not isinstance(pmat, DeviceCSR):
                raise ValueError(f"pc_type {self.pc_type!r} needs an assembled preconditioner matrix "
                                 "(setOperators(A, P))")
            prec = self.pc_precision or pmat.precision
            if self._amg is None or self._amg.precision != prec:
                prepared = getattr(self, "_amg_prepared", None)
                if prepared is not None and prepared.precision == prec:
                    self._amg = prepared
                else:
                    self._amg = self._new_multigrid(prec)
            B = self.near_nullspace_fn() if self.near_nullspace_fn is not None else None
            self._amg.setup(pmat, near_nullspace=B)
            self._inv_diag = None
        else:
            raise ValueError(f"unsupported pc_type {self.pc_type!r} (jacobi | gamg | mg | python | none)")
        self.pc_setup_count += 1

    def _new_multigrid(self, precision: str):
        from .amg import SmoothedAggregationPC

        if self.pc_type == "mg":
            if not self.mg_prolongators:
                raise ValueError("pc_type 'mg' needs the level transfers (KSP.mg_prolongators)")
            return SmoothedAggregationPC(self.near_nullspace, precision=precision,
                                         fixed_prolongators=self.mg_prolongators,
                                         coarse_near_nullspace=self.mg_coarse_near_nullspace,
                                         sweeps=self.mg_sweeps, smoother=self.mg_smoother,
                                         coarse_factor=self.mg_coarse_factor)
        return SmoothedAggregationPC(self.near_nullspace, precision=precision, sweeps=self.mg_sweeps,
                                     smoother=self.mg_smoother, coarse_factor=self.mg_coarse_factor)

    def prepare_pc(self, A: DeviceCSR) -> float:
        """Mesh preprocessing for the multigrid preconditioners: the symbolic phase (patterns of
        every level's products) for the operator pattern of ``A``.  Returns its host time; the
        first setUp then only runs the numeric phase."""
        if self.pc_type not in MULTIGRID_PC_TYPES:
            return 0.0
        pc = self._new_multigrid(self.pc_precision or A.precision)
        k = 1 if self.near_nullspace is None else int(np.asarray(self.near_nullspace).shape[1])
        pc.prepare(A, k)
        self._amg_prepared = pc
        return pc.symbolic_time

    def getIterationNumber(self) -> int:
        return int(self.its)

    def getConvergedReason(self) -> int:
        return int(self.reason)

    def getResidualNorm(self) -> float:
        return float(self.rnorm)

    def destroy(self) -> None:
        self.A = None
        self.P = None
        self._inv_diag = None

    # -------------------------------------------------------------------------------------
    def _pc(self, r: torch.Tensor, z: torch.Tensor) -> None:
        if self._amg is not None:
            self._amg.apply(r, z)
        elif self.pc_type == "python":
            self.python_pc.apply(None, r, z)
        elif self._inv_diag is None:
            z.copy_(r)
        else:
            torch.mul(r, self._inv_diag, out=z)

    def _dot(self, a: torch.Tensor, b: torch.Tensor) -> float:
        return float(dot_device(a, b).item())

    def _converged(self, it: int, rnorm: float, rnorm0: float) -> int:
        if not math.isfinite(rnorm):
            return KSP_DIVERGED_NANORINF
        if rnorm <= max(self.rtol * rnorm0, self.atol):
            return KSP_CONVERGED_RTOL if rnorm > self.atol else KSP_CONVERGED_ATOL
        if it > 0 and rnorm >= self.dtol * rnorm0:
            return KSP_DIVERGED_DTOL
        return 0

    def solve(self, b: DeviceVec, x: DeviceVec) -> None:
        if (self._inv_diag is None and self.pc_type == "jacobi") or (
            self._amg is None and self.pc_type in MULTIGRID_PC_TYPES
        ):
            self.setUp()
        t = self.ksp_type.lower()
        if t == "cg":
            self._solve_cg(b.t, x.t)
        elif t == "gmres":
            self._solve_gmres(b.t, x.t)
        elif t == "fgmres":
            self._solve_fgmres(b.t, x.t)
        elif t == "stcg":
            self._solve_stcg(b.t, x.t)
        else:
            raise ValueError(f"unsupported ksp_type {self.ksp_type!r} (cg | gmres | fgmres | stcg)")

    # KSPCG on the device (fne_pcg): scalars stay on the GPU, the host polls the stop flag
    def _solve_cg(self, b: torch.Tensor, x: torch.Tensor) -> None:
        import ctypes as C

        amg = self._amg
        if not self._device_loop() or (amg is not None and getattr(amg, "_graph", None) is None):
            # a shell operator, a separate preconditioner matrix or a multigrid without a
            # recorded cycle
            self._solve_cg_host(b, x)
            return

        A = self.A
        lib = native.lib()
        nbytes = int(lib.fne_pcg_workspace_bytes(A.n, A.nnz))
        if nbytes < 0:
            native.check(-1, "fne_pcg_workspace_bytes")
        if getattr(self, "_cg_work", None) is None or self._cg_work.numel() < nbytes:
            self._cg_work = torch.empty(nbytes, dtype=torch.uint8, device=b.device)
        its, reason, rnorm = C.c_int32(0), C.c_int32(0), C.c_double(0.0)
        if amg is not None:
            # the V-cycle graph is launched from inside the device iteration (fne_pcg_graph);
            # solves are short with a multigrid preconditioner, so the stop flag is read often
            amg.check_live()
            if amg._A0 is not A:
                raise RuntimeError("the KSP operator changed since the multigrid set-up; call setUp()")
            native.check(
                lib.fne_pcg_graph(A.n, native.ptr(A.indptr), native.ptr(A.indices), native.ptr(A.data),
                                  A.fne_precision, C.c_void_p(amg._graph.raw_cuda_graph_exec()), native.ptr(amg._g_in),
                                  native.ptr(amg._g_out), native.ptr(b), native.ptr(x), self.rtol, self.atol,
                                  self.dtol, self.max_it, int(self.check_every_graph_pc),
                                  native.ptr(self._cg_work), nbytes, C.byref(its), C.byref(reason), C.byref(rnorm),
                                  native.stream_ptr()),
                "fne_pcg_graph",
            )
        else:
            native.check(
                lib.fne_pcg(A.n, native.ptr(A.indptr), native.ptr(A.indices), native.ptr(A.data), A.fne_precision,
                            native.ptr(self._inv_diag), native.ptr(b), native.ptr(x), self.rtol, self.atol, self.dtol,
                            self.max_it, int(self.check_every), native.ptr(self._cg_work), nbytes, C.byref(its),
                            C.byref(reason), C.byref(rnorm), native.stream_ptr()),
                "fne_pcg",
            )
        self.its = int(its.value)
        self.reason = int(reason.value)
        self.rnorm = float(rnorm.value)

    # KSPCG, host-driven restatement (kept for checking the device solver) ----------------
    def _solve_cg_host(self, b: torch.Tensor, x: torch.Tensor) -> None:
        A = self.A
        x.zero_()
        r = b.clone()
        z = torch.empty_like(b)
        w = torch.empty_like(b)
        self._pc(r, z)
        dp = math.sqrt(max(self._dot(z, z), 0.0))
        rnorm0 = dp
        self.its = 0
        self.rnorm = dp
        self.reason = self._converged(0, dp, rnorm0)
        if self.reason:
            return
        beta = self._dot(z, r)
        if beta < 0.0:
            self.reason = KSP_DIVERGED_INDEFINITE_PC
            return
        p = z.clone()
        dpiold = None
        betaold = beta
        for i in range(self.max_it):
            if i > 0:
                p.mul_(beta / betaold).add_(z)
            A.mult_tensor(p, w)
            dpi = self._dot(p, w)
            if dpi == 0.0 or (dpiold is not None and dpi * dpiold <= 0.0) or not math.isfinite(dpi):
                self.reason = KSP_DIVERGED_INDEFINITE_MAT if math.isfinite(dpi) else KSP_DIVERGED_NANORINF
                return
            dpiold = dpi
            a = beta / dpi
            x.add_(p, alpha=a)
            r.add_(w, alpha=-a)
            self._pc(r, z)
            dp = math.sqrt(max(self._dot(z, z), 0.0))
            self.its = i + 1
            self.rnorm = dp
            self.reason = self._converged(i + 1, dp, rnorm0)
            if self.reason:
                return
            betaold = beta
            beta = self._dot(z, r)
            if beta < 0.0:
                self.reason = KSP_DIVERGED_INDEFINITE_PC
                return
        self.reason = KSP_DIVERGED_ITS

    # GMRES / FGMRES on the device (fne_gmres) whenever the operator is the assembled matrix and
    # the preconditioner Jacobi, none or a recorded multigrid cycle; else the host-driven loops
    def _device_gmres(self, b: torch.Tensor, x: torch.Tensor, flexible: bool) -> bool:
        import ctypes as C

        amg = self._amg
        if not self._device_loop() or (amg is not None and getattr(amg, "_graph", None) is None):
            return False
        m = max(1, self.restart)
        if m > 512:
            return False
        A = self.A
        lib = native.lib()
        nbytes = int(lib.fne_gmres_workspace_bytes(A.n, m, int(flexible)))
        if nbytes < 0:
            native.check(-1, "fne_gmres_workspace_bytes")
        if getattr(self, "_gm_work", None) is None or self._gm_work.numel() < nbytes:
            self._gm_work = torch.empty(nbytes, dtype=torch.uint8, device=b.device)
        its, reason, rnorm = C.c_int32(0), C.c_int32(0), C.c_double(0.0)
        graph = pc_in = pc_out = None
        if amg is not None:
            amg.check_live()
            if amg._A0 is not A:
                raise RuntimeError("the KSP operator changed since the multigrid set-up; call setUp()")
            graph = C.c_void_p(amg._graph.raw_cuda_graph_exec())
            pc_in, pc_out = amg._g_in, amg._g_out
        native.check(
            lib.fne_gmres(A.n, native.ptr(A.indptr), native.ptr(A.indices), native.ptr(A.data), A.fne_precision,
                          native.ptr(self._inv_diag) if amg is None else None, graph, native.ptr(pc_in),
                          native.ptr(pc_out), native.ptr(b), native.ptr(x), self.rtol, self.atol, self.dtol,
                          self.max_it, m, int(flexible), int(self.check_every_gmres), native.ptr(self._gm_work),
                          nbytes, C.byref(its), C.byref(reason), C.byref(rnorm), native.stream_ptr()),
            "fne_gmres",
        )
        self.its = int(its.value)
        self.reason = int(reason.value)
        self.rnorm = float(rnorm.value)
        return True

    def _solve_gmres(self, b: torch.Tensor, x: torch.Tensor) -> None:
        if not self._device_gmres(b, x, flexible=False):
            self._solve_gmres_host(b, x)

    def _solve_fgmres(self, b: torch.Tensor, x: torch.Tensor) -> None:
        if not self._device_gmres(b, x, flexible=True):
            self._solve_fgmres_host(b, x)

    # KSPGMRES(restart), left preconditioning, classical Gram-Schmidt: host-driven on device
    # vectors (shell operators, Python preconditioners; kept for checking the device solver) --
    def _solve_gmres_host(self, b: torch.Tensor, x: torch.Tensor) -> None:
        A = self.A
        m = max(1, self.restart)
        n = b.numel()
        x.zero_()
        V = torch.empty((m + 1, n), dtype=torch.float64, device=b.device)
        w = torch.empty_like(b)
        tmp = torch.empty_like(b)
        self.its = 0
        rnorm0 = None
        while True:
            # preconditioned residual r = B (b - A x)
            if self.its == 0:
                tmp.copy_(b)
            else:
                A.mult_tensor(x, tmp)
                torch.sub(b, tmp, out=tmp)
            self._pc(tmp, V[0])
            beta = math.sqrt(max(self._dot(V[0], V[0]), 0.0))
            if rnorm0 is None:
                rnorm0 = beta
                self.rnorm = beta
                self.reason = self._converged(0, beta, rnorm0)
                if self.reason:
                    return
            if beta == 0.0:
                self.reason = KSP_CONVERGED_ATOL
                return
            V[0].mul_(1.0 / beta)
            H = np.zeros((m + 1, m), dtype=np.float64)
            g = np.zeros(m + 1, dtype=np.float64)
            g[0] = beta
            cs = np.zeros(m, dtype=np.float64)
            sn = np.zeros(m, dtype=np.float64)
            k_used = 0
            for k in range(m):
                A.mult_tensor(V[k], tmp)
                self._pc(tmp, w)
                # classical Gram-Schmidt, PETSc's default (all projections of the same w, no
                # refinement): one multi-dot, one multi-axpy and one read-back per iteration
                hk = torch.mv(V[: k + 1], w)
                w.sub_(torch.mv(V[: k + 1].T, hk))
                hv = torch.cat((hk, torch.dot(w, w).view(1))).cpu().numpy()
                H[: k + 1, k] = hv[: k + 1]
                hn = math.sqrt(max(float(hv[k + 1]), 0.0))
                H[k + 1, k] = hn
                if hn > 0.0:
                    torch.mul(w, 1.0 / hn, out=V[k + 1])
                for j in range(k):
                    t0 = cs[j] * H[j, k] + sn[j] * H[j + 1, k]
                    H[j + 1, k] = -sn[j] * H[j, k] + cs[j] * H[j + 1, k]
                    H[j, k] = t0
                den = math.hypot(H[k, k], H[k + 1, k])
                if den == 0.0:
                    cs[k], sn[k] = 1.0, 0.0
                else:
                    cs[k], sn[k] = H[k, k] / den, H[k + 1, k] / den
                H[k, k] = cs[k] * H[k, k] + sn[k] * H[k + 1, k]
                H[k + 1, k] = 0.0
                g[k + 1] = -sn[k] * g[k]
                g[k] = cs[k] * g[k]
                k_used = k + 1
                self.its += 1
                self.rnorm = abs(g[k + 1])
                self.reason = self._converged(self.its, self.rnorm, rnorm0)
                if self.reason or self.its >= self.max_it or hn == 0.0:
                    break
            # x += V_k y,  H_k y = g_k
            y = np.zeros(k_used, dtype=np.float64)
            for i in range(k_used - 1, -1, -1):
                s = g[i] - float(np.dot(H[i, i + 1 : k_used], y[i + 1 : k_used]))
                y[i] = s / H[i, i] if H[i, i] != 0.0 else 0.0
            if k_used:
                coeff = torch.as_tensor(y, dtype=torch.float64, device=b.device)
                x.add_(coeff @ V[:k_used])
            if self.reason:
                return
            if self.its >= self.max_it:
                self.reason = KSP_DIVERGED_ITS
                return

    # KSPFGMRES(restart): flexible GMRES, right preconditioning (the preconditioner may change
    # between iterations, so its outputs z_k are kept and x += Z y), classical Gram-Schmidt as
    # KSPGMRES, convergence on the unpreconditioned residual norm ||b - A x|| (PETSc's norm for
    # right preconditioning) -- the outer solver of the reference's slope p-multigrid runs
    # (ksp_type "fgmres", docs/problems/Plasticity.md:240-250) -------------------------------
    def _solve_fgmres_host(self, b: torch.Tensor, x: torch.Tensor) -> None:
        A = self.A
        m = max(1, self.restart)
        n = b.numel()
        x.zero_()
        V = torch.empty((m + 1, n), dtype=torch.float64, device=b.device)
        Z = torch.empty((m, n), dtype=torch.float64, device=b.device)
        w = torch.empty_like(b)
        self.its = 0
        rnorm0 = None
        while True:
            if self.its == 0:
                V[0].copy_(b)
            else:
                A.mult_tensor(x, V[0])
                torch.sub(b, V[0], out=V[0])
            beta = math.sqrt(max(self._dot(V[0], V[0]), 0.0))
            if rnorm0 is None:
                rnorm0 = beta
                self.rnorm = beta
                self.reason = self._converged(0, beta, rnorm0)
                if self.reason:
                    return
            if beta == 0.0:
                self.reason = KSP_CONVERGED_ATOL
                return
            V[0].mul_(1.0 / beta)
            H = np.zeros((m + 1, m), dtype=np.float64)
            g = np.zeros(m + 1, dtype=np.float64)
            g[0] = beta
            cs = np.zeros(m, dtype=np.float64)
            sn = np.zeros(m, dtype=np.float64)
            k_used = 0
            for k in range(m):
                self._pc(V[k], Z[k])
                A.mult_tensor(Z[k], w)
                hk = torch.mv(V[: k + 1], w)
                w.sub_(torch.mv(V[: k + 1].T, hk))
                hv = torch.cat((hk, torch.dot(w, w).view(1))).cpu().numpy()
                H[: k + 1, k] = hv[: k + 1]
                hn = math.sqrt(max(float(hv[k + 1]), 0.0))
                H[k + 1, k] = hn
                if hn > 0.0:
                    torch.mul(w, 1.0 / hn, out=V[k + 1])
                for j in range(k):
                    t0 = cs[j] * H[j, k] + sn[j] * H[j + 1, k]
                    H[j + 1, k] = -sn[j] * H[j, k] + cs[j] * H[j + 1, k]
                    H[j, k] = t0
                den = math.hypot(H[k, k], H[k + 1, k])
                if den == 0.0:
                    cs[k], sn[k] = 1.0, 0.0
                else:
                    cs[k], sn[k] = H[k, k] / den, H[k + 1, k] / den
                H[k, k] = cs[k] * H[k, k] + sn[k] * H[k + 1, k]
                H[k + 1, k] = 0.0
                g[k + 1] = -sn[k] * g[k]
                g[k] = cs[k] * g[k]
                k_used = k + 1
                self.its += 1
                self.rnorm = abs(g[k + 1])
                self.reason = self._converged(self.its, self.rnorm, rnorm0)
                if self.reason or self.its >= self.max_it or hn == 0.0:
                    break
            y = np.zeros(k_used, dtype=np.float64)
            for i in range(k_used - 1, -1, -1):
                s_ = g[i] - float(np.dot(H[i, i + 1 : k_used], y[i + 1 : k_used]))
                y[i] = s_ / H[i, i] if H[i, i] != 0.0 else 0.0
            if k_used:
                x.add_(torch.as_tensor(y, dtype=torch.float64, device=b.device) @ Z[:k_used])
            if self.reason:
                return
            if self.its >= self.max_it:
                self.reason = KSP_DIVERGED_ITS
                return

    # Steihaug-Toint truncated CG inside radius: PETSc KSPSTCG as the trust-region Newton uses
    # it (core/petsc/trust_ksp.py:55-72), restated for the serial path in
    # core/serial/sparse_solvers.py:20-93.  Euclidean norm on the step (PETSc's default
    # direction type), convergence on the unpreconditioned residual norm ||r|| (PETSc STCG's
    # default norm, the serial restatement's test), boundary / negative-curvature stops ----
    def _solve_stcg(self, b: torch.Tensor, x: torch.Tensor) -> None:
        # on the device (fne_pstcg / fne_pstcg_graph): the CG kernels in trust mode
        import ctypes as C

        amg = self._amg
        if not self._device_loop() or (amg is not None and getattr(amg, "_graph", None) is None):
            self._solve_stcg_host(b, x)
            return
        A = self.A
        delta = float("inf") if self.radius is None else float(self.radius)
        lib = native.lib()
        nbytes = int(lib.fne_pcg_workspace_bytes(A.n, A.nnz))
        if nbytes < 0:
            native.check(-1, "fne_pcg_workspace_bytes")
        if getattr(self, "_cg_work", None) is None or self._cg_work.numel() < nbytes:
            self._cg_work = torch.empty(nbytes, dtype=torch.uint8, device=b.device)
        its, reason, rnorm = C.c_int32(0), C.c_int32(0), C.c_double(0.0)
        if amg is not None:
            amg.check_live()
            if amg._A0 is not A:
                raise RuntimeError("the KSP operator changed since the multigrid set-up; call setUp()")
            native.check(
                lib.fne_pstcg_graph(A.n, native.ptr(A.indptr), native.ptr(A.indices), native.ptr(A.data),
                                    A.fne_precision, C.c_void_p(amg._graph.raw_cuda_graph_exec()),
                                    native.ptr(amg._g_in), native.ptr(amg._g_out), native.ptr(b), native.ptr(x), delta,
                                    self.rtol, self.atol, self.dtol, self.max_it, int(self.check_every_graph_pc),
                                    native.ptr(self._cg_work), nbytes, C.byref(its), C.byref(reason), C.byref(rnorm),
                                    native.stream_ptr()),
                "fne_pstcg_graph",
            )
        else:
            native.check(
                lib.fne_pstcg(A.n, native.ptr(A.indptr), native.ptr(A.indices), native.ptr(A.data), A.fne_precision,
                              native.ptr(self._inv_diag), native.ptr(b), native.ptr(x), delta, self.rtol, self.atol,
                              self.dtol, self.max_it, int(self.check_every), native.ptr(self._cg_work), nbytes,
                              C.byref(its), C.byref(reason), C.byref(rnorm), native.stream_ptr()),
                "fne_pstcg",
            )
        self.its = int(its.value)
        self.reason = int(reason.value)
        self.rnorm = float(rnorm.value)

    # the same Steihaug-Toint CG driven from the host (kept for checking the device solver)
    def _solve_stcg_host(self, b: torch.Tensor, x: torch.Tensor) -> None:
        A = self.A
        delta = float("inf") if self.radius is None else float(self.radius)
        x.zero_()
        r = b.clone()
        z = torch.empty_like(b)
        w = torch.empty_like(b)
        self._pc(r, z)
        dp = math.sqrt(max(self._dot(r, r), 0.0))
        rnorm0 = dp
        self.its = 0
        self.rnorm = dp
        self.reason = self._converged(0, dp, rnorm0)
        if self.reason:
            return
        rz = self._dot(r, z)
        if rz < 0.0:
            self.reason = KSP_DIVERGED_INDEFINITE_PC
            return
        p = z.clone()
        xx = 0.0  # ||x||^2
        for i in range(self.max_it):
            A.mult_tensor(p, w)
            kappa = self._dot(p, w)
            if not math.isfinite(kappa):
                self.its = i + 1
                self.reason = KSP_DIVERGED_NANORINF
                return
            xp = self._dot(x, p)
            pp = self._dot(p, p)
            if kappa <= 0.0:
                # negative curvature: follow p to the boundary (no boundary: keep the iterate)
                if math.isfinite(delta):
                    tau = (-xp + math.sqrt(max(xp * xp + pp * (delta * delta - xx), 0.0))) / pp
                    x.add_(p, alpha=tau)
                self.its = i + 1
                self.reason = KSP_CONVERGED_CG_NEG_CURVE
                return
            alpha = rz / kappa
            xx_new = xx + 2.0 * alpha * xp + alpha * alpha * pp
            if xx_new >= delta * delta:
                tau = (-xp + math.sqrt(max(xp * xp + pp * (delta * delta - xx), 0.0))) / pp
                x.add_(p, alpha=tau)
                self.its = i + 1
                self.reason = KSP_CONVERGED_STEP_LENGTH
                return
            x.add_(p, alpha=alpha)
            xx = xx_new
            r.add_(w, alpha=-alpha)
            self._pc(r, z)
            dp = math.sqrt(max(self._dot(r, r), 0.0))
            self.its = i + 1
            self.rnorm = dp
            self.reason = self._converged(i + 1, dp, rnorm0)
            if self.reason:
                return
            rz_new = self._dot(r, z)
            if rz_new < 0.0:
                self.reason = KSP_DIVERGED_INDEFINITE_PC
                return
            p.mul_(rz_new / rz).add_(z)
            rz = rz_new
        self.reason = KSP_DIVERGED_ITS
