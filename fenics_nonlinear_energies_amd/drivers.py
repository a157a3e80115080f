"""Assemble-and-solve closures and the single-step Newton driver.

Mirrors ``run_scalar_problem``'s ``_assemble_and_solve`` / ``hessian_solve_fn`` /
``trust_subproblem_solve_fn`` and its retry loop (core/petsc/scalar_problem_driver.py:154-330)
on one GPU: every callback works on device vectors, the Hessian is assembled into the
assembler's CSR by the HIP kernels and solved by the GPU KSP.
"""

from __future__ import annotations

import math
import time

import numpy as np
import torch

from . import native
from .minimizers import newton
from .repair import as_policy, primary_rung, solve_with_repair

NEWTON_DEFAULTS = {
    "tolf": 1e-5,
    "tolg": 1e-3,
    "tolg_rel": 0.0,
    "tolx_rel": 1e-6,
    "tolx_abs": 1e-10,
    "maxit": 100,
    "linesearch_tol": 1e-3,
    "linesearch_interval": (-0.5, 2.0),
    "require_all_convergence": True,
    "fail_on_nonfinite": True,
}


class DeviceGoldenSearch:
    """Golden-section line search on the device (``fne_golden_trial`` / ``fne_energy`` /
    ``fne_golden_step`` per point): the same points and comparisons as
    ``minimizers.golden_section_search`` (``core/petsc/minimizers.py:14-64``), with one host
    synchronisation per search instead of one per energy."""

    def __init__(self, assembler):
        self.asm = assembler
        dev = assembler.device
        n_state = int(native.lib().fne_golden_state_len())
        self.state = torch.zeros(n_state, dtype=torch.float64, device=dev)
        self.trial = torch.empty(assembler.layout.n_free, dtype=torch.float64, device=dev)
        self.energy = torch.zeros(1, dtype=torch.float64, device=dev)
        self._init = torch.zeros(n_state, dtype=torch.float64)
        self.searches = 0

    def _point(self, xt: torch.Tensor, pt: torch.Tensor) -> None:
        lib = native.lib()
        sp = native.stream_ptr()
        native.check(lib.fne_golden_trial(int(xt.numel()), native.ptr(xt), native.ptr(pt), native.ptr(self.state),
                                          native.ptr(self.trial), sp), "fne_golden_trial")
        self.asm.energy_device(self.trial, out=self.energy)
        native.check(lib.fne_golden_step(native.ptr(self.state), native.ptr(self.energy), sp), "fne_golden_step")

    def __call__(self, x_prev, p, lo: float, hi: float, tol: float):
        lo, hi, tol = float(lo), float(hi), float(tol)
        gamma = 0.5 + math.sqrt(5.0) / 2.0
        init = self._init
        init.zero_()
        init[0], init[1], init[10], init[13] = lo, hi, lo, tol  # lo, hi, first point, tol
        self.state.copy_(init)
        # 4 initial points, the shrink steps until hi - lo <= tol (+2 for rounding), the midpoint;
        # steps after the search has finished are no-ops
        n_loop = int(math.ceil(math.log(tol / (hi - lo)) / math.log(1.0 / gamma))) + 2 if hi - lo > tol else 0
        xt, pt = x_prev.t, p.t
        for _ in range(4 + max(n_loop, 0) + 1):
            self._point(xt, pt)
        st = self.state.cpu().numpy()
        extra = 0
        while st[11] < 7.0:  # not finished (rounding made the bracket shrink more slowly)
            if extra >= 200:
                raise RuntimeError("device golden-section search did not finish")
            self._point(xt, pt)
            extra += 1
            st = self.state.cpu().numpy()
        self.searches += 1
        return float(st[8]), float(st[9]), int(st[12])


class LinearSolveContext:
    """State shared by the assemble-and-solve callbacks (PC set-up policy, records)."""

    # Eisenstat-Walker choice 2 (Eisenstat & Walker, SIAM J. Sci. Comput. 17 (1996) 16-32):
    # eta_k = gamma (|g_k| / |g_{k-1}|)^alpha, safeguarded by gamma eta_{k-1}^alpha when that
    # exceeds 0.1, capped by ksp_rtol, floored by EW_MIN
    EW_GAMMA, EW_ALPHA, EW_MIN = 0.9, 2.0, 1e-8

    def __init__(self, assembler, *, ksp_rtol, ksp_max_it, pc_setup_on_ksp_cap=False, save_linear_timing=False,
                 keep_pc=False, pc_refresh_its=None, ksp_forcing=None):
        self.asm = assembler
        if ksp_forcing not in (None, "ew"):
            raise ValueError("ksp_forcing must be None (fixed ksp_rtol) or 'ew' (Eisenstat-Walker)")
        self.ksp_forcing = ksp_forcing
        self._ew_prev = None  # (|g|, eta) of the previous solve
        self.ksp_rtol = float(ksp_rtol)
        self.ksp_max_it = int(ksp_max_it)
        self.pc_setup_on_ksp_cap = bool(pc_setup_on_ksp_cap)
        # keep_pc: with pc_setup_on_ksp_cap, carry an existing preconditioner over from the
        # previous solve (e.g. the previous load step) instead of rebuilding it first
        self.keep_pc = bool(keep_pc)
        # with pc_setup_on_ksp_cap: also rebuild once a solve needed more than this many iterations
        self.pc_refresh_its = None if pc_refresh_its is None else int(pc_refresh_its)
        self.force_pc_setup_next = not (self.keep_pc and self._has_pc())
        self.save_linear_timing = bool(save_linear_timing)
        self.records: list[dict] = []
        self.iters: list[int] = []

    def _has_pc(self) -> bool:
        ksp = self.asm.ksp
        return getattr(ksp, "_amg", None) is not None or getattr(ksp, "_inv_diag", None) is not None

    def reset(self, ksp_rtol=None, ksp_max_it=None) -> None:
        if ksp_rtol is not None:
            self.ksp_rtol = float(ksp_rtol)
        if ksp_max_it is not None:
            self.ksp_max_it = int(ksp_max_it)
        self.force_pc_setup_next = not (self.keep_pc and self._has_pc())
        self.records = []
        self.iters = []
        self._ew_prev = None

    def _rtol_for(self, rhs) -> float:
        """KSP relative tolerance of this solve: ksp_rtol, or the Eisenstat-Walker forcing term."""
        if self.ksp_forcing is None:
            return self.ksp_rtol
        gnorm = float(rhs.norm())
        eta = self.ksp_rtol
        if self._ew_prev is not None and self._ew_prev[0] > 0.0:
            g_prev, eta_prev = self._ew_prev
            eta = self.EW_GAMMA * (gnorm / g_prev) ** self.EW_ALPHA
            guard = self.EW_GAMMA * eta_prev ** self.EW_ALPHA
            if guard > 0.1:
                eta = max(eta, guard)
            eta = min(max(eta, self.EW_MIN), self.ksp_rtol)
        self._ew_prev = (gnorm, eta)
        return eta

    def assemble_and_solve(self, vec, rhs, sol, trust_radius=None, assemble=True) -> int:
        asm = self.asm
        ksp = asm.ksp
        t0 = time.perf_counter()
        if assemble:
            asm.assemble_hessian(vec)
        t_asm = time.perf_counter() - t0
        if trust_radius is not None:
            ksp.setCGRadius(float(trust_radius))
        ksp.setOperators(asm.A)
        ksp.setTolerances(rtol=self._rtol_for(rhs), max_it=self.ksp_max_it)
        t1 = time.perf_counter()
        if self.pc_setup_on_ksp_cap:
            if self.force_pc_setup_next:
                ksp.setUp()
                self.force_pc_setup_next = False
        else:
            ksp.setUp()
        t_setup = time.perf_counter() - t1
        t2 = time.perf_counter()
        ksp.solve(rhs, sol)
        t_solve = time.perf_counter() - t2
        its = ksp.getIterationNumber()
        self.iters.append(int(its))
        if self.pc_setup_on_ksp_cap and (its >= self.ksp_max_it or (
                self.pc_refresh_its is not None and its > self.pc_refresh_its)):
            self.force_pc_setup_next = True
        if self.save_linear_timing:
            self.records.append(
                {
                    "assemble_total_time": t_asm,
                    "pc_setup_time": t_setup,
                    "solve_time": t_solve,
                    "linear_total_time": t_asm + t_setup + t_solve,
                    "ksp_its": int(its),
                    "ksp_reason_code": int(ksp.getConvergedReason()),
                    "trust_radius": None if trust_radius is None else float(trust_radius),
                }
            )
        return int(its)

    def hessian_solve_fn(self, vec, rhs, sol) -> int:
        return self.assemble_and_solve(vec, rhs, sol)

    def trust_subproblem_solve_fn(self, vec, rhs, sol, radius) -> int:
        return self.assemble_and_solve(vec, rhs, sol, trust_radius=float(radius))


def solve_newton(
    assembler,
    x,
    *,
    ksp_rtol=1e-3,
    ksp_max_it=10000,
    pc_setup_on_ksp_cap=False,
    retry_on_failure=False,
    use_trust_region=False,
    save_history=False,
    save_linear_timing=False,
    first_direction_fn=None,
    first_step_laplacian=False,
    verbose=False,
    keep_pc=False,
    pc_refresh_its=None,
    ksp_forcing=None,
    device_linesearch=True,
    **newton_kwargs,
):
    """Run Newton (with the reference's retry schedule) from device vector ``x`` in place.

    ``first_step_laplacian``: compute the first Newton direction with the assembler's
    ``assemble_first_step_hessian`` (the p = 2 Hessian for p-Laplace), the stated policy for
    zero initial guesses where the exact p > 2 Hessian vanishes identically.
    ``device_linesearch``: run the golden-section line search on the device
    (``DeviceGoldenSearch``; same points and comparisons, one synchronisation per search).
    ``ksp_forcing="ew"``: the KSP tolerance of each Newton iteration follows the
    Eisenstat-Walker forcing term, capped by ``ksp_rtol`` (the reference uses ``ksp_rtol``
    throughout, which is the default here too).
    """
    opts = dict(NEWTON_DEFAULTS)
    opts.update(newton_kwargs)
    ls = tuple(opts.pop("linesearch_interval"))
    ctx = LinearSolveContext(
        assembler, ksp_rtol=ksp_rtol, ksp_max_it=ksp_max_it, pc_setup_on_ksp_cap=pc_setup_on_ksp_cap,
        save_linear_timing=save_linear_timing, keep_pc=keep_pc, pc_refresh_its=pc_refresh_its,
        ksp_forcing=ksp_forcing,
    )
    if first_step_laplacian and first_direction_fn is None:
        def first_direction_fn(vec, rhs, sol):
            assembler.assemble_first_step_hessian(vec)
            return ctx.assemble_and_solve(vec, rhs, sol, assemble=False)

    trust_ksp = bool(use_trust_region and str(assembler.ksp.ksp_type).lower() in {"stcg", "nash", "gltr"})
    policy = as_policy(retry_on_failure)
    primary = primary_rung(ksp_rtol, ksp_max_it, ls)
    rungs = [primary] if policy is None else policy.ladder(primary)
    dls = None
    if device_linesearch and getattr(assembler, "device", None) is not None and hasattr(assembler, "energy_device"):
        dls = getattr(assembler, "_device_golden", None)
        if dls is None:
            dls = assembler._device_golden = DeviceGoldenSearch(assembler)
    timing = {"t": 0.0}

    def solve_once(rung):
        ctx.reset(rung.ksp_rtol, rung.ksp_max_it)
        t0 = time.perf_counter()
        res = newton(
            energy_fn=assembler.energy_fn,
            gradient_fn=assembler.gradient_fn,
            hessian_solve_fn=ctx.hessian_solve_fn,
            x=x,
            linesearch_interval=rung.linesearch_interval,
            verbose=verbose,
            hessian_matvec_fn=lambda _x, vin, vout: assembler.A.mult(vin, vout),
            trust_subproblem_solve_fn=ctx.trust_subproblem_solve_fn if trust_ksp else None,
            save_history=save_history,
            trust_region=use_trust_region,
            first_direction_fn=first_direction_fn,
            device_linesearch=dls,
            **opts,
        )
        timing["t"] = time.perf_counter() - t0
        return res

    result, used, tries = solve_with_repair(
        x, rungs, solve_once, policy.wants_repair if policy is not None else (lambda _r: False))
    result = dict(result)
    result["linear_iters"] = int(sum(ctx.iters))
    result["linear_iters_list"] = list(ctx.iters)
    result["solve_time"] = float(timing["t"])
    result["attempt"] = used.label
    result["tries"] = int(tries)
    result["ksp_rtol_used"] = float(used.ksp_rtol)
    result["ksp_max_it_used"] = int(used.ksp_max_it)
    result["linesearch_interval_used"] = list(used.linesearch_interval)
    result["linear_timing"] = list(ctx.records)
    return result
