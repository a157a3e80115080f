"""Strength-reduction continuation for the Mohr-Coulomb slope (rows A5, A14).

The reference solves one reduced-strength problem per run
(``slope_stability/jax_petsc/solver.py:590-603``: Davis-B reduction at ``lambda_target``,
zero initial guess, Newton with tolf 1e-4, tolg 1e-3, tolg_rel 1e-3, line search on
[-0.5, 2], ``solve_slope_stability_dof.py:327-357``) and chains runs only in 3D, warm-starting
each lambda from the previous converged state
(``slope_stability_3d/jax_petsc/solve_slope_stability_3d_direct_continuation.py:187-266``).
``run_lambda_continuation`` is that loop on one GPU: per step it applies the reduction,
takes the initial guess from ``predictor`` (previous state by default, the fallback of the
learned warm start) and runs the Newton driver on device vectors.
"""

from __future__ import annotations

from typing import Callable, Sequence

import torch

from .drivers import solve_newton
from .problems.slope_stability import davis_b_reduction

SLOPE_NEWTON = dict(tolf=1e-4, tolg=1e-3, tolg_rel=1e-3, maxit=100, linesearch_interval=(-0.5, 2.0),
                    linesearch_tol=1e-3)


def elastic_initial_guess(asm, *, rtol: float = 1e-8, max_it: int = 20000):
    """Solve K u = f with the elastic tangent (``build_elastic_initial_guess``, ``mesh.py:785-799``).

    In elastic mode the energy is 1/2 u.K u - f.u, so one Newton step from u = 0 is exact.
    Returns the device vector and the CG iteration count.
    """
    x = asm.create_vec(None)
    g = x.duplicate()
    d = x.duplicate()
    prev = getattr(asm, "_constitutive_mode", 0)
    asm.set_constitutive_mode(1)
    try:
        asm.hessian_device(x.t, g.t)  # gradient at 0 is -f
        g.scale(-1.0)
        ksp = asm.ksp
        ksp.setOperators(asm.A)
        ksp.setTolerances(rtol=rtol, max_it=max_it)
        ksp.setUp()
        ksp.solve(g, d)
        its = ksp.getIterationNumber()
    finally:
        asm.set_constitutive_mode(prev)
    return d, int(its)


def run_lambda_continuation(
    asm,
    lambdas: Sequence[float],
    x0,
    *,
    predictor: Callable[[list, float], torch.Tensor | None] | None = None,
    on_step: Callable[[dict, list], None] | None = None,
    ksp_rtol: float = 1e-3,
    ksp_max_it: int = 2000,
    pc_setup_on_ksp_cap: bool = True,
    pc_refresh_its: int | None = None,
    retry_on_failure=False,
    stop_on_fail: bool = False,
    newton_kwargs: dict | None = None,
) -> list[dict]:
    """Solve the reduced-strength problem for each lambda, warm-starting from the history.

    ``x0``: device vector (e.g. the elastic initial guess), overwritten with the last state.
    ``predictor(history, lam)`` returns an initial guess tensor or None (then the state the
    previous step left is used).  ``history`` holds (lam, state tensor) pairs of converged
    steps.  With ``pc_setup_on_ksp_cap`` the preconditioner is kept across Newton iterations
    and steps and rebuilt when a solve hits ``ksp_max_it`` or needs more than
    ``pc_refresh_its``.  ``retry_on_failure`` is off by default, as in the reference's slope
    solver (``solve_slope_stability_dof.py:378-381``).
    """
    from .load_path import march

    p = asm.params
    phi_raw = float(p.get("phi_deg_raw", p["phi_deg"]))
    p["phi_deg_raw"] = phi_raw
    opts = dict(SLOPE_NEWTON)
    opts.update(newton_kwargs or {})
    strength = {}

    def apply(lam):
        c, phi = davis_b_reduction(float(p["c0"]), phi_raw, float(p["psi_deg"]), float(lam))
        asm.set_strength(c, phi)
        strength.update(cohesion=float(c), phi_deg=float(phi))

    def guess(history, lam):
        return predictor(history, lam) if (predictor is not None and history) else None

    def solve(vec, _lam):
        return solve_newton(asm, vec, ksp_rtol=ksp_rtol, ksp_max_it=ksp_max_it, pc_setup_on_ksp_cap=pc_setup_on_ksp_cap,
                            keep_pc=pc_setup_on_ksp_cap, pc_refresh_its=pc_refresh_its,
                            retry_on_failure=retry_on_failure, **opts)

    def notify(rec, history):
        rec.extra.update(strength)
        rec.extra["u_max"] = float(x0.t.abs().max().item())
        if predictor is not None and hasattr(predictor, "accepted") and rec.guessed and predictor.accepted:
            rec.extra["learned_accepted"] = bool(predictor.accepted[-1])
        if on_step is not None:
            on_step(rec.as_dict("lambda"), history)

    recs, _ = march(x0, list(enumerate(float(v) for v in lambdas)), apply=apply, solve=solve, guess=guess,
                    stop=(lambda r: not r.converged) if stop_on_fail else None, on_step=notify,
                    sync=torch.cuda.synchronize)
    return [r.as_dict("lambda") for r in recs]
