"""Retry policy for failed nonlinear solves (core/benchmark/repair.py:8-80 of the reference)."""

from __future__ import annotations

import numpy as np


def needs_solver_repair(result: dict, *, retry_on_nonfinite: bool = True, retry_on_maxit: bool = True) -> bool:
    msg = str(result.get("message", "")).lower()
    nonfinite = (
        not np.isfinite(float(result.get("fun", np.nan)))
        or "non-finite" in msg
        or "nonfinite" in msg
        or "nan" in msg
    )
    maxit = "maximum number of iterations reached" in msg
    linear_failure = "linear solve failed" in msg
    if retry_on_nonfinite and nonfinite:
        return True
    return bool(retry_on_maxit and (maxit or linear_failure))


def build_retry_attempts(
    *,
    retry_on_failure: bool | None = None,
    retry_on_nonfinite: bool | None = None,
    retry_on_maxit: bool | None = None,
    linesearch_interval: tuple[float, float],
    ksp_rtol: float,
    ksp_max_it: int,
    retry_rtol_factor: float = 0.1,
    retry_linesearch_b: float = 1.0,
    retry_ksp_max_it_factor: float = 2.0,
    min_rtol: float = 1e-12,
) -> list[tuple[str, tuple[float, float], float, int]]:
    """Primary attempt plus, when repair is wanted, a tighter 'repair' attempt."""
    a, b = float(linesearch_interval[0]), float(linesearch_interval[1])
    attempts = [("primary", (a, b), float(ksp_rtol), int(ksp_max_it))]
    wants = bool(retry_on_failure)
    if retry_on_nonfinite is not None or retry_on_maxit is not None:
        wants = bool(retry_on_nonfinite) or bool(retry_on_maxit)
    if not wants:
        return attempts
    rb = min(b, float(retry_linesearch_b))
    if rb <= a:
        rb = b
    rtol = max(float(min_rtol), float(ksp_rtol) * float(retry_rtol_factor))
    max_it = max(int(ksp_max_it) + 1, int(round(float(ksp_max_it) * float(retry_ksp_max_it_factor))))
    if rb < b or rtol < float(ksp_rtol) or max_it > int(ksp_max_it):
        attempts.append(("repair", (a, rb), rtol, max_it))
    return attempts
