"""Re-running failed nonlinear solves with tighter settings.

A solve is described by a ``Rung`` (KSP relative tolerance, KSP iteration cap, line-search
interval).  ``RepairPolicy`` decides from a Newton result whether the solve failed in a way a
tighter rung can cure, and derives the ladder of rungs from the primary one.  Which failures
are repairable, and how the repair rung is tightened, follow the reference's benchmark policy
(``core/benchmark/repair.py:8-80``): non-finite energies, the Newton iteration cap and linear
solve failures are repaired; the repair rung divides the KSP tolerance by 10 (never below
1e-12), doubles the KSP cap and ends the line-search interval at 1.  Those rules decide which
of the reference's runs were re-run, so they are part of its results and are kept; the
representation (a failure table and a policy object) is this package's.

One class is added to the reference's: a trust-region solve that stops because every trial
step was rejected ("Trust-region radius exhausted before full convergence", "Trust-region
rejected all candidate steps").  With the line search inside the radius (the reference's
``trust_subproblem_line_search``, ``core/petsc/minimizers.py:462-481``) this happens when a
short, nearly exact Newton step is taken at alpha = 2: the quadratic model predicts no
reduction there (m(2) = m(0) for the exact step), the ratio test rejects every trial, and
shrinking the radius does not change the choice while it exceeds 2 ||p||.  The repair rung's
line-search interval ends at 1, which removes that case.  ``RepairPolicy(repair_trust=False)``
is the reference's policy exactly.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Sequence

# failure class -> phrases of the Newton driver's messages that put a result in it
FAILURE_PHRASES: dict[str, tuple[str, ...]] = {
    "nonfinite": ("non-finite", "nonfinite", "nan"),
    "iteration_cap": ("maximum number of iterations reached",),
    "linear_failure": ("linear solve failed",),
    "trust_exhausted": ("trust-region radius exhausted", "trust-region rejected all candidate steps"),
}


@dataclass(frozen=True)
class Rung:
    """Linear and line-search settings of one Newton solve."""

    label: str
    ksp_rtol: float
    ksp_max_it: int
    linesearch_interval: tuple[float, float]


def classify_failure(result: dict) -> str | None:
    """The failure class of a Newton result (``FAILURE_PHRASES`` key), or None."""
    fun = result.get("fun", math.nan)
    try:
        finite = math.isfinite(float(fun))
    except (TypeError, ValueError):
        finite = False
    msg = str(result.get("message", "")).lower()
    if not finite:
        return "nonfinite"
    for kind, phrases in FAILURE_PHRASES.items():
        if any(p in msg for p in phrases):
            return kind
    return None


@dataclass(frozen=True)
class RepairPolicy:
    """Which failures get a second solve, and how much tighter that solve is."""

    repair_nonfinite: bool = True
    repair_caps: bool = True  # the Newton iteration cap and linear solve failures
    repair_trust: bool = True  # trust-region solves ended by rejected trials (not in the reference's policy)
    rtol_factor: float = 0.1
    rtol_floor: float = 1e-12
    linesearch_upper: float = 1.0
    ksp_cap_factor: float = 2.0

    def wants_repair(self, result: dict) -> bool:
        kind = classify_failure(result)
        if kind == "nonfinite":
            return self.repair_nonfinite
        if kind in ("iteration_cap", "linear_failure"):
            return self.repair_caps
        if kind == "trust_exhausted":
            return self.repair_trust
        return False

    def repair_rung(self, primary: Rung) -> Rung | None:
        """The tightened rung, or None when the primary one cannot be tightened."""
        lo, hi = (float(v) for v in primary.linesearch_interval)
        hi_r = min(hi, float(self.linesearch_upper))
        if hi_r <= lo:
            hi_r = hi
        rtol = max(float(self.rtol_floor), float(primary.ksp_rtol) * float(self.rtol_factor))
        cap = max(int(primary.ksp_max_it) + 1, int(round(primary.ksp_max_it * float(self.ksp_cap_factor))))
        if hi_r < hi or rtol < primary.ksp_rtol or cap > primary.ksp_max_it:
            return Rung("repair", rtol, cap, (lo, hi_r))
        return None

    def ladder(self, primary: Rung) -> list[Rung]:
        if not (self.repair_nonfinite or self.repair_caps or self.repair_trust):
            return [primary]
        rep = self.repair_rung(primary)
        return [primary] if rep is None else [primary, rep]


def as_policy(repair) -> RepairPolicy | None:
    """``True`` -> the default policy, ``False``/``None`` -> no repair, a policy -> itself."""
    if repair is None or repair is False:
        return None
    if repair is True:
        return RepairPolicy()
    if isinstance(repair, RepairPolicy):
        return repair
    raise TypeError("repair must be a bool, None or a RepairPolicy")


def solve_with_repair(x, rungs: Sequence[Rung], solve_once: Callable[[Rung], dict],
                      wants_repair: Callable[[dict], bool]) -> tuple[dict, Rung, int]:
    """Solve from ``x`` with each rung in turn until a result needs no repair.

    ``x`` is solved in place; before every rung after the first it is reset to its value at
    entry (``x.copy()`` snapshots, ``snapshot.copy(x)`` restores, the PETSc ``Vec`` protocol).
    Returns (result, rung used, number of solves).
    """
    if not rungs:
        raise ValueError("at least one rung is needed")
    start = x.copy() if len(rungs) > 1 else None
    result, used = None, rungs[0]
    for k, rung in enumerate(rungs):
        if k:
            start.copy(x)
        result, used = solve_once(rung), rung
        if k + 1 == len(rungs) or not wants_repair(result):
            return result, used, k + 1
    raise AssertionError("unreachable")


def primary_rung(ksp_rtol: float, ksp_max_it: int, linesearch_interval) -> Rung:
    lo, hi = linesearch_interval
    return Rung("primary", float(ksp_rtol), int(ksp_max_it), (float(lo), float(hi)))
