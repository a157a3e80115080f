"""Load-step / continuation orchestration (core/petsc/load_step_driver.py:58-135 of the reference).

``run_load_steps`` walks the steps, calls ``prepare_step`` (e.g. rotate the Dirichlet face
or apply the next strength-reduction factor), tries the attempt schedule with retries and
collects one record per step.  ``warm_start`` (optional, not in the reference) returns the
initial guess for a step from the previous converged states; it is how the learned warm
start plugs in, with the previous state as the fallback.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Any, Callable, Mapping, Sequence


@dataclass(frozen=True)
class LoadStepAttempt:
    name: str
    linesearch_interval: tuple[float, float]
    linear_rtol: float | None = None
    linear_max_it: int | None = None
    metadata: Mapping[str, object] = field(default_factory=dict)


@dataclass
class LoadStepContext:
    step: int
    angle: float
    state: dict = field(default_factory=dict)


def attempts_from_tuples(attempts) -> list[LoadStepAttempt]:
    return [
        LoadStepAttempt(
            name=str(name),
            linesearch_interval=(float(iv[0]), float(iv[1])),
            linear_rtol=None if rt is None else float(rt),
            linear_max_it=None if mi is None else int(mi),
        )
        for name, iv, rt, mi in attempts
    ]


def run_load_steps(
    *,
    start_step: int,
    num_steps: int,
    rotation_per_step: float,
    prepare_step: Callable[[LoadStepContext], None],
    build_attempts: Callable[[LoadStepContext], Sequence[LoadStepAttempt]],
    solve_attempt: Callable[[LoadStepContext, LoadStepAttempt], tuple[dict, float]],
    should_retry: Callable[[dict, LoadStepContext], bool],
    build_step_record: Callable[[LoadStepContext, dict, float, LoadStepAttempt], dict],
    should_stop: Callable[[dict, dict, LoadStepContext], bool] | None = None,
    on_retry: Callable[[LoadStepContext, LoadStepAttempt, int, int], None] | None = None,
    on_step_complete: Callable[[dict, LoadStepContext], None] | None = None,
) -> list[dict]:
    records: list[dict] = []
    for step in range(int(start_step), int(start_step) + int(num_steps)):
        ctx = LoadStepContext(step=int(step), angle=float(step * float(rotation_per_step)))
        prepare_step(ctx)
        attempts = list(build_attempts(ctx))
        if not attempts:
            raise ValueError("build_attempts() must return at least one attempt")
        result = None
        step_time = 0.0
        used = attempts[0]
        for idx, attempt in enumerate(attempts):
            result, step_time = solve_attempt(ctx, attempt)
            used = attempt
            if should_retry(result, ctx) and idx + 1 < len(attempts):
                if on_retry is not None:
                    on_retry(ctx, attempt, idx + 1, len(attempts))
                continue
            break
        if result is None:
            raise RuntimeError("Load-step driver did not receive a nonlinear result")
        rec = build_step_record(ctx, result, step_time, used)
        records.append(rec)
        if on_step_complete is not None:
            on_step_complete(rec, ctx)
        if should_stop is not None and should_stop(rec, result, ctx):
            break
    return records


def build_load_step_result(*, mesh_level, total_dofs, setup_time, total_runtime_start, steps, extra=None) -> dict[str, Any]:
    payload: dict[str, Any] = {
        "mesh_level": int(mesh_level),
        "total_dofs": int(total_dofs),
        "setup_time": float(round(setup_time, 6)),
        "solve_time_total": float(round(sum(float(s.get("time", 0.0)) for s in steps), 6)),
        "total_time": float(round(time.perf_counter() - total_runtime_start, 6)),
        "steps": [dict(s) for s in steps],
    }
    if extra:
        payload.update(extra)
    return payload
