"""Quasi-static load paths: one nonlinear solve per load value.

The beam is loaded by rotating its right face (``hyperelasticity/jax_petsc/solver.py:225-359``
in the reference: angle_k = k * total / n_steps, each step started from the previous state,
retried once with tighter settings when it fails, ``core/petsc/load_step_driver.py:58-110``).
The slope is loaded by lowering the strength-reduction factor lambda, each value started from
the previous one (the reference's 3D continuation,
``slope_stability_3d/.../solve_slope_stability_3d_direct_continuation.py:187-266``).
``march`` is the one loop behind both:

* ``apply(load)`` imposes the load (Dirichlet face, reduced strength);
* ``guess(history, load)`` may return an initial state (a device tensor); otherwise the step
  starts from the state the previous step left, converged or not, as in the reference;
* ``solve(x, load)`` runs the nonlinear solve in place on ``x`` and returns its result dict,
  which carries the rung used and the number of tries when the solve repairs failures itself;
* ``history`` holds ``(load, state)`` of the converged steps only, starting with
  ``initial_history`` -- the pairs a learned predictor trains on and extrapolates from;
* ``stop(record)`` ends the path after a step (the reference's ``--stop_on_fail``).
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, Iterable, Sequence


def converged(result: dict) -> bool:
    """A Newton result counts as converged when its message says so (the reference's test)."""
    return "converged" in str(result.get("message", "")).lower()


@dataclass
class StepRecord:
    step: int
    load: float
    result: dict
    seconds: float
    guessed: bool
    extra: dict = field(default_factory=dict)

    @property
    def converged(self) -> bool:
        return converged(self.result)

    def as_dict(self, load_key: str = "load") -> dict:
        r = self.result
        out = {
            "step": int(self.step),
            load_key: float(self.load),
            "nit": int(r.get("nit", 0)),
            "message": str(r.get("message", "")),
            "energy": float(r.get("fun", float("nan"))),
            "linear_iters": int(r.get("linear_iters", 0)),
            "attempt": str(r.get("attempt", "primary")),
            "tries": int(r.get("tries", 1)),
            "time": float(self.seconds),
            "predictor": bool(self.guessed),
            "converged": self.converged,
        }
        out.update(self.extra)
        return out


def march(x, loads: Iterable[tuple[int, float]], *, apply: Callable[[float], None],
          solve: Callable[[object, float], dict],
          guess: Callable[[list, float], object | None] | None = None,
          initial_history: Sequence[tuple[float, object]] = (),
          stop: Callable[[StepRecord], bool] | None = None,
          on_step: Callable[[StepRecord, list], None] | None = None,
          sync: Callable[[], None] | None = None) -> tuple[list[StepRecord], list]:
    """Walk the load path ``loads`` ((step index, load value) pairs) from state ``x``.

    ``x`` has a ``.t`` tensor holding the free-DOF state; it is updated in place and holds the
    last step's state at the end.  ``sync`` (e.g. ``torch.cuda.synchronize``) is called around
    each solve so ``seconds`` is the step's wall time.  Returns (records, history).
    """
    history = list(initial_history)
    records: list[StepRecord] = []
    for step, load in loads:
        apply(float(load))
        g = guess(history, float(load)) if guess is not None else None
        if g is not None:
            x.t.copy_(g)
        if sync is not None:
            sync()
        t0 = time.perf_counter()
        result = solve(x, float(load))
        if sync is not None:
            sync()
        rec = StepRecord(int(step), float(load), dict(result), time.perf_counter() - t0, g is not None)
        records.append(rec)
        if rec.converged:
            history.append((float(load), x.t.clone()))
        if on_step is not None:
            on_step(rec, history)
        if stop is not None and stop(rec):
            break
    return records, history


def uniform_loads(total: float, n_total: int, n_run: int | None = None, first: int = 1) -> list[tuple[int, float]]:
    """Steps ``first .. first + n_run - 1`` of ``n_total`` equal increments of ``total``."""
    n_run = int(n_total if n_run is None else n_run)
    return [(k, k * float(total) / float(n_total)) for k in range(int(first), int(first) + n_run)]
