"""Host logic around the Newton driver: the preconditioner set-up policy and the
Eisenstat-Walker forcing term of the linear solves, the retry schedule, the line-search
interval repair and the load-step loop.

Set-up policy: by default before every solve; with pc_setup_on_ksp_cap only on the first
solve and after a solve that hit ksp_max_it (core/petsc/scalar_problem_driver.py:150-202),
here also after one above pc_refresh_its, and keep_pc skips the first set-up when a
preconditioner exists.

Forcing term: eta_0 = ksp_rtol; eta_k = gamma (|g_k| / |g_{k-1}|)^alpha with gamma = 0.9, alpha = 2,
raised to gamma eta_{k-1}^alpha when that exceeds 0.1, capped by ksp_rtol and floored at
1e-8 (Eisenstat & Walker 1996, choice 2); the fixed policy always returns ksp_rtol.
"""

import numpy as np
import pytest

from fenics_nonlinear_energies_amd.drivers import LinearSolveContext


class _Vec:
    def __init__(self, n):
        self.n = n

    def norm(self):
        return self.n


class _Asm:
    class ksp:  # noqa: N801 - attribute bag standing in for the assembler's KSP
        _amg = None
        _inv_diag = None


def _ctx(forcing, rtol=0.1):
    return LinearSolveContext(_Asm(), ksp_rtol=rtol, ksp_max_it=100, ksp_forcing=forcing)


def test_fixed_policy_returns_ksp_rtol():
    c = _ctx(None)
    assert [c._rtol_for(_Vec(g)) for g in (10.0, 1.0, 1e-3)] == [0.1, 0.1, 0.1]


def test_eisenstat_walker_sequence():
    c = _ctx("ew")
    assert c._rtol_for(_Vec(1.0)) == 0.1  # first solve: ksp_rtol
    # fast decrease: 0.9 * (0.01)^2 = 9e-5, safeguard 0.9 * 0.1^2 = 9e-3 <= 0.1 -> not applied
    assert c._rtol_for(_Vec(0.01)) == pytest.approx(9e-5)
    # slow decrease: 0.9 * 0.9^2 = 0.729 -> capped by ksp_rtol
    assert c._rtol_for(_Vec(0.009)) == pytest.approx(0.1)
    # after eta = 0.1 the safeguard 0.9 * 0.1^2 = 9e-3 is below 0.1, so the plain term holds
    assert c._rtol_for(_Vec(0.009 * 1e-3)) == pytest.approx(max(0.9 * 1e-6, 1e-8))
    c.reset()
    assert c._rtol_for(_Vec(5.0)) == 0.1  # reset starts over


def test_eisenstat_walker_safeguard_and_floor():
    c = _ctx("ew", rtol=0.9)
    c._rtol_for(_Vec(1.0))  # eta_0 = 0.9
    # plain term 0.9 * (1e-6)^2 is tiny, safeguard 0.9 * 0.9^2 = 0.729 > 0.1 keeps it up
    assert c._rtol_for(_Vec(1e-6)) == pytest.approx(0.729)
    c2 = _ctx("ew", rtol=0.05)
    c2._rtol_for(_Vec(1.0))
    assert c2._rtol_for(_Vec(1e-9)) == pytest.approx(1e-8)  # floor


def test_unknown_forcing_rejected():
    with pytest.raises(ValueError):
        _ctx("bogus")


class _ScriptedKsp:
    """Records set-ups and radii; each solve reports the next scripted iteration count."""

    def __init__(self, its, has_pc=False):
        self.script = list(its)
        self.setups = []
        self.radii = []
        self.k = 0
        self._amg = object() if has_pc else None
        self._inv_diag = None
        self._its = 0

    def setCGRadius(self, r):
        self.radii.append(r)

    def setOperators(self, A):
        pass

    def setTolerances(self, rtol=None, max_it=None):
        self.max_it = max_it

    def setUp(self):
        self.setups.append(self.k)

    def solve(self, rhs, sol):
        self._its = self.script[self.k]
        self.k += 1

    def getIterationNumber(self):
        return self._its

    def getConvergedReason(self):
        return 2


class _ScriptedAsm:
    def __init__(self, ksp):
        self.ksp = ksp
        self.A = None

    def assemble_hessian(self, vec):
        pass


def _solves(its, **kw):
    ksp = _ScriptedKsp(its, has_pc=kw.pop("has_pc", False))
    c = LinearSolveContext(_ScriptedAsm(ksp), ksp_rtol=0.1, ksp_max_it=30, **kw)
    for _ in its:
        c.hessian_solve_fn(None, _Vec(1.0), None)
    return ksp.setups, c.iters


def test_pc_set_up_every_solve_by_default():
    """pc_setup_on_ksp_cap off (core/petsc/scalar_problem_driver.py:186-192): set up before
    every solve."""
    setups, iters = _solves([5, 30, 7, 12])
    assert setups == [0, 1, 2, 3] and iters == [5, 30, 7, 12]


def test_pc_rebuilt_only_after_a_solve_hits_the_cap():
    """pc_setup_on_ksp_cap (scalar_problem_driver.py:186-202): set up on the first solve, then
    only after a solve reached ksp_max_it (30)."""
    setups, _ = _solves([5, 30, 7, 12, 31, 4], pc_setup_on_ksp_cap=True)
    assert setups == [0, 2, 5]


def test_pc_refresh_above_a_threshold_and_kept_across_calls():
    """Additionally rebuilt after a solve above pc_refresh_its; keep_pc carries an existing
    preconditioner over, so the first solve sets nothing up."""
    setups, _ = _solves([5, 21, 7, 20, 22, 3], pc_setup_on_ksp_cap=True, pc_refresh_its=20)
    assert setups == [0, 2, 5]
    setups, _ = _solves([5, 21, 7], pc_setup_on_ksp_cap=True, pc_refresh_its=20, keep_pc=True, has_pc=True)
    assert setups == [2]


def test_trust_radius_reaches_the_ksp():
    ksp = _ScriptedKsp([3, 4])
    c = LinearSolveContext(_ScriptedAsm(ksp), ksp_rtol=0.1, ksp_max_it=30)
    c.trust_subproblem_solve_fn(None, _Vec(1.0), None, 0.5)
    c.hessian_solve_fn(None, _Vec(1.0), None)
    assert ksp.radii == [0.5]


def test_repair_ladder():
    """Primary rung, plus with repair a tighter rung: rtol x 0.1 (floored at 1e-12), the
    line-search interval's upper end capped at 1, the KSP cap doubled (the reference's policy,
    core/benchmark/repair.py:29-80); nothing to tighten -> no repair rung."""
    from fenics_nonlinear_energies_amd.repair import Rung, RepairPolicy, primary_rung

    prim = primary_rung(1e-1, 30, (-0.5, 2.0))
    assert prim == Rung("primary", 1e-1, 30, (-0.5, 2.0))
    lad = RepairPolicy().ladder(prim)
    assert len(lad) == 2 and lad[0] == prim
    rep = lad[1]
    assert rep.label == "repair" and rep.linesearch_interval == (-0.5, 1.0) and rep.ksp_max_it == 60
    assert rep.ksp_rtol == pytest.approx(1e-2)
    assert RepairPolicy(repair_nonfinite=False, repair_caps=False, repair_trust=False).ladder(prim) == [prim]
    # rtol already at the floor, interval already inside [., 1]: only the cap moves (1 -> 2)
    tiny = RepairPolicy().ladder(primary_rung(1e-12, 1, (-0.5, 0.5)))
    assert tiny[1] == Rung("repair", 1e-12, 2, (-0.5, 0.5))
    # an interval that the cap would empty keeps its upper end
    assert RepairPolicy().ladder(primary_rung(1e-3, 10, (1.5, 2.0)))[1].linesearch_interval == (1.5, 2.0)


def test_which_results_trigger_a_repair():
    from fenics_nonlinear_energies_amd.repair import RepairPolicy, classify_failure

    pol = RepairPolicy()
    assert classify_failure({"fun": float("nan"), "message": "Converged"}) == "nonfinite"
    assert pol.wants_repair({"fun": float("nan"), "message": "Converged"})
    assert pol.wants_repair({"fun": 1.0, "message": "Non-finite energy after update at Newton iteration 3"})
    assert pol.wants_repair({"fun": 1.0, "message": "Maximum number of iterations reached"})
    assert pol.wants_repair({"fun": 1.0, "message": "Linear solve failed at Newton iteration 2"})
    assert not RepairPolicy(repair_caps=False).wants_repair({"fun": 1.0, "message": "Maximum number of iterations reached"})
    assert not RepairPolicy(repair_nonfinite=False).wants_repair({"fun": float("inf"), "message": "x"})
    assert not pol.wants_repair({"fun": 1.0, "message": "Converged (energy, step, gradient)"})
    # trust-region solves ended by rejected trials: repaired here, not by the reference's policy
    # (RepairPolicy(repair_trust=False))
    for msg in ("Trust-region rejected all candidate steps at Newton iteration 4",
                "Trust-region radius exhausted before full convergence"):
        assert classify_failure({"fun": 1.0, "message": msg}) == "trust_exhausted"
        assert pol.wants_repair({"fun": 1.0, "message": msg})
        assert not RepairPolicy(repair_trust=False).wants_repair({"fun": 1.0, "message": msg})
    assert not pol.wants_repair({"fun": 1.0, "message": "Trust-region step converged"})


class _HostVec:
    """numpy-backed stand-in with the PETSc copy protocol and a torch view (``.t``)."""

    def __init__(self, a):
        import torch

        self.t = torch.as_tensor(np.asarray(a, dtype=np.float64)).clone()

    def copy(self, other=None):
        if other is None:
            return _HostVec(self.t.numpy())
        other.t.copy_(self.t)
        return other


def test_solve_with_repair_restarts_from_the_entry_state():
    """The repair rung starts from the state at entry, not from where the failed solve ended;
    a result that needs no repair ends the ladder; the last rung's result is kept either way."""
    from fenics_nonlinear_energies_amd.repair import RepairPolicy, primary_rung, solve_with_repair

    pol = RepairPolicy()
    rungs = pol.ladder(primary_rung(1e-1, 30, (-0.5, 2.0)))
    x = _HostVec([1.0, 2.0])
    seen = []

    def fail_then_ok(rung):
        seen.append((rung.label, x.t.tolist()))
        x.t.add_(10.0)
        return {"fun": 1.0, "message": "Maximum number of iterations reached" if rung.label == "primary"
                else "Converged (energy, step, gradient)"}

    res, used, tries = solve_with_repair(x, rungs, fail_then_ok, pol.wants_repair)
    assert seen == [("primary", [1.0, 2.0]), ("repair", [1.0, 2.0])]
    assert used.label == "repair" and tries == 2 and x.t.tolist() == [11.0, 12.0]
    res, used, tries = solve_with_repair(x, rungs, lambda r: {"fun": 0.0, "message": "Converged"}, pol.wants_repair)
    assert used.label == "primary" and tries == 1
    res, used, tries = solve_with_repair(x, rungs, lambda r: {"fun": float("nan"), "message": ""}, pol.wants_repair)
    assert used.label == "repair" and tries == 2 and res["message"] == ""
    with pytest.raises(ValueError):
        solve_with_repair(x, [], fail_then_ok, pol.wants_repair)


def test_line_search_interval_repair_bisects_to_the_finite_region():
    """f is finite for alpha < 0.7 and for alpha > -0.3 only: both ends of [-0.5, 2] are pulled
    back to within tol of the boundaries, from the finite side; finite ends are kept."""
    from fenics_nonlinear_energies_amd.minimizers import repair_linesearch_interval

    f = lambda a: a * a if -0.3 < a < 0.7 else float("inf")
    a, b, evals, repaired = repair_linesearch_interval(f, -0.5, 2.0, center=0.0, tol=1e-6)
    assert repaired and -0.3 < a <= -0.3 + 1e-6 and 0.7 - 1e-6 <= b < 0.7
    assert evals > 2
    a, b, _, repaired = repair_linesearch_interval(lambda t: t * t, -0.5, 2.0, tol=1e-6)
    assert (a, b, repaired) == (-0.5, 2.0, False)
    # a non-finite centre leaves the interval alone
    a, b, _, repaired = repair_linesearch_interval(f, -0.5, 2.0, center=1.0, tol=1e-6)
    assert (a, b, repaired) == (-0.5, 2.0, False)


def test_load_path_history_guess_and_stop():
    """march: loads applied in order, the guess hook sees only converged steps (an unconverged
    step is recorded but not added to the history, and the next step starts from the state it
    left), the stop test ends the path after the step it fires on."""
    from fenics_nonlinear_energies_amd.load_path import march, uniform_loads

    loads = uniform_loads(1.0, 4, 5)  # last one beyond the total: allowed, k * total / n
    assert [k for k, _ in loads] == [1, 2, 3, 4, 5] and loads[1][1] == pytest.approx(0.5)
    applied, starts, guesses = [], [], []
    x = _HostVec([0.0])

    def solve(vec, load):
        starts.append(float(vec.t[0]))
        vec.t.fill_(load)
        return {"fun": load, "nit": 3, "message": "Maximum number of iterations reached" if load == 0.5
                else "Converged (energy, step, gradient)", "attempt": "primary"}

    def guess(history, load):
        guesses.append([h[0] for h in history])
        return None if load < 0.75 else history[-1][1] * 2.0

    recs, hist = march(x, loads, apply=applied.append, solve=solve, guess=guess, initial_history=[(0.0, x.t.clone())],
                       stop=lambda r: r.load >= 1.0)
    assert applied == [0.25, 0.5, 0.75, 1.0]
    assert guesses == [[0.0], [0.0, 0.25], [0.0, 0.25], [0.0, 0.25, 0.75]]
    assert starts == [0.0, 0.25, 0.5, 1.5]  # step 3 guessed 2 x 0.25, step 4 guessed 2 x 0.75
    assert [r.converged for r in recs] == [True, False, True, True]
    assert [h[0] for h in hist] == [0.0, 0.25, 0.75, 1.0]
    d = recs[1].as_dict("angle")
    assert d["angle"] == 0.5 and d["nit"] == 3 and not d["converged"] and d["predictor"] is False
    assert recs[2].as_dict()["predictor"] is True


def test_load_path_repairs_a_failed_step():
    """A load path whose solve repairs failures: the failed step is solved again from its own
    start with the repair rung, and the record names the rung and the number of tries."""
    from fenics_nonlinear_energies_amd.load_path import march, uniform_loads
    from fenics_nonlinear_energies_amd.repair import RepairPolicy, primary_rung, solve_with_repair

    pol = RepairPolicy()
    calls = []
    x = _HostVec([0.0])

    def solve(vec, load):
        def once(rung):
            calls.append((load, rung.label, float(vec.t[0])))
            vec.t.add_(1.0)
            bad = load == 2.0 and rung.label == "primary"
            return {"fun": float("nan") if bad else load, "message": "Converged"}
        res, rung, tries = solve_with_repair(vec, pol.ladder(primary_rung(1e-1, 30, (-0.5, 2.0))), once, pol.wants_repair)
        return dict(res, attempt=rung.label, tries=tries)

    recs, _ = march(x, uniform_loads(3.0, 3), apply=lambda v: None, solve=solve)
    assert calls == [(1.0, "primary", 0.0), (2.0, "primary", 1.0), (2.0, "repair", 1.0), (3.0, "primary", 2.0)]
    assert [r.as_dict()["attempt"] for r in recs] == ["primary", "repair", "primary"]
    assert [r.as_dict()["tries"] for r in recs] == [1, 2, 1]
