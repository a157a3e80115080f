"""Newton minimiser with golden-section / Armijo / bisection line search and trust regions.

Same control flow, stopping rules, messages and history keys as the reference's
core/petsc/minimizers.py (``golden_section_search`` :14-62, ``_repair_linesearch_interval``
:74-126, ``_solve_trust_1d/2d`` :129-178, ``newton`` :181-1038), so iteration counts are
comparable and the retry policy (``repair.needs_solver_repair``) reads the same messages.
Vectors are any objects with the PETSc.Vec methods used here -- on the GPU these are
:class:`~fenics_nonlinear_energies_amd.vec.DeviceVec`; scalars (energies, norms, dots) are
FP64 host values read back from deterministic device reductions.
"""

from __future__ import annotations

import time

import numpy as np

NORM_2 = 2


def golden_section_search(f, a, b, tol):
    """Minimise ``f`` on [a, b] by golden-section search; returns (alpha, n_evals)."""
    gamma = 0.5 + np.sqrt(5.0) / 2.0
    lo, hi = float(a), float(b)
    f_lo, f_hi = f(lo), f(hi)
    d = (hi - lo) / gamma + lo
    c = lo + hi - d
    f_c, f_d = f(c), f(d)
    n_evals = 4
    best_a, best_f = lo, f_lo
    for cand, val in ((hi, f_hi), (c, f_c), (d, f_d)):
        if val < best_f:
            best_a, best_f = cand, val
    while hi - lo > tol:
        if f_c < f_d:
            hi = d
            d, c = c, lo + hi - c
            f_d, f_c = f_c, f(c)
            probes = ((d, f_d), (c, f_c))
        else:
            lo = c
            c, d = d, lo + hi - d
            f_c, f_d = f_d, f(d)
            probes = ((c, f_c), (d, f_d))
        for cand, val in probes:
            if val < best_f:
                best_a, best_f = cand, val
        n_evals += 1
    mid = 0.5 * (lo + hi)
    f_mid = f(mid)
    n_evals += 1
    if f_mid < best_f:
        best_a = mid
    return float(best_a), n_evals


def repair_linesearch_interval(f, a, b, center=0.0, center_value=None, tol=1e-6, max_bisect=60):
    """Shrink non-finite interval ends back toward a finite centre by bisection."""
    center, a, b = float(center), float(a), float(b)
    tol = max(float(tol), 1e-12)
    state = {"evals": 0, "repaired": False}
    if center_value is None:
        center_value = f(center)
        state["evals"] += 1
    if not np.isfinite(center_value):
        return a, b, state["evals"], False

    def side(bound, sign):
        if (sign < 0 and not bound < center) or (sign > 0 and not bound > center):
            return bound
        state["evals"] += 1
        if np.isfinite(f(bound)):
            return bound
        state["repaired"] = True
        good, bad = center, bound
        for _ in range(max_bisect):
            mid = 0.5 * (good + bad)
            state["evals"] += 1
            if np.isfinite(f(mid)):
                good = mid
            else:
                bad = mid
            if abs(bad - good) <= tol:
                break
        return good

    a = side(a, -1.0)
    b = side(b, +1.0)
    return a, b, state["evals"], state["repaired"]


def solve_trust_1d(rhs, curv, delta):
    """argmin rhs*t + curv*t^2/2 subject to |t| <= delta."""
    delta = max(0.0, float(delta))
    if delta <= 0.0:
        return 0.0
    if np.isfinite(curv) and curv > 1e-14:
        t = -rhs / curv
    else:
        t = -delta if rhs > 0.0 else (delta if rhs < 0.0 else 0.0)
    return float(np.clip(t, -delta, delta))


def _quad(rhs, mm, s):
    return float(rhs.dot(s) + 0.5 * s.dot(mm.dot(s)))


def solve_trust_2d(rhs, mm, delta):
    """2-D trust-region model: interior stationary point vs a 181-point boundary scan."""
    delta = max(0.0, float(delta))
    s0 = np.zeros(2)
    best_s, best_q = s0, _quad(rhs, mm, s0)
    if delta <= 0.0:
        return best_s, best_q
    try:
        su = np.linalg.solve(mm, -rhs)
        if np.all(np.isfinite(su)) and np.linalg.norm(su) <= delta:
            qu = _quad(rhs, mm, su)
            if qu < best_q:
                best_s, best_q = su, qu
    except np.linalg.LinAlgError:
        pass
    for theta in np.linspace(0.0, 2.0 * np.pi, num=181, endpoint=False):
        st = delta * np.array([np.cos(theta), np.sin(theta)])
        qt = _quad(rhs, mm, st)
        if qt < best_q:
            best_s, best_q = st, qt
    return best_s, best_q


def _sanitize_trust(shrink, expand, eta_s, eta_e, max_reject):
    shrink = float(shrink)
    expand = float(expand)
    if not np.isfinite(shrink) or shrink <= 0.0 or shrink >= 1.0:
        shrink = 0.5
    if not np.isfinite(expand) or expand < 1.0:
        expand = 1.5
    eta_s = 0.05 if not np.isfinite(eta_s) else float(eta_s)
    eta_e = 0.75 if not np.isfinite(eta_e) else float(eta_e)
    eta_s = min(max(eta_s, 0.0), 1.0)
    eta_e = min(max(eta_e, eta_s), 1.0)
    return shrink, expand, eta_s, eta_e, max(0, int(max_reject))


def newton(
    energy_fn,
    gradient_fn,
    hessian_solve_fn,
    x,
    tolf=1e-5,
    tolg=1e-3,
    tolg_rel=0.0,
    linesearch_tol=1e-3,
    linesearch_interval=(-0.5, 2.0),
    line_search="golden_fixed",
    armijo_alpha0=1.0,
    armijo_c1=1e-4,
    armijo_shrink=0.5,
    armijo_max_ls=40,
    armijo_gradient_fallback=False,
    maxit=100,
    tolx_rel=1e-6,
    tolx_abs=1e-10,
    require_all_convergence=False,
    fail_on_nonfinite=True,
    verbose=False,
    comm=None,
    ghost_update_fn=None,
    project_fn=None,
    hessian_matvec_fn=None,
    trust_subproblem_solve_fn=None,
    trust_subproblem_line_search=False,
    save_history=False,
    trust_region=False,
    trust_radius_init=1.0,
    trust_radius_min=1e-8,
    trust_radius_max=1e6,
    trust_shrink=0.5,
    trust_expand=1.5,
    trust_eta_shrink=0.05,
    trust_eta_expand=0.75,
    trust_max_reject=6,
    step_time_limit_s=None,
    iteration_callback=None,
    first_direction_fn=None,
    device_linesearch=None,
):
    """Newton's method for energy minimisation; returns the reference's result dict.

    ``first_direction_fn(x, rhs, sol) -> ksp_its`` (optional, not in the reference) replaces
    ``hessian_solve_fn`` on the first iteration only; it carries the stated first-step policy
    for starts where the exact Hessian vanishes (p-Laplace, p > 2, from u = 0).
    ``device_linesearch(x_prev, p, lo, hi, tol) -> (alpha, J(x_prev + alpha p), n_evals)``
    (optional) runs the golden-section search on the device with the same points and
    comparisons as ``golden_section_search``; used only without ``project_fn`` /
    ``ghost_update_fn``.
    """
    trust_shrink, trust_expand, trust_eta_shrink, trust_eta_expand, trust_max_reject = _sanitize_trust(
        trust_shrink, trust_expand, trust_eta_shrink, trust_eta_expand, trust_max_reject
    )
    ghost = ghost_update_fn or (lambda _v: None)
    project = project_fn or (lambda _v: None)
    if hessian_matvec_fn is None:
        def hessian_matvec_fn(_x, _vin, vout):
            vout.set(0.0)
    rank = 0 if comm is None else comm.Get_rank()

    g = x.duplicate()
    h = x.duplicate()
    g_trial = x.duplicate()
    x_trial = x.duplicate()
    x_prev = x.duplicate()
    p = x.duplicate()

    start = time.perf_counter()
    fx = energy_fn(x)
    nit = 0
    message = "Maximum number of iterations reached"
    history = []
    initial_grad_norm = None
    trust_radius = float(trust_radius_init)
    if trust_region:
        trust_radius = min(max(trust_radius, trust_radius_min), trust_radius_max)

    for _ in range(maxit):
        t_iter_start = time.perf_counter()
        if fail_on_nonfinite and not np.isfinite(fx):
            message = f"Non-finite energy before Newton iteration {nit + 1}"
            break
        t0 = time.perf_counter()
        gradient_fn(x, g)
        normg = g.norm(NORM_2)
        t_grad = time.perf_counter() - t0
        if fail_on_nonfinite and not np.isfinite(normg):
            message = f"Non-finite gradient norm at Newton iteration {nit + 1}"
            break
        if initial_grad_norm is None:
            initial_grad_norm = normg
        grad_target = tolg
        if tolg_rel > 0.0 and np.isfinite(initial_grad_norm):
            grad_target = max(tolg, tolg_rel * initial_grad_norm)
        if verbose and rank == 0:
            print(f"it={nit + 1}: gradient ready, ||g||={normg:.5e}, target={grad_target:.5e}", flush=True)
        if (not require_all_convergence) and normg < grad_target:
            message = "Gradient norm converged"
            break
        nit += 1

        trust_ksp_active = bool(trust_region and trust_subproblem_solve_fn is not None)
        ksp_its = 0
        hnorm = np.nan
        t_hess = 0.0
        if not trust_ksp_active:
            t0 = time.perf_counter()
            g.copy(p)
            p.scale(-1.0)
            solver = first_direction_fn if (first_direction_fn is not None and nit == 1) else hessian_solve_fn
            ksp_its = solver(x, p, h)
            ghost(h)
            hnorm = h.norm(NORM_2)
            t_hess = time.perf_counter() - t0
            if verbose and rank == 0:
                print(f"it={nit}: linear step ready, ksp_its={ksp_its}, ||p||={hnorm:.5e}", flush=True)
            if fail_on_nonfinite and not np.isfinite(hnorm):
                message = f"Non-finite Newton direction norm at Newton iteration {nit}"
                break

        x.copy(x_prev)
        x_prev_norm = x_prev.norm(NORM_2)
        fx_old = fx
        it = {
            "alpha": 0.0, "ls_evals": 0, "ls_repaired": False, "dE": 0.0, "rho": np.nan, "step_norm": 0.0,
            "step_rel": 0.0, "trust_qp": np.nan, "pred": np.nan, "actual": np.nan, "rejects": 0,
            "accepted": False, "grad_fallback": False, "terminate": False, "term_msg": None, "t_update": 0.0,
        }
        t_ls_start = time.perf_counter()
        ls_a, ls_b = linesearch_interval
        if ls_b <= ls_a:
            ls_a, ls_b = -0.5, 2.0

        def energy_at(alpha_local):
            x_trial.waxpy(alpha_local, p, x_prev)
            project(x_trial)
            ghost(x_trial)
            val = energy_fn(x_trial)
            return val if np.isfinite(val) else np.inf

        def bounded_armijo(alpha_lo, alpha_hi, dirderiv):
            alpha_lo, alpha_hi = float(alpha_lo), float(alpha_hi)
            if not np.isfinite(alpha_lo) or not np.isfinite(alpha_hi) or alpha_hi <= max(alpha_lo, 0.0) + 1e-14:
                return 0.0, np.inf, 0, False
            a_t = min(max(float(armijo_alpha0), max(alpha_lo, 1e-12)), alpha_hi)
            n_eval = 0
            for _ls in range(max(1, int(armijo_max_ls))):
                val = energy_at(a_t)
                n_eval += 1
                if np.isfinite(val) and val <= fx_old + armijo_c1 * a_t * dirderiv:
                    return float(a_t), float(val), n_eval, True
                nxt = a_t * float(armijo_shrink)
                floor_a = max(alpha_lo, 1e-12)
                if nxt <= floor_a + 1e-16:
                    a_t = floor_a
                    if a_t <= floor_a + 1e-16:
                        break
                else:
                    a_t = nxt
            return 0.0, np.inf, n_eval, False

        def accept(newval, actual, step_size):
            nonlocal fx
            fx = newval
            it["dE"] = actual
            it["step_norm"] = step_size
            it["step_rel"] = step_size / max(1.0, x_prev_norm)
            t_up = time.perf_counter()
            x_trial.copy(x)
            project(x)
            ghost(x)
            it["t_update"] = time.perf_counter() - t_up
            it["accepted"] = True

        def trust_exhausted(step_size):
            x_prev.copy(x)
            project(x)
            ghost(x)
            it.update(alpha=0.0, dE=0.0, step_norm=0.0, step_rel=0.0, t_update=0.0, terminate=True)
            step_tol = max(tolx_abs, tolx_rel * max(1.0, x_prev_norm))
            small_step = np.isfinite(step_size) and step_size <= step_tol
            small_pred = np.isfinite(it["pred"]) and abs(it["pred"]) <= max(tolf, 1e-16)
            if small_step or small_pred:
                if require_all_convergence and normg >= grad_target:
                    it["term_msg"] = "Trust-region radius exhausted before full convergence"
                else:
                    it["term_msg"] = "Trust-region step converged"
            else:
                it["term_msg"] = f"Trust-region rejected all candidate steps at Newton iteration {nit}"

        def trust_update(trial_radius, newval, snod_norm):
            """rho test; returns True when the step was accepted."""
            nonlocal trust_radius
            rho = np.nan
            if np.isfinite(it["pred"]) and it["pred"] > 0.0 and np.isfinite(it["actual"]):
                rho = it["actual"] / it["pred"]
            it["rho"] = rho
            step_size = abs(it["alpha"]) * snod_norm if np.isfinite(snod_norm) else 0.0
            if np.isfinite(newval) and np.isfinite(rho) and rho >= trust_eta_shrink and it["actual"] > 0.0:
                accept(newval, it["actual"], step_size)
                if rho >= trust_eta_expand and it["step_norm"] >= 0.9 * trial_radius:
                    trust_radius = min(trust_radius_max, trial_radius * trust_expand)
                else:
                    trust_radius = trial_radius
                return True, step_size
            trust_radius = max(trust_radius_min, trial_radius * trust_shrink)
            return False, step_size

        def golden_on(alpha_lo, alpha_hi):
            alpha_lo, alpha_hi, n_rep, rep = repair_linesearch_interval(
                energy_at, alpha_lo, alpha_hi, center=0.0, center_value=fx_old, tol=linesearch_tol
            )
            it["ls_evals"] += n_rep
            it["ls_repaired"] = rep
            if (device_linesearch is not None and project_fn is None and ghost_update_fn is None
                    and np.isfinite(alpha_lo) and np.isfinite(alpha_hi) and linesearch_tol > 0.0):
                alpha, val, n_ev = device_linesearch(x_prev, p, alpha_lo, alpha_hi, linesearch_tol)
                it["ls_evals"] += n_ev
                x_trial.waxpy(alpha, p, x_prev)  # the accepted point, as energy_at(alpha) leaves it
                return alpha, val if np.isfinite(val) else np.inf
            alpha, n_ev = golden_section_search(energy_at, alpha_lo, alpha_hi, linesearch_tol)
            it["ls_evals"] += n_ev
            return alpha, energy_at(alpha)

        if trust_region:
            trust_radius = min(max(trust_radius, trust_radius_min), trust_radius_max)
            max_attempts = trust_max_reject + 1
            if trust_ksp_active:
                for attempt in range(max_attempts):
                    trial_radius = float(min(max(trust_radius, trust_radius_min), trust_radius_max))
                    t_h0 = time.perf_counter()
                    g.copy(h)
                    h.scale(-1.0)
                    ksp_its = trust_subproblem_solve_fn(x, h, p, trial_radius)
                    ghost(p)
                    snod_norm = p.norm(NORM_2)
                    t_hess += time.perf_counter() - t_h0
                    if fail_on_nonfinite and not np.isfinite(snod_norm):
                        message = f"Non-finite trust-region step norm at Newton iteration {nit}"
                        it["terminate"] = True
                        it["term_msg"] = message
                        break
                    it["grad_fallback"] = False
                    if np.isfinite(snod_norm) and snod_norm > 1e-20:
                        hessian_matvec_fn(x, p, h)
                        ghost(h)
                        lin = float(g.dot(p))
                        curv = float(p.dot(h))
                        it["trust_qp"] = float(lin + 0.5 * curv)
                        if trust_subproblem_line_search:
                            cap = trial_radius / snod_norm
                            a_lo, a_hi = max(float(ls_a), -cap), min(float(ls_b), cap)
                            if np.isfinite(a_lo) and np.isfinite(a_hi) and a_hi > a_lo + 1e-14:
                                if str(line_search) == "armijo":
                                    alpha, newval, n_ev, ok = bounded_armijo(a_lo, a_hi, lin)
                                    it["ls_evals"] += n_ev
                                    if ok:
                                        it["pred"] = -(alpha * lin + 0.5 * alpha**2 * curv)
                                        it["actual"] = fx_old - newval if np.isfinite(newval) else -np.inf
                                    else:
                                        alpha, newval, it["pred"], it["actual"] = 0.0, np.inf, 0.0, -np.inf
                                else:
                                    alpha, newval = golden_on(a_lo, a_hi)
                                    it["pred"] = -(alpha * lin + 0.5 * alpha**2 * curv)
                                    it["actual"] = fx_old - newval if np.isfinite(newval) else -np.inf
                            else:
                                alpha, newval, it["pred"], it["actual"] = 0.0, np.inf, 0.0, -np.inf
                        else:
                            alpha = 1.0
                            it["pred"] = -(lin + 0.5 * curv)
                            newval = energy_at(1.0)
                            it["actual"] = fx_old - newval if np.isfinite(newval) else -np.inf
                    else:
                        it["trust_qp"] = 0.0
                        alpha, newval, it["pred"], it["actual"] = 0.0, np.inf, 0.0, -np.inf
                    it["alpha"] = alpha
                    ok, step_size = trust_update(trial_radius, newval, snod_norm)
                    if ok:
                        break
                    it["rejects"] = attempt + 1
                    if attempt + 1 >= max_attempts:
                        trust_exhausted(step_size)
                        break
            else:
                z1 = x.duplicate()
                z2 = x.duplicate()
                hz1 = x.duplicate()
                hz2 = x.duplicate()
                sg = x.duplicate()
                has_z2 = False
                rhs_red = mm = None
                rhs_g = curv_g = np.nan
                if np.isfinite(hnorm) and hnorm > 1e-20:
                    h.copy(z1)
                    z1.scale(1.0 / hnorm)
                elif normg > 1e-20:
                    g.copy(z1)
                    z1.scale(-1.0 / normg)
                else:
                    z1.set(0.0)
                if normg > 1e-20:
                    g.copy(z2)
                    z2.scale(1.0 / normg)
                    z2.axpy(-z1.dot(z2), z1)
                    z2n = z2.norm(NORM_2)
                    if z2n > 1e-12:
                        z2.scale(1.0 / z2n)
                        has_z2 = True
                hessian_matvec_fn(x, z1, hz1)
                ghost(hz1)
                if has_z2:
                    hessian_matvec_fn(x, z2, hz2)
                    ghost(hz2)
                    rhs_red = np.array([g.dot(z1), g.dot(z2)])
                    mm = np.array([[z1.dot(hz1), z1.dot(hz2)], [z2.dot(hz1), z2.dot(hz2)]])
                rhs1 = float(g.dot(z1))
                curv1 = float(z1.dot(hz1))
                if normg > 1e-20:
                    g.copy(sg)
                    sg.scale(1.0 / normg)
                    hessian_matvec_fn(x, sg, hz2)
                    ghost(hz2)
                    rhs_g = float(g.dot(sg))
                    curv_g = float(sg.dot(hz2))

                def build_step(delta):
                    p.set(0.0)
                    used_grad = False
                    if has_z2:
                        st, q2d = solve_trust_2d(rhs_red, mm, delta)
                        p.axpy(float(st[0]), z1)
                        p.axpy(float(st[1]), z2)
                        snorm, lin, curv, q = float(np.linalg.norm(st)), float(rhs_red.dot(st)), float(st.dot(mm.dot(st))), float(q2d)
                    else:
                        t1 = solve_trust_1d(rhs1, curv1, delta)
                        z1.copy(p)
                        p.scale(t1)
                        snorm, lin, curv = abs(t1), float(rhs1 * t1), float(curv1 * t1 * t1)
                        q = float(lin + 0.5 * curv)
                    if normg > 1e-20:
                        tg = solve_trust_1d(rhs_g, curv_g, delta)
                        qg = float(rhs_g * tg + 0.5 * curv_g * tg * tg)
                        if qg <= q:
                            g.copy(sg)
                            sg.scale(tg / normg)
                            sg.copy(p)
                            snorm, lin, curv, q, used_grad = abs(tg), float(rhs_g * tg), float(curv_g * tg * tg), qg, True
                    return snorm, q, lin, curv, used_grad

                for attempt in range(max_attempts):
                    trial_radius = float(min(max(trust_radius, trust_radius_min), trust_radius_max))
                    snod_norm, it["trust_qp"], lin, curv, it["grad_fallback"] = build_step(trial_radius)
                    a_lo, a_hi = float(ls_a), float(ls_b)
                    if snod_norm > 1e-20:
                        cap = trial_radius / snod_norm
                        a_lo, a_hi = max(a_lo, -cap), min(a_hi, cap)
                    if (
                        not np.isfinite(snod_norm) or snod_norm <= 1e-20 or not np.isfinite(a_lo)
                        or not np.isfinite(a_hi) or a_hi <= a_lo + 1e-14
                    ):
                        newval, alpha, it["pred"], it["actual"] = np.inf, 0.0, 0.0, -np.inf
                    else:
                        alpha, newval = golden_on(a_lo, a_hi)
                        it["pred"] = -(alpha * lin + 0.5 * alpha**2 * curv)
                        it["actual"] = fx_old - newval if np.isfinite(newval) else -np.inf
                    it["alpha"] = alpha
                    ok, step_size = trust_update(trial_radius, newval, snod_norm)
                    if ok:
                        break
                    it["rejects"] = attempt + 1
                    if attempt + 1 >= max_attempts:
                        trust_exhausted(step_size)
                        break
                for v in (z1, z2, hz1, hz2, sg):
                    v.destroy()
        else:
            h.copy(p)
            pnorm = p.norm(NORM_2)
            if pnorm > 1e-20:
                mode = str(line_search)
                if mode == "armijo":
                    dd = float(g.dot(p))
                    alpha, fx_trial, n_ev, ok = bounded_armijo(ls_a, ls_b, dd)
                    it["ls_evals"] += n_ev
                    if (not ok) and bool(armijo_gradient_fallback):
                        g.copy(p)
                        p.scale(-1.0)
                        pnorm = p.norm(NORM_2)
                        alpha, fx_trial, n_ev, ok = bounded_armijo(ls_a, ls_b, -float(normg * normg))
                        it["ls_evals"] += n_ev
                        it["grad_fallback"] = bool(ok)
                elif mode in {"residual_bisection", "residual_bisection_tol"}:
                    dd = float(g.dot(p))
                    if not np.isfinite(dd) or dd >= 0.0:
                        alpha, fx_trial = 0.0, np.inf
                    else:
                        a_min, a_max, alpha, acc_a = 0.0, 1.0, 1.0, 0.0
                        btol = max(float(linesearch_tol), 1e-12)
                        for _ls in range(max(1, int(armijo_max_ls))):
                            x_trial.waxpy(float(alpha), p, x_prev)
                            project(x_trial)
                            ghost(x_trial)
                            gradient_fn(x_trial, g_trial)
                            ghost(g_trial)
                            it["ls_evals"] += 1
                            dtr = float(g_trial.dot(p))
                            if np.isfinite(dtr) and dtr < 0.0:
                                acc_a = float(alpha)
                                if abs(alpha - 1.0) <= 1e-15:
                                    break
                                a_min = float(alpha)
                            else:
                                a_max = float(alpha)
                            if mode == "residual_bisection_tol" and acc_a > 0.0 and (a_max - a_min) <= btol:
                                break
                            nxt = 0.5 * (a_min + a_max)
                            if nxt <= a_min + 1e-16:
                                break
                            alpha = float(nxt)
                        if acc_a > 0.0:
                            alpha = acc_a
                            x_trial.waxpy(alpha, p, x_prev)
                            project(x_trial)
                            ghost(x_trial)
                            fx_trial = energy_fn(x_trial)
                        else:
                            alpha, fx_trial = 0.0, np.inf
                else:
                    alpha, fx_trial = golden_on(ls_a, ls_b)
                if verbose and rank == 0:
                    print(f"it={nit}: line search done, alpha={alpha:.5e}, ls_evals={it['ls_evals']}", flush=True)
                it["alpha"] = alpha
                if (mode == "residual_bisection" and alpha > 0.0 and np.isfinite(fx_trial)) or (
                    np.isfinite(fx_trial) and fx_trial < fx_old
                ):
                    accept(fx_trial, fx_old - fx_trial, abs(alpha) * pnorm)
                else:
                    x_prev.copy(x)
                    project(x)
                    ghost(x)
                    it["alpha"] = 0.0
            else:
                x_prev.copy(x)
                project(x)
                ghost(x)
                it["alpha"] = 0.0

        t_ls = time.perf_counter() - t_ls_start
        t_iter = time.perf_counter() - t_iter_start
        rt =float(normg / max(abs(float(grad_target)), 1e-30))
        if verbose and rank == 0:
            print(
                f"it={nit}, J={fx:.10f}, dJ={it['dE']:.5e}, ||g||={normg:.5e}, alpha={it['alpha']:.5e}, "
                f"ksp_its={ksp_its}, ls_evals={it['ls_evals']}, accepted={it['accepted']}, t_it={t_iter:.3f}s",
                flush=True,
            )
        post_grad_norm = np.nan
        converged_energy = it["accepted"] and np.isfinite(it["dE"]) and abs(it["dE"]) < tolf
        converged_step = it["accepted"] and (it["step_norm"] < tolx_abs or it["step_rel"] < tolx_rel)
        if require_all_convergence and converged_energy and converged_step:
            gradient_fn(x, g)
            post_grad_norm = g.norm(NORM_2)
            if fail_on_nonfinite and not np.isfinite(post_grad_norm):
                x_prev.copy(x)
                ghost(x)
                message = f"Non-finite post-update gradient norm at Newton iteration {nit}"
                break
        entry = {
            "it": int(nit), "energy": float(fx), "dE": float(it["dE"]), "grad_norm": float(normg),
            "grad_target": float(grad_target), "rt": rt, "grad_norm_post": float(post_grad_norm),
            "alpha": float(it["alpha"]), "ksp_its": int(ksp_its), "ls_evals": int(it["ls_evals"]),
            "ls_repaired": bool(it["ls_repaired"]), "used_gradient_fallback": bool(it["grad_fallback"]),
            "accepted_step": bool(it["accepted"]), "step_norm": float(it["step_norm"]),
            "step_rel": float(it["step_rel"]), "pred_reduction": float(it["pred"]),
            "actual_reduction": float(it["actual"]), "trust_rejects": int(it["rejects"]),
            "t_grad": float(t_grad), "t_hess": float(t_hess), "t_ls": float(t_ls),
            "t_update": float(it["t_update"]), "t_iter": float(t_iter),
            "runtime_s": float(time.perf_counter() - start), "trust_radius": float(trust_radius),
            "trust_ratio": float(it["rho"]), "trust_qp": float(it["trust_qp"]), "line_search": str(line_search),
        }
        if save_history:
            history.append(entry)
        if iteration_callback is not None:
            iteration_callback(dict(entry), list(history))
        if step_time_limit_s is not None and np.isfinite(step_time_limit_s) and (
            time.perf_counter() - start
        ) > float(step_time_limit_s):
            message = (
                f"Step time limit exceeded ({time.perf_counter() - start:.3f}s > {float(step_time_limit_s):.3f}s)"
            )
            break
        if it["terminate"]:
            message = it["term_msg"]
            break
        if fail_on_nonfinite and not np.isfinite(fx):
            x_prev.copy(x)
            ghost(x)
            message = f"Non-finite energy after update at Newton iteration {nit}"
            break
        if require_all_convergence:
            if converged_energy and converged_step and post_grad_norm < grad_target:
                message = "Converged (energy, step, gradient)"
                break
        elif converged_energy:
            message = "Energy change converged"
            break

    runtime = time.perf_counter() - start
    for v in (g, h, g_trial, x_trial, x_prev, p):
        v.destroy()
    return {
        "x": x,
        "fun": fx,
        "nit": nit,
        "time": runtime,
        "message": message,
        "history": history,
        "success": "converged" in message.lower(),
    }
