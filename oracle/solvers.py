"""Oracle Krylov solvers and Newton loop (numpy, FP64) -- TEST INFRASTRUCTURE ONLY.

* ``cg``: PETSc KSPCG semantics (zero initial guess, left Jacobi preconditioning with zero
  diagonal entries replaced by 1, convergence on the preconditioned residual norm
  ||B r_k|| <= rtol ||B r_0||, indefinite-curvature stop, iteration cap).
* ``gmres``: PETSc KSPGMRES(restart) with left preconditioning and its default
  orthogonalization, classical Gram-Schmidt without refinement
  (KSPGMRESClassicalGramSchmidtOrthogonalization, KSP_GMRES_CGS_REFINE_NEVER).
* ``fgmres``: PETSc KSPFGMRES(restart), right preconditioning by any (possibly varying)
  callable, the same orthogonalization, convergence on the unpreconditioned residual norm.
* ``stcg``: PETSc KSPSTCG, the trust-region (Steihaug-Toint) CG of the trust-region Newton.
* ``newton_golden``: the line-search branch of core/petsc/minimizers.py:181-1038
  (golden-section search on [a, b] after non-finite interval repair, convergence on
  |dE| < tolf and small step, then a gradient check when require_all_convergence).
* ``newton_trust``: the trust-region branch of the same function (:405-790): radius-limited
  subproblem (a trust KSP such as ``stcg`` or the 1-D / 2-D model of :570-700), optional
  golden-section search inside the radius, the rho test with shrink / expand and the stop
  after trust_max_reject + 1 rejected trials.
* ``trust_1d`` / ``trust_2d``: the model subproblems of :129-178.

Pins (tests/test_oracle_solvers.py): Krylov solvers by exact solves and termination counts,
``fgmres`` also by its residual history against brute-force least squares over the
right-preconditioned Krylov space x in M K_k(A M, b) (restart larger than the run);
``newton_golden`` by a quadratic (one exact Newton step, convergence at the second
iteration), the quartic x^4/4 whose line search is clipped at alpha = 2 (x_k = x_0 / 3^k) and
the tolg / tolg_rel targets; ``newton_trust`` by the isotropic quadratic whose boundary steps
grow by 1.5 per iteration (closed-form iteration count), a model that overshoots (radius
halves, then the reference's exhaustion messages) and agreement of its model branch with the
closed-form 1-D / 2-D subproblems.
"""

from __future__ import annotations

import math

import numpy as np


def jacobi_inverse(A) -> np.ndarray:
    d = A.diagonal().astype(np.float64)
    d[d == 0.0] = 1.0
    return 1.0 / d


def cg(A, b, rtol=1e-3, max_it=10000, atol=1e-50, dtol=1e5, inv_diag=None):
    """Returns (x, its, reason)."""
    inv_diag = jacobi_inverse(A) if inv_diag is None else inv_diag
    x = np.zeros_like(b)
    r = b.copy()
    z = inv_diag * r
    dp = math.sqrt(float(z @ z))
    r0 = dp
    if dp <= max(rtol * r0, atol):
        return x, 0, 2
    beta = float(z @ r)
    p = z.copy()
    dpiold = None
    betaold = beta
    for i in range(max_it):
        if i > 0:
            p = z + (beta / betaold) * p
        w = A @ p
        dpi = float(p @ w)
        if dpi == 0.0 or (dpiold is not None and dpi * dpiold <= 0.0) or not math.isfinite(dpi):
            return x, i, -10
        dpiold = dpi
        a = beta / dpi
        x += a * p
        r -= a * w
        z = inv_diag * r
        dp = math.sqrt(float(z @ z))
        if dp <= max(rtol * r0, atol):
            return x, i + 1, 2
        if dp >= dtol * r0:
            return x, i + 1, -4
        betaold = beta
        beta = float(z @ r)
    return x, max_it, -3


def gmres(A, b, rtol=1e-3, max_it=10000, restart=30, atol=1e-50, inv_diag=None):
    inv_diag = jacobi_inverse(A) if inv_diag is None else inv_diag
    n = b.size
    x = np.zeros(n)
    its = 0
    r0 = None
    while True:
        r = inv_diag * (b - A @ x) if its else inv_diag * b
        beta = math.sqrt(float(r @ r))
        if r0 is None:
            r0 = beta
            if beta <= max(rtol * r0, atol):
                return x, 0, 2
        if beta == 0.0:
            return x, its, 3
        m = restart
        V = np.zeros((m + 1, n))
        V[0] = r / beta
        H = np.zeros((m + 1, m))
        g = np.zeros(m + 1)
        g[0] = beta
        cs = np.zeros(m)
        sn = np.zeros(m)
        k_used = 0
        done = 0
        for k in range(m):
            w = inv_diag * (A @ V[k])
            H[: k + 1, k] = V[: k + 1] @ w  # classical Gram-Schmidt
            w = w - H[: k + 1, k] @ V[: k + 1]
            hn = math.sqrt(float(w @ w))
            H[k + 1, k] = hn
            if hn > 0.0:
                V[k + 1] = w / hn
            for j in range(k):
                t0 = cs[j] * H[j, k] + sn[j] * H[j + 1, k]
                H[j + 1, k] = -sn[j] * H[j, k] + cs[j] * H[j + 1, k]
                H[j, k] = t0
            den = math.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = (1.0, 0.0) if den == 0.0 else (H[k, k] / den, H[k + 1, k] / den)
            H[k, k] = cs[k] * H[k, k] + sn[k] * H[k + 1, k]
            H[k + 1, k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            k_used = k + 1
            its += 1
            res = abs(g[k + 1])
            if res <= max(rtol * r0, atol):
                done = 2
            if done or its >= max_it or hn == 0.0:
                break
        y = np.zeros(k_used)
        for i in range(k_used - 1, -1, -1):
            s = g[i] - float(H[i, i + 1 : k_used] @ y[i + 1 : k_used])
            y[i] = s / H[i, i] if H[i, i] != 0.0 else 0.0
        x += y @ V[:k_used]
        if done:
            return x, its, done
        if its >= max_it:
            return x, its, -3


def fgmres(A, b, precond, rtol=1e-3, max_it=10000, restart=30, atol=1e-50, history=None):
    """KSPFGMRES(restart): right preconditioning z_k = precond(v_k) (any callable, it may vary),
    x = x0 + Z y, classical Gram-Schmidt, convergence on ||b - A x|| (the unpreconditioned norm
    PETSc tests under right preconditioning).  ``history`` (a list) receives the Givens
    residual estimate after each iteration.  Returns (x, its, reason)."""
    n = b.size
    x = np.zeros(n)
    its = 0
    r0 = None
    while True:
        r = b - A @ x if its else b.copy()
        beta = math.sqrt(float(r @ r))
        if r0 is None:
            r0 = beta
            if beta <= max(rtol * r0, atol):
                return x, 0, 2
        if beta == 0.0:
            return x, its, 3
        m = restart
        V = np.zeros((m + 1, n))
        Z = np.zeros((m, n))
        V[0] = r / beta
        H = np.zeros((m + 1, m))
        g = np.zeros(m + 1)
        g[0] = beta
        cs = np.zeros(m)
        sn = np.zeros(m)
        k_used = 0
        done = 0
        for k in range(m):
            Z[k] = precond(V[k])
            w = A @ Z[k]
            H[: k + 1, k] = V[: k + 1] @ w  # classical Gram-Schmidt
            w = w - H[: k + 1, k] @ V[: k + 1]
            hn = math.sqrt(float(w @ w))
            H[k + 1, k] = hn
            if hn > 0.0:
                V[k + 1] = w / hn
            for j in range(k):
                t0 = cs[j] * H[j, k] + sn[j] * H[j + 1, k]
                H[j + 1, k] = -sn[j] * H[j, k] + cs[j] * H[j + 1, k]
                H[j, k] = t0
            den = math.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = (1.0, 0.0) if den == 0.0 else (H[k, k] / den, H[k + 1, k] / den)
            H[k, k] = cs[k] * H[k, k] + sn[k] * H[k + 1, k]
            H[k + 1, k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            k_used = k + 1
            its += 1
            res = abs(g[k + 1])
            if history is not None:
                history.append(res)
            if res <= max(rtol * r0, atol):
                done = 2
            if done or its >= max_it or hn == 0.0:
                break
        y = np.zeros(k_used)
        for i in range(k_used - 1, -1, -1):
            s = g[i] - float(H[i, i + 1 : k_used] @ y[i + 1 : k_used])
            y[i] = s / H[i, i] if H[i, i] != 0.0 else 0.0
        x += y @ Z[:k_used]
        if done:
            return x, its, done
        if its >= max_it:
            return x, its, -3


def stcg(A, b, radius=math.inf, rtol=1e-3, max_it=10000, atol=1e-50, dtol=1e5, inv_diag=None):
    """Steihaug-Toint truncated CG: approximately minimise q(s) = -b.s + s.A s / 2 over ||s|| <= radius.

    The trust-region Krylov solve the reference requests with ksp_type "stcg" and a radius set
    through KSPCGSetRadius (core/petsc/trust_ksp.py:55-72, core/petsc/scalar_problem_driver.py:
    143-170).  The algorithm is PETSc's KSPSTCG (PETSc 3.2x, absent here; Steihaug 1983, Toint
    1981), which the reference restates for its serial path in
    core/serial/sparse_solvers.py:20-93: preconditioned CG from s = 0 with the constraint on the
    Euclidean norm of s; on curvature p.Ap <= 0 (KSP_CONVERGED_CG_NEG_CURVE, 5) or on a step
    that would leave the ball (KSP_CONVERGED_STEP_LENGTH, 6) the iterate moves along p to the
    boundary and stops; convergence on the unpreconditioned residual norm
    ||r_k|| <= max(rtol ||b||, atol) (the serial restatement's test and PETSc STCG's default
    norm), divergence at dtol ||b||.  With an infinite radius a negative curvature stop keeps
    the current iterate.  Returns (s, its, reason).
    """
    inv_diag = jacobi_inverse(A) if inv_diag is None else inv_diag
    delta = float(radius)
    x = np.zeros_like(b)
    r = b.copy()
    z = inv_diag * r
    r0 = math.sqrt(float(r @ r))
    if r0 <= max(rtol * r0, atol):
        return x, 0, 2 if r0 > atol else 3
    rz = float(r @ z)
    if rz < 0.0:
        return x, 0, -8
    p = z.copy()
    xx = 0.0
    for i in range(max_it):
        w = A @ p
        kappa = float(p @ w)
        if not math.isfinite(kappa):
            return x, i + 1, -9
        xp, pp = float(x @ p), float(p @ p)

        def to_boundary():
            # the positive root tau of ||x + tau p||^2 = delta^2
            return (-xp + math.sqrt(max(xp * xp + pp * (delta * delta - xx), 0.0))) / pp

        if kappa <= 0.0:
            if math.isfinite(delta):
                x = x + to_boundary() * p
            return x, i + 1, 5
        alpha = rz / kappa
        if xx + 2.0 * alpha * xp + alpha * alpha * pp >= delta * delta:
            return x + to_boundary() * p, i + 1, 6
        x = x + alpha * p
        xx = float(x @ x)
        r = r - alpha * w
        z = inv_diag * r
        dp = math.sqrt(float(r @ r))
        if dp <= max(rtol * r0, atol):
            return x, i + 1, 2 if dp > atol else 3
        if dp >= dtol * r0:
            return x, i + 1, -4
        rz_new = float(r @ z)
        if rz_new < 0.0:
            return x, i + 1, -8
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, max_it, -3


def golden_section(f, a, b, tol):
    """core/petsc/minimizers.py:14-62."""
    gam = 0.5 + math.sqrt(5.0) / 2.0
    an, bn = float(a), float(b)
    fa, fb = f(an), f(bn)
    dn = (bn - an) / gam + an
    cn = an + bn - dn
    fc, fd = f(cn), f(dn)
    n = 4
    best, bestv = an, fa
    for al, v in ((bn, fb), (cn, fc), (dn, fd)):
        if v < bestv:
            best, bestv = al, v
    while bn - an > tol:
        if fc < fd:
            bn = dn
            dn, cn = cn, an + bn - cn
            fd, fc = fc, f(cn)
            cand = ((dn, fd), (cn, fc))
        else:
            an = cn
            cn, dn = dn, an + bn - dn
            fc, fd = fd, f(dn)
            cand = ((cn, fc), (dn, fd))
        for al, v in cand:
            if v < bestv:
                best, bestv = al, v
        n += 1
    mid = 0.5 * (an + bn)
    n += 1
    if f(mid) < bestv:
        best = mid
    return best, n


def newton_golden(energy, gradient, solve, x0, *, tolf=1e-5, tolg=1e-3, tolx_rel=1e-6, tolx_abs=1e-10,
                  maxit=100, ls=(-0.5, 2.0), ls_tol=1e-3, require_all=True, first_solve=None, tolg_rel=0.0):
    """Line-search Newton (golden section).  solve(x, rhs) -> (direction, its).

    Non-finite energies along the line count as +inf; interval repair bisects each non-finite
    end toward alpha = 0 (core/petsc/minimizers.py:74-126).
    """
    x = np.array(x0, dtype=np.float64)
    fx = energy(x)
    nit = 0
    msg = "Maximum number of iterations reached"
    lin = []
    for _ in range(maxit):
        if not np.isfinite(fx):
            msg = f"Non-finite energy before Newton iteration {nit + 1}"
            break
        g = gradient(x)
        ng = float(np.linalg.norm(g))
        g0 = ng if nit == 0 else g0
        target = max(tolg, tolg_rel * g0) if (tolg_rel > 0.0 and np.isfinite(g0)) else tolg
        if not require_all and ng < target:
            msg = "Gradient norm converged"
            break
        nit += 1
        s = first_solve if (first_solve is not None and nit == 1) else solve
        h, its = s(x, -g)
        lin.append(its)
        x_prev = x.copy()
        xn = float(np.linalg.norm(x_prev))
        fold = fx

        def f_at(al):
            v = energy(x_prev + al * h)
            return v if np.isfinite(v) else np.inf

        pn = float(np.linalg.norm(h))
        accepted, dE, step = False, 0.0, 0.0
        if pn > 1e-20:
            a, b = ls

            def repair(bound, sign):
                if (sign < 0 and not bound < 0.0) or (sign > 0 and not bound > 0.0):
                    return bound
                if np.isfinite(f_at(bound)):
                    return bound
                good, bad = 0.0, bound
                for _ in range(60):
                    mid = 0.5 * (good + bad)
                    if np.isfinite(f_at(mid)):
                        good = mid
                    else:
                        bad = mid
                    if abs(bad - good) <= max(ls_tol, 1e-12):
                        break
                return good

            if np.isfinite(fold):
                a, b = repair(a, -1.0), repair(b, 1.0)
            al, _ = golden_section(f_at, a, b, ls_tol)
            ft = f_at(al)
            if np.isfinite(ft) and ft < fold:
                x = x_prev + al * h
                fx = ft
                dE = fold - ft
                step = abs(al) * pn
                accepted = True
            else:
                x = x_prev
        conv_e = accepted and abs(dE) < tolf
        conv_s = accepted and (step < tolx_abs or step / max(1.0, xn) < tolx_rel)
        post = np.nan
        if require_all and conv_e and conv_s:
            post = float(np.linalg.norm(gradient(x)))
        if not np.isfinite(fx):
            msg = f"Non-finite energy after update at Newton iteration {nit}"
            x = x_prev
            break
        if require_all:
            if conv_e and conv_s and post < target:
                msg = "Converged (energy, step, gradient)"
                break
        elif conv_e:
            msg = "Energy change converged"
            break
    return {"x": x, "fun": fx, "nit": nit, "message": msg, "linear_iters": lin}


def trust_1d(rhs, curv, delta):
    """argmin rhs t + curv t^2 / 2 over |t| <= delta (core/petsc/minimizers.py:129-143)."""
    delta = max(0.0, float(delta))
    if delta == 0.0:
        return 0.0
    if math.isfinite(curv) and curv > 1e-14:
        t = -rhs / curv
    else:
        t = -delta if rhs > 0.0 else (delta if rhs < 0.0 else 0.0)
    return float(min(max(t, -delta), delta))


def trust_2d(rhs, mm, delta):
    """2-D model: the interior stationary point if it lies in the disc and lowers the model,
    else the best of 181 equally spaced boundary points (core/petsc/minimizers.py:150-178)."""
    rhs = np.asarray(rhs, dtype=np.float64)
    mm = np.asarray(mm, dtype=np.float64)
    q = lambda s: float(rhs @ s + 0.5 * s @ (mm @ s))
    delta = max(0.0, float(delta))
    best = np.zeros(2)
    best_q = q(best)
    if delta == 0.0:
        return best, best_q
    try:
        su = np.linalg.solve(mm, -rhs)
        if np.all(np.isfinite(su)) and np.linalg.norm(su) <= delta and q(su) < best_q:
            best, best_q = su, q(su)
    except np.linalg.LinAlgError:
        pass
    for th in np.arange(181) * (2.0 * np.pi / 181):
        s = delta * np.array([np.cos(th), np.sin(th)])
        if q(s) < best_q:
            best, best_q = s, q(s)
    return best, best_q


def _repair_interval(f, a, b, tol, fcenter):
    """Bisect a non-finite end of [a, b] toward 0 (core/petsc/minimizers.py:74-126)."""
    if not np.isfinite(fcenter):
        return a, b
    tol = max(float(tol), 1e-12)

    def side(bound, sign):
        if (sign < 0 and not bound < 0.0) or (sign > 0 and not bound > 0.0):
            return bound
        if np.isfinite(f(bound)):
            return bound
        good, bad = 0.0, bound
        for _ in range(60):
            mid = 0.5 * (good + bad)
            if np.isfinite(f(mid)):
                good = mid
            else:
                bad = mid
            if abs(bad - good) <= tol:
                break
        return good

    return side(float(a), -1.0), side(float(b), 1.0)


def newton_trust(energy, gradient, hess_mv, x0, *, trust_solve=None, solve=None, tolf=1e-5, tolg=1e-3,
                 tolg_rel=0.0, tolx_rel=1e-6, tolx_abs=1e-10, maxit=100, ls=(-0.5, 2.0), ls_tol=1e-3,
                 line_search=False, radius=1.0, radius_min=1e-8, radius_max=1e6, shrink=0.5, expand=1.5,
                 eta_shrink=0.05, eta_expand=0.75, max_reject=6, require_all=True):
    """Trust-region Newton (core/petsc/minimizers.py:181-1038 with trust_region=True).

    ``trust_solve(x, rhs, radius) -> (step, its)``: the radius-limited linear subproblem (the
    trust KSP, e.g. ``stcg``).  Without it the Newton direction ``solve(x, rhs) -> (h, its)``
    and the gradient span the 2-D model space (or 1-D when they are parallel), and the
    scaled steepest-descent step replaces the model step when its model value is lower.
    ``line_search``: golden-section search on [ls] clipped to the radius (always on in the
    model branch).  ``hess_mv(x, v)`` is the Hessian product at x.
    Returns dict(x, fun, nit, message, linear_iters, radii).
    """
    x = np.array(x0, dtype=np.float64)
    fx = float(energy(x))
    nit = 0
    msg = "Maximum number of iterations reached"
    lin, radii = [], []
    g0 = None
    a0, b0 = (float(ls[0]), float(ls[1])) if ls[1] > ls[0] else (-0.5, 2.0)
    rad = min(max(float(radius), radius_min), radius_max)
    for _ in range(int(maxit)):
        if not np.isfinite(fx):
            msg = f"Non-finite energy before Newton iteration {nit + 1}"
            break
        g = gradient(x)
        ng = float(np.linalg.norm(g))
        if not np.isfinite(ng):
            msg = f"Non-finite gradient norm at Newton iteration {nit + 1}"
            break
        g0 = ng if g0 is None else g0
        target = max(tolg, tolg_rel * g0) if (tolg_rel > 0.0 and np.isfinite(g0)) else tolg
        if not require_all and ng < target:
            msg = "Gradient norm converged"
            break
        nit += 1
        x_prev = x.copy()
        xn = float(np.linalg.norm(x_prev))
        fold = fx
        rad = min(max(rad, radius_min), radius_max)

        def f_at(al, d):
            v = float(energy(x_prev + al * d))
            return v if np.isfinite(v) else np.inf

        def search(d, dnorm, r):
            lo, hi = a0, b0
            cap = r / dnorm
            lo, hi = max(lo, -cap), min(hi, cap)
            if not (np.isfinite(lo) and np.isfinite(hi) and hi > lo + 1e-14):
                return None
            lo, hi = _repair_interval(lambda al: f_at(al, d), lo, hi, ls_tol, fold)
            al, _ = golden_section(lambda t: f_at(t, d), lo, hi, ls_tol)
            return al, f_at(al, d)

        model = None
        if trust_solve is None:
            h, its = solve(x, -g)
            lin.append(its)
            hn = float(np.linalg.norm(h))
            if not np.isfinite(hn):
                msg = f"Non-finite Newton direction norm at Newton iteration {nit}"
                break
            if hn > 1e-20:
                z1 = h / hn
            elif ng > 1e-20:
                z1 = -g / ng
            else:
                z1 = np.zeros_like(g)
            z2 = None
            if ng > 1e-20:
                z2 = g / ng
                z2 = z2 - float(z1 @ z2) * z1
                nz2 = float(np.linalg.norm(z2))
                z2 = z2 / nz2 if nz2 > 1e-12 else None
            hz1 = hess_mv(x, z1)
            if z2 is not None:
                hz2 = hess_mv(x, z2)
                red = np.array([g @ z1, g @ z2])
                mm = np.array([[z1 @ hz1, z1 @ hz2], [z2 @ hz1, z2 @ hz2]])
            sg = g / ng if ng > 1e-20 else None
            if sg is not None:
                rg, cg_ = float(g @ sg), float(sg @ hess_mv(x, sg))

            def model(r):
                if z2 is not None:
                    st, q = trust_2d(red, mm, r)
                    d = st[0] * z1 + st[1] * z2
                    lin_, curv_ = float(red @ st), float(st @ (mm @ st))
                    dn = float(np.linalg.norm(st))
                else:
                    r1, c1 = float(g @ z1), float(z1 @ hz1)
                    t = trust_1d(r1, c1, r)
                    d, dn, lin_, curv_ = t * z1, abs(t), r1 * t, c1 * t * t
                    q = lin_ + 0.5 * curv_
                if sg is not None:
                    tg = trust_1d(rg, cg_, r)
                    if rg * tg + 0.5 * cg_ * tg * tg <= q:
                        d, dn, lin_, curv_ = tg * sg, abs(tg), rg * tg, cg_ * tg * tg
                return d, dn, lin_, curv_

        accepted, dE, step = False, 0.0, 0.0
        terminate = None
        for trial in range(max_reject + 1):
            r = min(max(rad, radius_min), radius_max)
            if model is None:
                rhs = -g
                d, its = trust_solve(x, rhs, r)
                lin.append(its)
                dn = float(np.linalg.norm(d))
                if not np.isfinite(dn):
                    msg = f"Non-finite trust-region step norm at Newton iteration {nit}"
                    terminate = msg
                    break
                if dn > 1e-20:
                    hd = hess_mv(x, d)
                    lin_, curv_ = float(g @ d), float(d @ hd)
                    if line_search:
                        found = search(d, dn, r)
                        if found is None:
                            al, new, pred, act = 0.0, np.inf, 0.0, -np.inf
                        else:
                            al, new = found
                            pred = -(al * lin_ + 0.5 * al * al * curv_)
                            act = fold - new if np.isfinite(new) else -np.inf
                    else:
                        al = 1.0
                        new = f_at(1.0, d)
                        pred = -(lin_ + 0.5 * curv_)
                        act = fold - new if np.isfinite(new) else -np.inf
                else:
                    al, new, pred, act = 0.0, np.inf, 0.0, -np.inf
            else:
                d, dn, lin_, curv_ = model(r)
                found = search(d, dn, r) if (np.isfinite(dn) and dn > 1e-20) else None
                if found is None:
                    al, new, pred, act = 0.0, np.inf, 0.0, -np.inf
                else:
                    al, new = found
                    pred = -(al * lin_ + 0.5 * al * al * curv_)
                    act = fold - new if np.isfinite(new) else -np.inf
            rho = act / pred if (np.isfinite(pred) and pred > 0.0 and np.isfinite(act)) else np.nan
            size = abs(al) * dn if np.isfinite(dn) else 0.0
            if np.isfinite(new) and np.isfinite(rho) and rho >= eta_shrink and act > 0.0:
                x = x_prev + al * d
                fx, dE, step, accepted = new, act, size, True
                rad = min(radius_max, r * expand) if (rho >= eta_expand and size >= 0.9 * r) else r
                break
            rad = max(radius_min, r * shrink)
            if trial == max_reject:
                x = x_prev
                small = (np.isfinite(size) and size <= max(tolx_abs, tolx_rel * max(1.0, xn))) or (
                    np.isfinite(pred) and abs(pred) <= max(tolf, 1e-16))
                if small:
                    terminate = ("Trust-region radius exhausted before full convergence"
                                 if (require_all and ng >= target) else "Trust-region step converged")
                else:
                    terminate = f"Trust-region rejected all candidate steps at Newton iteration {nit}"
        radii.append(rad)
        if terminate is not None and terminate.startswith("Non-finite"):
            break
        conv_e = accepted and np.isfinite(dE) and abs(dE) < tolf
        conv_s = accepted and (step < tolx_abs or step / max(1.0, xn) < tolx_rel)
        post = np.nan
        if require_all and conv_e and conv_s:
            post = float(np.linalg.norm(gradient(x)))
            if not np.isfinite(post):
                x = x_prev
                msg = f"Non-finite post-update gradient norm at Newton iteration {nit}"
                break
        if terminate is not None:
            msg = terminate
            break
        if not np.isfinite(fx):
            msg = f"Non-finite energy after update at Newton iteration {nit}"
            x = x_prev
            break
        if require_all:
            if conv_e and conv_s and post < target:
                msg = "Converged (energy, step, gradient)"
                break
        elif conv_e:
            msg = "Energy change converged"
            break
    return {"x": x, "fun": fx, "nit": nit, "message": msg, "linear_iters": lin, "radii": radii}
