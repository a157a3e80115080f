"""Mohr-Coulomb strength-reduction continuation on the GPU against the CPU oracle.

The elastic initial guess is checked against a direct solve of the oracle's elastic
stiffness (rel. 1e-7, CG tolerance 1e-12).  Each continuation step then runs a fixed number
of Newton iterations (stopping tests off, KSP rtol 1e-10) from the previous step's state,
on the GPU and with the oracle's golden-section Newton (direct solves); the two follow the
same iterates up to rounding, so the energies after the last iteration agree to 1e-9
relative and the gradient norm has dropped by three orders of magnitude.  (Perfect
plasticity is non-smooth: Newton on it stalls near 1e-4..1e-6 gradient norms, so a fixed
iteration budget is the reproducible comparison.)
"""

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from fenics_nonlinear_energies_amd.continuation import elastic_initial_guess, run_lambda_continuation
from fenics_nonlinear_energies_amd.problems.slope_stability import (
    SlopeStabilityReorderedElementAssembler,
    davis_b_reduction,
)
from oracle import assembly, solvers
from tests.test_oracle_mc import mc_params

pytestmark = pytest.mark.gpu


# P2 below the factor of safety (about 1.2 for this slope); P4 over the strength-reduction
# range of the P4 continuation workload (lambda 0.5 -> 1.0)
@pytest.mark.parametrize("degree,lambdas", [(2, (1.0, 1.05, 1.1)), (4, (0.5, 0.75, 1.0))])
def test_elastic_guess_and_continuation_match_oracle(require_gpu, degree, lambdas):
    p, adj = mc_params(h=4.0, degree=degree, lam=1.0)
    asm = SlopeStabilityReorderedElementAssembler(p, None, adj, hessian_precision="fp64", ksp_type="cg")
    lay = asm.layout
    x, its = elastic_initial_guess(asm, rtol=1e-12, max_it=50000)
    K = assembly.assemble("mc", p, np.zeros(lay.n_free), want=("h",), elastic_only=True)["h"]
    f = p["force"][p["freedofs"]]
    u_el = spla.spsolve(K.tocsc(), f)
    assert np.abs(x.array[lay.iperm] - u_el).max() <= 1e-7 * np.abs(u_el).max()

    lambdas = list(lambdas)
    n_it = 8
    opts = dict(tolg=0.0, tolg_rel=0.0, tolf=0.0, tolx_rel=0.0, tolx_abs=0.0, maxit=n_it,
                require_all_convergence=False)
    recs = run_lambda_continuation(asm, lambdas, x, ksp_rtol=1e-10, ksp_max_it=50000, newton_kwargs=opts)
    assert all(r["nit"] == n_it for r in recs), [(r["message"], r["nit"]) for r in recs]
    u_prev = u_el
    for lam, rec in zip(lambdas, recs):
        q = dict(p)
        q["cohesion"], q["phi_deg"] = davis_b_reduction(p["c0"], p["phi_deg_raw"], p["psi_deg"], lam)
        E = lambda u: assembly.assemble("mc", q, u, want=("e",))["e"]
        G = lambda u: assembly.assemble("mc", q, u, want=("g",))["g"]

        def solve(u, rhs):
            H = assembly.assemble("mc", q, u, want=("h",))["h"]
            return spla.spsolve(H.tocsc(), rhs), 1

        g0 = np.linalg.norm(G(u_prev))
        res = solvers.newton_golden(E, G, solve, u_prev, tolf=-1.0, tolg=0.0, tolx_rel=0.0, tolx_abs=0.0,
                                    maxit=n_it, require_all=False)
        assert res["nit"] == n_it
        assert abs(rec["energy"] - res["fun"]) <= 1e-9 * abs(res["fun"]), (rec, res["fun"])
        assert np.linalg.norm(G(res["x"])) <= 1e-3 * g0
        u_prev = res["x"]
    asm.cleanup()
