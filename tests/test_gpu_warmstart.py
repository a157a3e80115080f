"""Learned warm start on the rotating Neo-Hookean beam: the trained network saves Newton
iterations over the extrapolated guess (its fallback) at the same converged energy.

Train: 3 load-step runs of an 80x4x4 beam (E, nu from a seeded draw), 24 steps to 4 pi,
per-node MLP 10-128-128-128-128-3 trained by the BF16-MFMA kernels.  Evaluate: an unseen
material on the same beam; every load step of every variant must converge (no step may
end at the 100-iteration cap), the learned guess (energy-guarded) must need fewer Newton
iterations in total than the extrapolated guess, and the runs end at the same energy
(rel. 1e-9).
"""

import numpy as np
import pytest

from fenics_nonlinear_energies_amd.experiments import he_warmstart_study

pytestmark = pytest.mark.gpu


def test_learned_warm_start_saves_newton_iterations(require_gpu):
    rng = np.random.default_rng(3)
    mats = [(float(rng.uniform(1e8, 4e8)), float(rng.uniform(0.25, 0.4))) for _ in range(3)]
    nk = dict(tolf=0.0, tolg=2e-4, tolg_rel=0.0, tolx_rel=0.0, tolx_abs=0.0, maxit=100,
              require_all_convergence=False)
    out = he_warmstart_study(train_materials=mats, eval_material=(2e8, 0.3), train_mesh=(80, 4, 4),
                             eval_mesh=(80, 4, 4), total_steps=24, total_angle=4 * np.pi, epochs=40,
                             newton_kwargs=nk, log=lambda *a: None)
    v = out["variants"]
    print({k: (x["newton_its"], round(x["wall_s"], 2), x["per_step_nit"]) for k, x in v.items()})
    for name, x in v.items():
        assert x["all_converged"], (name, x["per_step_nit"])
        assert max(x["per_step_nit"]) < nk["maxit"], (name, x["per_step_nit"])
    assert v["learned_fp32"]["newton_its"] < v["extrapolate_fp32"]["newton_its"]
    assert abs(v["learned_fp32"]["final_energy"] - v["previous_fp64"]["final_energy"]) <= 1e-9 * abs(
        v["previous_fp64"]["final_energy"])
