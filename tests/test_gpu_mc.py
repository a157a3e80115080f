"""GPU parity of the Mohr-Coulomb P2/P4 kernels against the CPU oracle.

Tolerances: FP64 mode -- energy 1e-12 relative, gradient 1e-11 and Hessian 1e-10 of their
largest entry (MFMA FP64 sums in another order than numpy).  FP32 Hessian -- 3e-5 of the
largest entry (FP32 MFMA accumulation over 2 n_q = 38 products).
"""

import numpy as np
import pytest
import torch

from fenics_nonlinear_energies_amd.problems.slope_stability import SlopeStabilityReorderedElementAssembler
from oracle import assembly
from tests.test_oracle_mc import mc_params

pytestmark = pytest.mark.gpu


def _state(p, lay, seed, scale=2e-3):
    rng = np.random.default_rng(seed)
    X = p["nodes"]
    v = np.zeros(2 * X.shape[0])
    v[0::2] = scale * X[:, 1] + 0.05 * scale * rng.standard_normal(X.shape[0])
    v[1::2] = -0.5 * scale * X[:, 1] + 0.05 * scale * rng.standard_normal(X.shape[0])
    return v[p["freedofs"]]


@pytest.mark.parametrize("degree", [4, 2])
def test_element_eval_matches_oracle(require_gpu, degree):
    p, adj = mc_params(h=4.0, degree=degree, lam=1.5)
    asm = SlopeStabilityReorderedElementAssembler(p, None, adj, hessian_precision="fp64")
    u = _state(p, asm.layout, 1)
    e, g, h = asm.element_eval(asm.create_vec(u[asm.layout.perm]))
    v = assembly.local_state(p, u)
    q = assembly.element_eval("mc", p, v)
    e, g, h = e.cpu().numpy(), g.cpu().numpy(), h.cpu().numpy()  # original element order
    assert np.allclose(e, q["e"], rtol=1e-12, atol=1e-12 * np.abs(q["e"]).max())
    assert np.abs(g - q["g"]).max() <= 1e-11 * np.abs(q["g"]).max()
    assert np.abs(h - q["h"]).max() <= 1e-10 * np.abs(q["h"]).max()
    asm.cleanup()


@pytest.mark.parametrize("degree", [4, 2])
def test_assembly_fp64_matches_oracle(require_gpu, degree):
    p, adj = mc_params(h=4.0, degree=degree, lam=1.5)
    asm = SlopeStabilityReorderedElementAssembler(p, None, adj, hessian_precision="fp64")
    lay = asm.layout
    u = _state(p, lay, 2)
    ref = assembly.assemble("mc", p, u)
    x = asm.create_vec(u[lay.perm])
    E = asm.energy_fn(x)
    assert abs(E - ref["e"]) <= 1e-12 * abs(ref["e"])
    g = x.duplicate()
    asm.gradient_fn(x, g)
    assert np.abs(g.array[lay.iperm] - ref["g"]).max() <= 1e-11 * np.abs(ref["g"]).max()
    asm.assemble_hessian(x)
    H = asm.A.to_scipy()[lay.iperm][:, lay.iperm]
    assert abs(H - ref["h"]).max() <= 1e-10 * abs(ref["h"]).max()
    # deterministic: bitwise identical on a second assembly
    v1 = asm.A.data.clone()
    asm.assemble_hessian(x)
    assert torch.equal(v1, asm.A.data)
    asm.cleanup()


def test_fp32_hessian_and_fused_call(require_gpu):
    p, adj = mc_params(h=4.0, degree=4, lam=1.5)
    asm = SlopeStabilityReorderedElementAssembler(p, None, adj, hessian_precision="fp32")
    lay = asm.layout
    u = _state(p, lay, 3)
    ref = assembly.assemble("mc", p, u)
    x = asm.create_vec(u[lay.perm])
    g1, g2 = x.duplicate(), x.duplicate()
    e_sep = float(asm.energy_device(x.t).item())
    asm.gradient_device(x.t, g1.t)
    asm.hessian_device(x.t)
    h_sep = asm.A.data.clone()
    e_f = float(asm.energy_grad_hessian_device(x.t, g2.t).item())
    assert e_f == e_sep
    assert torch.equal(g1.t, g2.t) and torch.equal(h_sep, asm.A.data)
    H = asm.A.to_scipy()[lay.iperm][:, lay.iperm]
    assert abs(H - ref["h"]).max() <= 3e-5 * abs(ref["h"]).max()
    assert np.abs(g1.array[lay.iperm] - ref["g"]).max() <= 1e-11 * np.abs(ref["g"]).max()
    asm.cleanup()


def test_elastic_mode_matches_elastic_stiffness(require_gpu):
    p, adj = mc_params(h=4.0, degree=4, lam=1.0)
    asm = SlopeStabilityReorderedElementAssembler(p, None, adj, hessian_precision="fp64")
    lay = asm.layout
    u = _state(p, lay, 4, scale=5e-3)
    asm.assemble_hessian_with_mode(asm.create_vec(u[lay.perm]), constitutive_mode="elastic")
    H = asm.A.to_scipy()[lay.iperm][:, lay.iperm]
    ref = assembly.assemble("mc", p, u, want=("h",), elastic_only=True)
    assert abs(H - ref["h"]).max() <= 1e-10 * abs(ref["h"]).max()
    # the mode switch is scoped to the call: the energy afterwards is the full Mohr-Coulomb one
    Ee = asm.energy_fn(asm.create_vec(u[lay.perm]))
    Eref = assembly.assemble("mc", p, u, want=("e",))["e"]
    assert abs(Ee - Eref) <= 1e-12 * abs(Eref)
    asm.cleanup()
