"""GPU parity of the Neo-Hookean tetrahedron kernels (A3) against the CPU oracle."""

import numpy as np
import pytest
import torch

from fenics_nonlinear_energies_amd.problems.hyperelasticity import (
    HEReorderedElementAssembler,
    build_beam_problem,
    rotate_right_face_from_reference,
)
from oracle import assembly
from tests import fixtures

pytestmark = pytest.mark.gpu


def _state(p, u0, seed, amp=2e-4):
    return u0 + amp * np.random.default_rng(seed).standard_normal(u0.size)


@pytest.mark.parametrize("amode", ["staged", "rowgather"])
@pytest.mark.parametrize("absdet", [False, True])
def test_he_energy_gradient_hessian_fp64(require_gpu, absdet, amode):
    p, adj, u0 = build_beam_problem(16, 3, 3)
    asm = HEReorderedElementAssembler(p, None, adj, use_abs_det=absdet, assembly_mode=amode)
    u = _state(p, u0, 1)
    lay = asm.layout
    x = asm.create_vec(u[lay.perm])
    ref = assembly.assemble("he", p, u, dirichlet_key="u_0_ref", use_abs_det=absdet)
    assert abs(asm.energy_fn(x) - ref["e"]) <= 1e-12 * abs(ref["e"])
    g = x.duplicate()
    asm.gradient_fn(x, g)
    assert np.max(np.abs(g.array[lay.iperm] - ref["g"])) <= 1e-10 * np.max(np.abs(ref["g"]))
    asm.assemble_hessian(x)
    H = asm.A.to_scipy()[lay.iperm][:, lay.iperm]
    assert abs(H - ref["h"]).max() <= 1e-11 * abs(ref["h"]).max()
    asm.cleanup()


@pytest.mark.parametrize("amode", ["staged", "rowgather"])
def test_he_fp32_hessian_and_rotated_face(require_gpu, amode):
    p, adj, u0 = build_beam_problem(20, 2, 2)
    asm = HEReorderedElementAssembler(p, None, adj, hessian_precision="fp32", assembly_mode=amode)
    u0_rot = rotate_right_face_from_reference(p["u_0_ref"], p["nodes2coord"], 0.3, p["right_nodes"])
    asm.update_dirichlet(u0_rot)
    pr = dict(p)
    pr["u_0_ref"] = u0_rot
    u = _state(p, u0, 2)
    x = asm.create_vec(u[asm.layout.perm])
    ref = assembly.assemble("he", pr, u, dirichlet_key="u_0_ref")
    assert abs(asm.energy_fn(x) - ref["e"]) <= 1e-12 * abs(ref["e"])
    asm.assemble_hessian(x)
    H = asm.A.to_scipy()[asm.layout.iperm][:, asm.layout.iperm]
    assert abs(H - ref["h"]).max() <= 2e-6 * abs(ref["h"]).max()
    h1 = asm.A.data.clone()
    asm.assemble_hessian(x)
    assert torch.equal(h1, asm.A.data)


def test_he_reference_mesh_element_eval(require_gpu):
    raw, adj = fixtures.he(1)
    from fenics_nonlinear_energies_amd.problems.hyperelasticity import problem_from_mesh

    p, a2, u0 = problem_from_mesh(raw["nodes2coord"], raw["elems2nodes"], float(raw["C1"]), float(raw["D1"]),
                                  dphi=(raw["dphix"], raw["dphiy"], raw["dphiz"], raw["vol"]))
    asm = HEReorderedElementAssembler(p, None, adj)
    u = _state(p, u0, 3)
    e, g, h = asm.element_eval(asm.create_vec(u[asm.layout.perm]))
    q = assembly.element_eval("he", p, assembly.local_state(p, u, "u_0_ref"))
    assert np.allclose(e.cpu().numpy(), q["e"], rtol=1e-11, atol=1e-12 * np.abs(q["e"]).max())
    assert np.allclose(g.cpu().numpy(), q["g"], rtol=1e-10, atol=1e-10 * np.abs(q["g"]).max())
    assert np.allclose(h.cpu().numpy(), q["h"], rtol=1e-10, atol=1e-10 * np.abs(q["h"]).max())


def test_he_nonpositive_jacobian_gives_nan_energy(require_gpu):
    """J <= 0 makes ln J non-finite; the energy is NaN/inf, never an error (line search repairs)."""
    p, adj, u0 = build_beam_problem(6, 1, 1)
    asm = HEReorderedElementAssembler(p, None, adj)
    u = u0.copy()
    u[0::3] *= -1.0  # mirror x: inverts every free-node element orientation
    E = asm.energy_fn(asm.create_vec(u[asm.layout.perm]))
    assert not np.isfinite(E)
