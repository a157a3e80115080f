"""SFD Hessian mode and element-by-element Hessian-vector products on the GPU.

* ``fne_hvp`` for the P1 families (staged plans) and the tetrahedra (row-gather plans): H(x) v
  equals the product with the assembled FP64 Hessian (1e-12 of its largest entry).
* ``local_hessian_mode`` "sfd_local" / "sfd_local_vmap": one Hessian-vector product per colour
  of the distance-2 colouring gives the same CSR values as the element assembly, the check the
  reference makes for its slope problem (element vs local-SFD Hessian to 1e-10,
  tests/test_slope_stability_petsc.py:588-600), here for all four energy families.
"""

import numpy as np
import pytest
import torch

from fenics_nonlinear_energies_amd.problems.ginzburg_landau import GLReorderedElementAssembler, build_gl_problem, sine_noise_init
from fenics_nonlinear_energies_amd.problems.hyperelasticity import HEReorderedElementAssembler, build_beam_problem
from fenics_nonlinear_energies_amd.problems.plaplace import PLaplaceReorderedElementAssembler, build_plaplace_problem
from fenics_nonlinear_energies_amd.problems.slope_stability import SlopeStabilityReorderedElementAssembler
from tests.test_gpu_mc import _state
from tests.test_oracle_mc import mc_params

pytestmark = pytest.mark.gpu


def _case(name, mode="element"):
    rng = np.random.default_rng(7)
    if name == "plaplace":
        p, adj, _ = build_plaplace_problem(24)
        asm = PLaplaceReorderedElementAssembler(p, None, adj, hessian_precision="fp64", local_hessian_mode=mode)
        u = rng.random(p["freedofs"].size)
    elif name == "gl":
        p, adj = build_gl_problem(24)
        asm = GLReorderedElementAssembler(p, None, adj, hessian_precision="fp64", local_hessian_mode=mode)
        u = sine_noise_init(p, seed=3)
    elif name == "he":
        p, adj, u0 = build_beam_problem(20, 3, 3)
        asm = HEReorderedElementAssembler(p, None, adj, hessian_precision="fp64", local_hessian_mode=mode)
        u = u0 + 1e-4 * rng.standard_normal(u0.size)
    else:
        degree = 4 if name == "mc_p4" else 2
        p, adj = mc_params(h=4.0, degree=degree, lam=1.5)
        asm = SlopeStabilityReorderedElementAssembler(p, None, adj, hessian_precision="fp64", local_hessian_mode=mode)
        u = _state(p, asm.layout, 5)
    return asm, asm.create_vec(u[asm.layout.perm])


@pytest.mark.parametrize("name", ["plaplace", "gl", "he"])
def test_hvp_matches_assembled_hessian(require_gpu, name):
    asm, x = _case(name)
    asm.assemble_hessian(x)
    H = asm.A.to_scipy()
    v = np.random.default_rng(1).standard_normal(asm.layout.n_free)
    vt = torch.as_tensor(v).cuda()
    y = torch.empty_like(vt)
    asm.hvp_device(x.t, vt, y)
    ref = H @ v
    assert np.abs(y.cpu().numpy() - ref).max() <= 1e-12 * np.abs(H).max() * np.abs(v).max() * 40
    y2 = torch.empty_like(vt)
    asm.hvp_device(x.t, vt, y2)
    assert torch.equal(y, y2)
    asm.cleanup()


@pytest.mark.parametrize("name", ["plaplace", "gl", "he", "mc_p2", "mc_p4"])
@pytest.mark.parametrize("mode", ["sfd_local", "sfd_local_vmap"])
def test_sfd_hessian_matches_element_hessian(require_gpu, name, mode):
    ref_asm, x = _case(name)
    ref_asm.assemble_hessian(x)
    Href = ref_asm.A.to_scipy()
    asm, x2 = _case(name, mode)
    t = asm.assemble_hessian(x2)
    H = asm.A.to_scipy()
    assert t["n_hvps"] == asm._sfd_n_colors > 0
    print(name, mode, "colours", asm._sfd_n_colors, "n_free", asm.layout.n_free, flush=True)
    np.testing.assert_array_equal(H.indices, Href.indices)
    assert np.abs(H.data - Href.data).max() <= 1e-10 * np.abs(Href.data).max()
    ref_asm.cleanup()
    asm.cleanup()


def test_hvp_rejects_plans_without_the_needed_arrays(require_gpu):
    """Row-gather P1 plans have no staged patches: fne_hvp says so (FNE_EUNSUPPORTED); the
    Hessian diagonal stays a Mohr-Coulomb-only operation."""
    from fenics_nonlinear_energies_amd import native

    p, adj, _ = build_plaplace_problem(8)
    asm = PLaplaceReorderedElementAssembler(p, None, adj, assembly_mode="rowgather")
    x = torch.zeros(asm.layout.n_free, dtype=torch.float64, device="cuda")
    v, y = x.clone(), torch.empty_like(x)
    lib = native.lib()
    assert lib.fne_hvp(asm._plan, native.ptr(x), native.ptr(asm.dvals), native.ptr(v), native.ptr(y),
                       native.stream_ptr()) == native.FNE_EUNSUPPORTED
    assert lib.fne_hessian_diag(asm._plan, native.ptr(x), native.ptr(asm.dvals), native.ptr(y),
                                native.stream_ptr()) == native.FNE_EUNSUPPORTED
    assert lib.fne_hvp(asm._plan, native.ptr(x), native.ptr(asm.dvals), native.ptr(y), native.ptr(y),
                       native.stream_ptr()) == native.FNE_EINVAL  # v and y alias
    asm.cleanup()
