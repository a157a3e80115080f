"""Matrix-free Mohr-Coulomb Hessian operator and diagonal (fne_hvp, fne_hessian_diag).

Mirrors the reference's operator tests (tests/test_slope_stability_petsc.py:475-556: the
matrix-free operator's product and diagonal equal those of the assembled Hessian): for P2 and
P4, in the full Mohr-Coulomb mode at a state with plastic quadrature points and in the elastic
mode, H v and diag(H) from the element-by-element kernels equal the product with and the
diagonal of the assembled FP64 Hessian and of the oracle's Hessian, to 1e-11 of their largest
entry (FP64 sums in another order).
"""

import numpy as np
import pytest
import torch

from fenics_nonlinear_energies_amd.problems.slope_stability import SlopeStabilityReorderedElementAssembler
from oracle import assembly
from tests.test_gpu_mc import _state
from tests.test_oracle_mc import mc_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("degree,mode", [(4, 0), (2, 0), (4, 1)])
def test_matrix_free_hvp_and_diagonal(require_gpu, degree, mode):
    p, adj = mc_params(h=4.0, degree=degree, lam=1.5)
    asm = SlopeStabilityReorderedElementAssembler(p, None, adj, hessian_precision="fp64")
    asm.set_constitutive_mode(mode)
    lay = asm.layout
    u = _state(p, lay, 3)
    x = asm.create_vec(u[lay.perm])
    asm.assemble_hessian(x)
    H = asm.A.to_scipy()
    v = np.random.default_rng(4).standard_normal(lay.n_free)
    vt = torch.as_tensor(v).cuda()
    y = torch.empty_like(vt)
    asm.hvp_device(x.t, vt, y)
    ref = H @ v
    assert np.abs(y.cpu().numpy() - ref).max() <= 1e-11 * np.abs(ref).max()
    dg = torch.empty_like(vt)
    asm.hessian_diagonal_device(x.t, dg)
    dref = H.diagonal()
    assert np.abs(dg.cpu().numpy() - dref).max() <= 1e-11 * np.abs(dref).max()
    if mode == 0:  # against the oracle's Hessian in the original numbering
        Ho = assembly.assemble("mc", p, u, want=("h",))["h"]
        yo = Ho @ v[lay.iperm]
        assert np.abs(y.cpu().numpy()[lay.iperm] - yo).max() <= 1e-10 * np.abs(yo).max()
    # deterministic
    y2 = torch.empty_like(vt)
    asm.hvp_device(x.t, vt, y2)
    assert torch.equal(y, y2)
    asm.cleanup()


@pytest.mark.parametrize("mode", ["matfree_element", "matfree_overlap"])
def test_prepared_operator_matches_assembled(require_gpu, mode):
    """The reference's operator API (prepare_matrix_free_operator / mult / getDiagonal,
    slope_stability/jax_petsc/reordered_element_assembler.py:313-562) against the assembled
    FP64 Hessian at the same point, as tests/test_slope_stability_petsc.py:475-556 check
    (1e-8 there; 1e-11 of the largest entry here)."""
    p, adj = mc_params(h=4.0, degree=4, lam=1.5)
    asm = SlopeStabilityReorderedElementAssembler(p, None, adj, hessian_precision="fp64")
    lay = asm.layout
    u = _state(p, lay, 5)
    trial = asm.create_vec(np.random.default_rng(12).standard_normal(lay.n_free))
    asm.assemble_hessian(u[lay.perm])
    ref = asm.create_vec()
    asm.A.mult(trial, ref)
    dref = asm.create_vec()
    asm.A.getDiagonal(dref)

    op = asm.prepare_matrix_free_operator(u[lay.perm], mode=mode)
    assert op is asm.get_matrix_free_operator(mode)
    assert op.getSize() == (lay.n_free, lay.n_free) and op.getType() == "python"
    out = asm.create_vec()
    op.mult(trial, out)
    scale = np.abs(ref.array).max()
    assert np.abs(out.array - ref.array).max() <= 1e-11 * scale
    d = asm.create_vec()
    op.getDiagonal(d)
    assert np.abs(d.array - dref.array).max() <= 1e-11 * np.abs(dref.array).max()
    st = asm.matrix_free_summary()
    assert st["mode"] == mode and st["mult_calls"] == 1 and st["diagonal_source"] == "element_hessian_diagonal"
    # another mode's operator has no linearisation point of its own
    other = [m for m in asm.MATRIX_FREE_MODES if m != mode][0]
    with pytest.raises(RuntimeError):
        asm.get_matrix_free_operator(other).mult(trial, out)
    asm.cleanup()


def test_matrix_free_krylov_and_python_pc(require_gpu):
    """A KSP on the shell operator: Jacobi CG from its diagonal takes the iterations of the
    device CG on the assembled FP64 matrix (same operator to rounding, +-1), and the
    "overlap_lu" Python PC (the factored operator) makes CG converge in one iteration."""
    from fenics_nonlinear_energies_amd.linalg import KSP

    p, adj = mc_params(h=4.0, degree=2, lam=1.5)
    asm = SlopeStabilityReorderedElementAssembler(p, None, adj, hessian_precision="fp64")
    lay = asm.layout
    u = _state(p, lay, 7)
    b = asm.create_vec(np.random.default_rng(3).standard_normal(lay.n_free))
    asm.assemble_hessian(u[lay.perm])
    ref = KSP("cg", "jacobi", rtol=1e-8, max_it=5000)
    ref.setOperators(asm.A)
    x_ref = asm.create_vec()
    ref.solve(b, x_ref)

    op = asm.prepare_matrix_free_operator(u[lay.perm], mode="matfree_element")
    ksp = KSP("cg", "jacobi", rtol=1e-8, max_it=5000)
    ksp.setOperators(op)
    x = asm.create_vec()
    ksp.solve(b, x)
    assert ksp.getConvergedReason() > 0 and ref.getConvergedReason() > 0
    assert abs(ksp.getIterationNumber() - ref.getIterationNumber()) <= 1
    assert np.abs(x.array - x_ref.array).max() <= 1e-6 * np.abs(x_ref.array).max()

    lu = KSP("cg", "python", rtol=1e-10, max_it=10)
    ctx = asm.make_matrix_free_python_pc_context("overlap_lu")
    assert "overlap_lu" in ctx.view()
    lu.setPythonContext(ctx)
    lu.setOperators(op)
    lu.setUp()
    x_lu = asm.create_vec()
    lu.solve(b, x_lu)
    assert lu.getConvergedReason() > 0 and lu.getIterationNumber() <= 2
    assert np.abs(x_lu.array - x_ref.array).max() <= 1e-6 * np.abs(x_ref.array).max()
    with pytest.raises(ValueError):
        asm.prepare_matrix_free_python_pc("ilu")
    asm.cleanup()
