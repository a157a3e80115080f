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


def test_matrix_free_rejects_other_families(require_gpu):
    from fenics_nonlinear_energies_amd import native
    from fenics_nonlinear_energies_amd.problems.plaplace import PLaplaceReorderedElementAssembler, build_plaplace_problem

    p, adj, _ = build_plaplace_problem(8)
    asm = PLaplaceReorderedElementAssembler(p, None, adj)
    x = torch.zeros(asm.layout.n_free, dtype=torch.float64, device="cuda")
    v, y = x.clone(), torch.empty_like(x)  # kept alive: distinct buffers
    rc = native.lib().fne_hvp(asm._plan, native.ptr(x), native.ptr(asm.dvals), native.ptr(v), native.ptr(y),
                              native.stream_ptr())
    assert rc == native.FNE_EUNSUPPORTED
    asm.cleanup()
