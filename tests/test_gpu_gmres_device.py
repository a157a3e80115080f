"""The device GMRES / FGMRES (fne_gmres) against the host-driven restatement of the same loop.

Both follow PETSc's KSPGMRES / KSPFGMRES (left / right preconditioning, classical Gram-Schmidt,
Givens rotations); the device keeps every scalar in device memory and polls its flags.  On
each system the two must stop at the same iteration with the same reason and agree on the
solution to 1e-9 relative: an assembled p-Laplace Hessian (Jacobi), a nonsymmetric sparse
matrix (Jacobi and none, restarts of 5 and 12), the Neo-Hookean beam Hessian with the
recorded multigrid cycle, the iteration cap, and a happy breakdown (a right-hand side in a
3-dimensional invariant subspace).  Repeated device solves are bitwise identical.  The
oracle comparisons of the same solvers are in test_gpu_trust_krylov.py and test_gpu_pmg.py.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from fenics_nonlinear_energies_amd.linalg import KSP, DeviceCSR
from fenics_nonlinear_energies_amd.problems.hyperelasticity import HEReorderedElementAssembler, build_beam_problem
from fenics_nonlinear_energies_amd.problems.plaplace import PLaplaceReorderedElementAssembler, build_plaplace_problem

pytestmark = pytest.mark.gpu


def _device_csr(M: sp.csr_matrix) -> DeviceCSR:
    M = M.tocsr()
    M.sort_indices()
    A = DeviceCSR(M.indptr, M.indices, M.shape[0], precision="fp64")
    A.data.copy_(torch.as_tensor(M.data, dtype=torch.float64))
    return A


def _both(ksp: KSP, b: torch.Tensor, flexible: bool):
    """(device x, its, reason), (host x, its, reason) of one configured KSP."""
    out = []
    for host in (False, True):
        x = torch.zeros_like(b)
        if host:
            (ksp._solve_fgmres_host if flexible else ksp._solve_gmres_host)(b, x)
        else:
            assert ksp._device_gmres(b, x, flexible)
        out.append((x.cpu().numpy(), ksp.its, ksp.reason, ksp.rnorm))
    return out


def _check(dev, host, rel=1e-9):
    (xd, itd, rd, rnd), (xh, ith, rh, rnh) = dev, host
    print("device", itd, rd, rnd, "host", ith, rh, rnh, flush=True)
    assert (itd, rd) == (ith, rh)
    assert np.linalg.norm(xd - xh) <= rel * max(np.linalg.norm(xh), 1e-300)
    # the final residual estimate sits many orders below ||r_0||, where rounding differences of
    # the two orthogonalisations (compensated device sums, torch.mv on the host) show
    assert abs(rnd - rnh) <= 1e-2 * max(rnh, rnd) + 1e-13


def _nonsymmetric(n=900, seed=3):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=8.0 / n, random_state=seed, format="csr")
    M = M + sp.diags(4.0 + rng.random(n)) + 0.5 * sp.eye(n, k=1) - 0.3 * sp.eye(n, k=-2)
    return M.tocsr(), rng.standard_normal(n)


# Restarted runs that creep down for hundreds of iterations (GMRES(5) at 1e-8 on this Hessian:
# 866 against 885) cross rtol at different iterations under different rounding; the restart
# cases below converge steadily.
@pytest.mark.parametrize("flexible", [False, True])
@pytest.mark.parametrize("restart,rtol", [(30, 1e-10), (30, 1e-4)])
def test_device_gmres_equals_host_on_hessian(require_gpu, flexible, restart, rtol):
    params, adj, _ = build_plaplace_problem(48)
    asm = PLaplaceReorderedElementAssembler(params, None, adj, ksp_type="gmres")
    u = np.random.default_rng(0).random(params["freedofs"].size)
    asm.assemble_hessian(asm.create_vec(u[asm.layout.perm]))
    ksp = KSP("fgmres" if flexible else "gmres", "jacobi", rtol=rtol, max_it=3000, restart=restart)
    ksp.setOperators(asm.A)
    ksp.setUp()
    b = torch.as_tensor(np.random.default_rng(1).standard_normal(asm.layout.n_free), device="cuda")
    dev, host = _both(ksp, b, flexible)
    assert dev[2] == 2
    _check(dev, host)
    asm.cleanup()


@pytest.mark.parametrize("flexible", [False, True])
@pytest.mark.parametrize("pc,restart", [("jacobi", 12), ("none", 12), ("jacobi", 5)])
def test_device_gmres_equals_host_nonsymmetric(require_gpu, flexible, pc, restart):
    M, b = _nonsymmetric()
    A = _device_csr(M)
    ksp = KSP("fgmres" if flexible else "gmres", pc, rtol=1e-11, max_it=500, restart=restart)
    ksp.setOperators(A)
    ksp.setUp()
    bt = torch.as_tensor(b, device="cuda")
    dev, host = _both(ksp, bt, flexible)
    assert dev[2] == 2
    _check(dev, host)
    res = np.linalg.norm(M @ dev[0] - b) / np.linalg.norm(b)
    assert res < 1e-8
    # repeated device solves: bitwise identical
    x2 = torch.zeros_like(bt)
    assert ksp._device_gmres(bt, x2, flexible)
    assert np.array_equal(x2.cpu().numpy(), dev[0])


@pytest.mark.parametrize("flexible", [False, True])
def test_device_gmres_with_multigrid_graph(require_gpu, flexible):
    p, adj, u0 = build_beam_problem(40, 4, 4)
    asm = HEReorderedElementAssembler(p, None, adj, ksp_type="fgmres" if flexible else "gmres", pc_type="gamg",
                                      ksp_rtol=1e-9, ksp_max_it=400)
    x0 = asm.create_vec((u0 + 1e-5 * np.random.default_rng(0).standard_normal(u0.size))[asm.layout.perm])
    asm.assemble_hessian(x0)
    ksp = asm.ksp
    ksp.setOperators(asm.A)
    ksp.setUp()
    assert ksp._amg is not None and ksp._amg._graph is not None
    b = torch.as_tensor(np.random.default_rng(2).standard_normal(asm.layout.n_free), device="cuda")
    dev, host = _both(ksp, b, flexible)
    assert dev[2] == 2 and dev[1] < 200
    _check(dev, host, rel=1e-8)
    asm.cleanup()


def test_device_gmres_iteration_cap_and_breakdown(require_gpu):
    M, b = _nonsymmetric(seed=5)
    A = _device_csr(M)
    ksp = KSP("gmres", "jacobi", rtol=1e-14, max_it=7, restart=4)
    ksp.setOperators(A)
    ksp.setUp()
    dev, host = _both(ksp, torch.as_tensor(b, device="cuda"), False)
    assert dev[1] == 7 and dev[2] == -3
    _check(dev, host)
    # b in the span of three eigenvectors of a diagonal operator: GMRES ends by happy
    # breakdown at iteration 3 and then meets the tolerance (atol or rtol) on the restart
    n = 200
    d = np.linspace(1.0, 3.0, n)
    D = sp.diags(d).tocsr()
    bb = np.zeros(n)
    bb[[3, 50, 120]] = [1.0, -2.0, 0.5]
    ksp = KSP("gmres", "none", rtol=1e-12, max_it=50, restart=30)
    ksp.setOperators(_device_csr(D))
    ksp.setUp()
    dev, host = _both(ksp, torch.as_tensor(bb, device="cuda"), False)
    assert dev[1] == 3 and dev[2] in (2, 3)
    np.testing.assert_allclose(dev[0], bb / d, rtol=0, atol=1e-12)
    _check(dev, host)
