"""Smoothed-aggregation AMG preconditioner (pc_type "gamg") on the GPU.

CG with the V-cycle reaches the solution of a direct solve (rel. 1e-6 at rtol 1e-10) in far
fewer iterations than point Jacobi, for an elasticity Hessian (rigid-body near-nullspace)
and a scalar p-Laplace Hessian (constant near-nullspace).
"""

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from fenics_nonlinear_energies_amd.problems.hyperelasticity import (
    HEReorderedElementAssembler,
    build_beam_problem,
    rigid_modes,
)
from fenics_nonlinear_energies_amd.problems.plaplace import PLaplaceReorderedElementAssembler, build_plaplace_problem

pytestmark = pytest.mark.gpu


def _solve(asm, pc, rtol, b):
    ksp = asm.ksp
    ksp.pc_type = pc
    ksp._amg = None
    ksp._inv_diag = None
    ksp.setOperators(asm.A)
    ksp.setTolerances(rtol=rtol, max_it=20000)
    ksp.setUp()
    x = b.duplicate()
    ksp.solve(b, x)
    return x, ksp.getIterationNumber()


@pytest.mark.parametrize("coarse", ["host", "device"])
@pytest.mark.parametrize("kind", ["he", "plaplace"])
def test_gamg_cg_matches_direct_solve(require_gpu, kind, coarse):
    """Multigrid-preconditioned CG reaches the direct solution in far fewer iterations than
    Jacobi, with the coarsest level factorised on the host (LAPACK) or on the GPU."""
    opts = {"mg_coarse_factor": coarse}
    if kind == "he":
        p, adj, u0 = build_beam_problem(40, 4, 4)
        asm = HEReorderedElementAssembler(p, None, adj, ksp_type="cg", pc_type="gamg", pc_options=opts)
        u = u0 + 1e-5 * np.random.default_rng(0).standard_normal(u0.size)
        x0 = asm.create_vec(u[asm.layout.perm])
    else:
        p, adj, _ = build_plaplace_problem(96)
        asm = PLaplaceReorderedElementAssembler(p, None, adj, ksp_type="cg", pc_type="gamg", pc_options=opts)
        x0 = asm.create_vec(np.random.default_rng(0).random(asm.layout.n_free))
    asm.assemble_hessian(x0)
    b = asm.create_vec(np.random.default_rng(1).standard_normal(asm.layout.n_free))
    xg, its_g = _solve(asm, "gamg", 1e-10, b)
    _, its_j = _solve(asm, "jacobi", 1e-10, b)
    ref = spla.spsolve(asm.A.to_scipy().tocsc(), b.array)
    assert np.linalg.norm(xg.array - ref) <= 1e-6 * np.linalg.norm(ref)
    assert its_g * 4 < its_j, (its_g, its_j)
    asm.cleanup()


def test_vcycle_graph_replays_match_eager(require_gpu):
    """The V-cycle recorded as a HIP graph gives bitwise the eager result on every replay
    (several replays, with unrelated eager work in between)."""
    import torch

    p, adj, _ = build_plaplace_problem(64)
    asm = PLaplaceReorderedElementAssembler(p, None, adj, ksp_type="cg", pc_type="gamg")
    asm.assemble_hessian(asm.create_vec(np.random.default_rng(0).random(asm.layout.n_free)))
    ksp = asm.ksp
    ksp.setOperators(asm.A)
    ksp.setUp()
    pc = ksp._amg
    assert pc._graph is not None
    rng = np.random.default_rng(5)
    for k in range(4):
        r = torch.as_tensor(rng.standard_normal(asm.layout.n_free)).cuda()
        eager = pc._cycle(0, r.clone())
        junk = torch.randn(10000, device="cuda").cumsum(0)  # unrelated eager work
        z = torch.empty_like(r)
        pc.apply(r, z)
        torch.cuda.synchronize()
        assert torch.equal(z, eager), k
        assert torch.isfinite(junk).all()
    asm.cleanup()


@pytest.mark.parametrize("a_prec", ["fp64", "fp32"])
def test_spgemm_numeric_matches_scipy(require_gpu, a_prec):
    """fne_spgemm_numeric against scipy's product, with result rows longer than one LDS chunk
    (512 entries) and A rows longer than one staging chunk (256), with signed values (cancelled entries of the fixed pattern come out 0)."""
    import scipy.sparse as sp
    import torch

    from fenics_nonlinear_energies_amd import amg

    A = sp.random(70, 3000, density=0.1, random_state=3, format="csr")
    B = sp.random(3000, 4000, density=0.003, random_state=4, format="csr")
    A.data -= 0.5
    B.data -= 0.5
    if a_prec == "fp32":
        A.data = A.data.astype(np.float32).astype(np.float64)
    C_pat = amg._pattern_product(A, B)
    assert np.diff(C_pat.indptr).max() > 1024  # more than two LDS chunks of result columns
    assert np.diff(A.indptr).max() > 256  # more than one staging chunk of A's row
    dev = torch.device("cuda")
    Bd = amg._DevMat(B.indptr, B.indices, B.shape[1], dev, "fp64")
    Bd.data.copy_(torch.as_tensor(B.data))
    Cd = amg._DevMat(C_pat.indptr, C_pat.indices, C_pat.shape[1], dev, "fp64")
    a_ptr = torch.as_tensor(A.indptr.astype(np.int32)).to(dev)
    a_idx = torch.as_tensor(A.indices.astype(np.int32)).to(dev)
    a_dat = torch.as_tensor(A.data).to(dev, torch.float32 if a_prec == "fp32" else torch.float64)
    Cd.data.fill_(np.nan)
    amg._spgemm(a_ptr, a_idx, a_dat, Bd, Bd.data, Cd, Cd.data)
    got = sp.csr_matrix((Cd.data.cpu().numpy(), C_pat.indices, C_pat.indptr), shape=C_pat.shape).toarray()
    ref = (A @ B).toarray()
    assert np.all(np.isfinite(got))
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())
    # deterministic: a second product is bitwise the same
    first = Cd.data.clone()
    amg._spgemm(a_ptr, a_idx, a_dat, Bd, Bd.data, Cd, Cd.data)
    assert torch.equal(first, Cd.data)


def _host_levels(pc, A0):
    """Operators of every level read back from the device (FP64 masters)."""
    import scipy.sparse as sp

    out = []
    A = A0
    for L in pc.levels:
        P = L["P64"]
        Pm = sp.csr_matrix((P.data.cpu().numpy(), P.indices.cpu().numpy(), P.indptr.cpu().numpy()),
                           shape=(L["n"], L["nc"]))
        Ac = L["Ac64"]
        Acm = sp.csr_matrix((Ac.data.cpu().numpy(), Ac.indices.cpu().numpy(), Ac.indptr.cpu().numpy()),
                            shape=(L["nc"], L["nc"]))
        out.append((A, Pm, Acm, L["wd"].cpu().numpy()))
        A = Acm
    return out


@pytest.mark.parametrize("smoother", ["jacobi", "chebyshev"])
def test_device_galerkin_setup_matches_host_products(require_gpu, smoother):
    """Every level of the device set-up: P = P_tent - omega D^-1 A P_tent with P_tent from the
    host QR of the (current-configuration) rigid modes, A_c = P^T A P, omega within 15 % of the
    host power-iteration value of 4 / (3 rho(D^-1 A)) (60 steps); and the V-cycle equals a host cycle
    on the same operators, with either smoother (Chebyshev in the textbook form of Saad,
    Iterative Methods, Alg. 12.1, on [0.1, 1.1] lambda_max with Jacobi scaling)."""
    import torch

    from fenics_nonlinear_energies_amd import amg

    p, adj, u0 = build_beam_problem(40, 4, 4)
    asm = HEReorderedElementAssembler(p, None, adj, ksp_type="cg", pc_type="gamg")
    u = u0 + 1e-4 * np.random.default_rng(0).standard_normal(u0.size)
    x = asm.create_vec(u[asm.layout.perm])
    asm.assemble_hessian(x)
    B = asm._rigid_modes_current().cpu().numpy()
    np.testing.assert_array_equal(B, rigid_modes(x.t.cpu().numpy().reshape(-1, 3)))
    pc = amg.SmoothedAggregationPC(B, precision="fp64", max_coarse=100, smoother=smoother)
    pc.setup(asm.A)
    assert len(pc.levels) >= 2
    Bl = B
    for lvl, (A, P, Ac, wd) in enumerate(_host_levels(pc, asm.A.to_scipy())):
        Pt, Bc = amg.tentative_prolongator(Bl, pc.levels[lvl]["agg"])
        d = A.diagonal()
        omega = wd[0] * d[0]
        np.testing.assert_allclose(wd, omega / d, rtol=1e-12)
        amg_rho = None  # host estimate of rho(D^-1 A)
        v = np.random.default_rng(1).standard_normal(A.shape[0])
        for _ in range(60):
            y = (A @ v) / d
            amg_rho = np.linalg.norm(y) / np.linalg.norm(v)
            v = y / np.linalg.norm(y)
        rho = amg_rho
        # omega from a 12-step device Lanczos estimate x 1.1: near 4 / (3 rho), convergent smoother
        assert 0.85 <= omega * 3.0 * rho / 4.0 <= 1.15 and omega * rho < 2.0, (omega, rho)
        Pdev = pc.levels[lvl]["Pt"].data.cpu().numpy()  # device QR (fne_aggregate_qr)
        np.testing.assert_allclose(Pdev, Pt.data, rtol=0, atol=1e-12)
        np.testing.assert_allclose(pc.levels[lvl]["Bc"].cpu().numpy(), Bc, rtol=0, atol=1e-12 * np.abs(Bc).max())
        Pref = (Pt - omega * (A @ Pt).multiply(1.0 / d[:, None])).toarray()
        np.testing.assert_allclose(P.toarray(), Pref, rtol=0, atol=1e-12 * np.abs(Pref).max())
        Acref = (P.T @ A @ P).toarray()
        Acref = 0.5 * (Acref + Acref.T)
        np.testing.assert_allclose(Ac.toarray(), Acref, rtol=0, atol=1e-12 * np.abs(Acref).max())
        Bl = Bc
    # V-cycle against the host restatement on the read-back operators
    levels = _host_levels(pc, asm.A.to_scipy())
    Ainv = pc.coarse_op.data.cpu().numpy().reshape(pc._coarse_n, pc._coarse_n)

    def cheb(lvl, A, r, x):
        # Saad Alg. 12.1 with M = D: d_0 = M^-1 r_0 / theta, d_k = rho_k rho_{k-1} d_{k-1}
        # + (2 rho_k / delta) M^-1 r_k, x_{k+1} = x_k + d_k
        dinv = 1.0 / A.diagonal()
        lam = 1.0 / (0.6 * float(pc.levels[lvl]["wcheb"][0, 0].cpu()) / dinv[0])
        a, b = 0.1 * lam, 1.1 * lam
        theta, delta = 0.5 * (b + a), 0.5 * (b - a)
        sigma = theta / delta
        rho = 1.0 / sigma
        dk = dinv * (r - A @ x) / theta
        x = x + dk
        for _ in range(pc.sweeps - 1):
            rho_new = 1.0 / (2.0 * sigma - rho)
            dk = rho_new * rho * dk + (2.0 * rho_new / delta) * dinv * (r - A @ x)
            x = x + dk
            rho = rho_new
        return x

    def cycle(lvl, r):
        if lvl == len(levels):
            return Ainv @ r
        A, P, _, wd = levels[lvl]
        if smoother == "chebyshev":
            xx = cheb(lvl, A, r, np.zeros_like(r))
        else:
            xx = wd * r
            for _ in range(pc.sweeps - 1):
                xx = xx + wd * (r - A @ xx)
        xx = xx + P @ cycle(lvl + 1, P.T @ (r - A @ xx))
        if smoother == "chebyshev":
            return cheb(lvl, A, r, xx)
        for _ in range(pc.sweeps):
            xx = xx + wd * (r - A @ xx)
        return xx

    r = np.random.default_rng(2).standard_normal(asm.layout.n_free)
    z = torch.empty(r.size, dtype=torch.float64, device="cuda")
    pc.apply(torch.as_tensor(r).cuda(), z)
    ref = cycle(0, r)
    np.testing.assert_allclose(z.cpu().numpy(), ref, rtol=0, atol=1e-10 * np.abs(ref).max())
    asm.cleanup()


def test_resetup_reuses_graph_and_follows_the_hessian(require_gpu):
    """A second set-up at another state rewrites the recorded cycle's buffers in place: the
    graph object is kept and its replays equal the eager cycle on the new operators."""
    import torch

    p, adj, u0 = build_beam_problem(40, 4, 4)
    asm = HEReorderedElementAssembler(p, None, adj, ksp_type="cg", pc_type="gamg")
    ksp = asm.ksp
    rng = np.random.default_rng(3)
    graph = None
    for k, scale in enumerate((1e-5, 3e-3)):
        x = asm.create_vec((u0 + scale * rng.standard_normal(u0.size))[asm.layout.perm])
        asm.assemble_hessian(x)
        ksp.setOperators(asm.A)
        ksp.setUp()
        pc = ksp._amg
        if graph is None:
            graph = pc._graph
        assert pc._graph is graph and pc.setup_count == k + 1
        r = torch.as_tensor(rng.standard_normal(asm.layout.n_free)).cuda()
        eager = pc._cycle(0, r.clone())
        z = torch.empty_like(r)
        pc.apply(r, z)
        torch.cuda.synchronize()
        assert torch.equal(z, eager), k
    asm.cleanup()


def test_device_cg_with_graph_preconditioner_matches_host_cg(require_gpu):
    """fne_pcg_graph (device iteration launching the recorded V-cycle) against the host-driven
    KSPCG restatement with the same preconditioner: same iteration count and reason, and the
    same solution to 1e-10."""
    import torch

    p, adj, u0 = build_beam_problem(40, 4, 4)
    asm = HEReorderedElementAssembler(p, None, adj, ksp_type="cg", pc_type="gamg")
    x = asm.create_vec((u0 + 1e-5 * np.random.default_rng(4).standard_normal(u0.size))[asm.layout.perm])
    asm.assemble_hessian(x)
    ksp = asm.ksp
    ksp.setOperators(asm.A)
    ksp.setUp()
    assert ksp._amg._graph is not None
    b = asm.create_vec(np.random.default_rng(5).standard_normal(asm.layout.n_free))
    for rtol in (1e-1, 1e-8):
        ksp.setTolerances(rtol=rtol, max_it=500)
        xd = b.duplicate()
        ksp.solve(b, xd)
        its_d, reason_d = ksp.getIterationNumber(), ksp.getConvergedReason()
        xh = torch.empty_like(b.t)
        ksp._solve_cg_host(b.t, xh)
        assert (its_d, reason_d) == (ksp.getIterationNumber(), ksp.getConvergedReason()), (its_d, reason_d)
        assert reason_d > 0
        torch.testing.assert_close(xd.t, xh, rtol=1e-10, atol=1e-10 * float(xh.abs().max()))
    asm.cleanup()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fused_vcycle_ops_match_numpy(require_gpu, prec):
    """fne_vcycle_op: first sweep, sweep, residual and in-place prolongation-add against numpy
    on a random CSR operator (FP32 values are compared with their FP32-rounded numpy twin)."""
    import scipy.sparse as sp
    import torch

    from fenics_nonlinear_energies_amd import amg

    rng = np.random.default_rng(7)
    A = sp.random(500, 500, density=0.03, random_state=8, format="csr") + sp.eye(500, format="csr")
    A.sort_indices()
    P = sp.random(500, 80, density=0.05, random_state=9, format="csr")
    P.sort_indices()
    if prec == "fp32":
        A.data = A.data.astype(np.float32).astype(np.float64)
        P.data = P.data.astype(np.float32).astype(np.float64)
    dev = torch.device("cuda")
    Ad = amg._DevMat(A.indptr, A.indices, 500, dev, prec)
    Ad.data.copy_(torch.as_tensor(A.data))
    Pd = amg._DevMat(P.indptr, P.indices, 80, dev, prec)
    Pd.data.copy_(torch.as_tensor(P.data))
    x, r, w = (rng.standard_normal(500) for _ in range(3))
    xc = rng.standard_normal(80)
    T = lambda v: torch.as_tensor(v).to(dev)  # noqa: E731
    y = torch.empty(500, dtype=torch.float64, device=dev)
    tol = dict(rtol=1e-12, atol=1e-12)
    Ad.op(amg._SWEEP_FIRST, None, T(r), T(w), None, y)
    x1 = w * r
    np.testing.assert_allclose(y.cpu().numpy(), x1 + w * (r - A @ x1), **tol)
    Ad.op(amg._SWEEP, T(x), T(r), T(w), None, y)
    np.testing.assert_allclose(y.cpu().numpy(), x + w * (r - A @ x), **tol)
    Ad.op(amg._RESIDUAL, T(x), T(r), None, None, y)
    np.testing.assert_allclose(y.cpu().numpy(), r - A @ x, **tol)
    yy = T(x)
    Pd.op(amg._ADD_PRODUCT, T(xc), None, None, yy, yy)
    np.testing.assert_allclose(yy.cpu().numpy(), x + P @ xc, **tol)
    from fenics_nonlinear_energies_amd.native import FneError

    with pytest.raises(FneError):
        Ad.op(amg._SWEEP, yy, T(r), T(w), None, yy)  # in-place sweep is rejected


def test_fp32_multigrid_with_fp64_operator(require_gpu):
    """pc_precision="fp32" with an FP64 Hessian: the hierarchy (finest copy included) is stored
    in FP32 while CG runs on the FP64 operator, and still reaches the direct solution."""
    import torch

    p, adj, u0 = build_beam_problem(40, 4, 4)
    asm = HEReorderedElementAssembler(p, None, adj, ksp_type="cg", pc_type="gamg", pc_precision="fp32")
    u = u0 + 1e-5 * np.random.default_rng(0).standard_normal(u0.size)
    asm.assemble_hessian(asm.create_vec(u[asm.layout.perm]))
    assert asm.A.precision == "fp64"
    b = asm.create_vec(np.random.default_rng(1).standard_normal(asm.layout.n_free))
    xg, its = _solve(asm, "gamg", 1e-10, b)
    pc = asm.ksp._amg
    assert pc.precision == "fp32" and pc.levels[0]["A"].data.dtype == torch.float32
    assert asm.ksp.getConvergedReason() > 0 and its < 200
    ref = spla.spsolve(asm.A.to_scipy().tocsc(), b.array)
    assert np.linalg.norm(xg.array - ref) <= 1e-6 * np.linalg.norm(ref)
    asm.cleanup()


@pytest.mark.parametrize("prec,density", [("fp64", 0.03), ("fp32", 0.03), ("fp64", 0.12)])
def test_chebyshev_step_matches_numpy(require_gpu, prec, density):
    """fne_vcycle_cheb: y = x + c (x - x_prev) + w.(r - A x) against numpy (the 0.12 density gives
    60-entry rows, the grouped-load path), in place over x_prev, and with x_prev absent for
    c = 0; an output aliasing x is rejected."""
    import scipy.sparse as sp
    import torch

    from fenics_nonlinear_energies_amd import amg
    from fenics_nonlinear_energies_amd.native import FneError

    rng = np.random.default_rng(11)
    A = sp.random(500, 500, density=density, random_state=12, format="csr") + sp.eye(500, format="csr")
    A.sort_indices()
    if prec == "fp32":
        A.data = A.data.astype(np.float32).astype(np.float64)
    dev = torch.device("cuda")
    Ad = amg._DevMat(A.indptr, A.indices, 500, dev, prec)
    Ad.data.copy_(torch.as_tensor(A.data))
    x, xp, r, w = (rng.standard_normal(500) for _ in range(4))
    T = lambda v: torch.as_tensor(v).to(dev)  # noqa: E731
    tol = dict(rtol=1e-12, atol=1e-12)
    y = torch.empty(500, dtype=torch.float64, device=dev)
    Ad.cheb(T(x), T(xp), T(r), T(w), 0.7, y)
    np.testing.assert_allclose(y.cpu().numpy(), x + 0.7 * (x - xp) + w * (r - A @ x), **tol)
    Ad.cheb(T(x), None, T(r), T(w), 0.0, y)
    np.testing.assert_allclose(y.cpu().numpy(), x + w * (r - A @ x), **tol)
    yp = T(xp)
    Ad.cheb(T(x), yp, T(r), T(w), 0.3, yp)  # in place over x_prev
    np.testing.assert_allclose(yp.cpu().numpy(), x + 0.3 * (x - xp) + w * (r - A @ x), **tol)
    xx = T(x)
    with pytest.raises(FneError):
        Ad.cheb(xx, T(xp), T(r), T(w), 0.5, xx)


@pytest.mark.parametrize("kind", ["he", "slope"])
def test_chebyshev_smoothed_multigrid_cg(require_gpu, kind):
    """CG with the Chebyshev-smoothed multigrid (mg_levels_ksp_type chebyshev, 2 steps) reaches
    the direct solution; its graph replays equal the eager cycle; iteration counts of both
    smoothers are printed."""
    import torch

    from fenics_nonlinear_energies_amd.problems.slope_stability import SlopeStabilityReorderedElementAssembler
    from tests.test_gpu_mc import _state
    from tests.test_oracle_mc import mc_params

    its = {}
    for smoother in ("richardson", "chebyshev"):
        opts = {"mg_levels_ksp_type": smoother, "mg_levels_ksp_max_it": 2}
        if kind == "he":
            p, adj, u0 = build_beam_problem(40, 4, 4)
            asm = HEReorderedElementAssembler(p, None, adj, ksp_type="cg", pc_type="gamg", pc_options=opts)
            x0 = asm.create_vec((u0 + 1e-5 * np.random.default_rng(0).standard_normal(u0.size))[asm.layout.perm])
        else:
            p, adj = mc_params(h=1.0, degree=4, lam=1.0)
            asm = SlopeStabilityReorderedElementAssembler(p, None, adj, hessian_precision="fp64", ksp_type="cg",
                                                          pc_type="mg", pc_options=opts)
            x0 = asm.create_vec(_state(p, asm.layout, 11)[asm.layout.perm])
        asm.assemble_hessian(x0)
        b = asm.create_vec(np.random.default_rng(1).standard_normal(asm.layout.n_free))
        ksp = asm.ksp
        ksp.setOperators(asm.A)
        ksp.setTolerances(rtol=1e-10, max_it=5000)
        ksp.setUp()
        assert ksp._amg.smoother == ("chebyshev" if smoother == "chebyshev" else "jacobi")
        x = b.duplicate()
        ksp.solve(b, x)
        assert ksp.getConvergedReason() > 0
        its[smoother] = ksp.getIterationNumber()
        ref = spla.spsolve(asm.A.to_scipy().tocsc(), b.array)
        assert np.linalg.norm(x.array - ref) <= 1e-6 * np.linalg.norm(ref)
        pc = ksp._amg
        r = torch.as_tensor(np.random.default_rng(2).standard_normal(asm.layout.n_free)).cuda()
        eager = pc._cycle(0, r.clone())
        z = torch.empty_like(r)
        pc.apply(r, z)
        torch.cuda.synchronize()
        assert torch.equal(z, eager)
        asm.cleanup()
    print(kind, "CG iterations", its, flush=True)
