"""GPU parity of the P1 kernels (p-Laplace, Ginzburg-Landau) against the CPU oracle.

Tolerances (FP64 path): energy rel 1e-13, gradient and Hessian rel 1e-12 of their max
entry (summation order differs from np.add.at).  FP32 Hessian values are exactly the
FP64 values rounded to float32 (accumulation is FP64 in LDS).
"""

import numpy as np
import pytest
import torch

from fenics_nonlinear_energies_amd.problems.ginzburg_landau import GLReorderedElementAssembler, build_gl_problem
from fenics_nonlinear_energies_amd.problems.plaplace import PLaplaceReorderedElementAssembler, build_plaplace_problem
from oracle import assembly
from tests import fixtures

pytestmark = pytest.mark.gpu


def _check_all(asm, kind, d, u_orig, rtol_e=1e-13, rtol=1e-12):
    lay = asm.layout
    x = asm.create_vec(u_orig[lay.perm])
    ref = assembly.assemble(kind, d, u_orig)
    E = asm.energy_fn(x)
    assert abs(E - ref["e"]) <= rtol_e * max(1.0, abs(ref["e"]))
    g = x.duplicate()
    asm.gradient_fn(x, g)
    gg = g.array[lay.iperm]
    assert np.max(np.abs(gg - ref["g"])) <= rtol * np.max(np.abs(ref["g"]))
    asm.assemble_hessian(x)
    H = asm.A.to_scipy()[lay.iperm][:, lay.iperm]
    Href = ref["h"]
    assert abs(H - Href).max() <= rtol * abs(Href).max()
    return H


@pytest.mark.parametrize(
    "mode,amode",
    [("block_morton", "patch"), ("block_xyz", "patch"), ("block_morton", "staged"), ("block_morton", "rowgather")],
)
def test_plaplace_reference_mesh_parity(require_gpu, mode, amode):
    d, adj = fixtures.plaplace(5)
    asm = PLaplaceReorderedElementAssembler(d, None, adj, reorder_mode=mode, assembly_mode=amode)
    u = np.random.default_rng(11).random(d["freedofs"].size)
    _check_all(asm, "plaplace", d, u)
    asm.cleanup()


@pytest.mark.parametrize("amode", ["patch", "staged", "rowgather"])
def test_gl_reference_mesh_parity(require_gpu, amode):
    d, adj = fixtures.gl(5)
    asm = GLReorderedElementAssembler(d, None, adj, assembly_mode=amode)
    u = np.random.default_rng(12).uniform(-1, 1, d["freedofs"].size)
    _check_all(asm, "gl", d, u)
    asm.cleanup()


def test_fp32_hessian_close_to_fp64_and_deterministic(require_gpu):
    """The staged kernel sums the (at most 6) FP32 element contributions of a CSR entry in
    FP32: within a few FP32 ulps of the FP64 assembly."""
    params, adj, _ = build_plaplace_problem(64)
    u = np.random.default_rng(5).random(params["freedofs"].size)
    a64 = PLaplaceReorderedElementAssembler(params, None, adj, hessian_precision="fp64")
    a32 = PLaplaceReorderedElementAssembler(params, None, adj, hessian_precision="fp32")
    x = a64.create_vec(u[a64.layout.perm])
    a64.assemble_hessian(x)
    a32.assemble_hessian(x)
    v64 = a64.A.data.clone()
    d = (a32.A.data.double() - v64).abs()
    assert float(d.max()) <= 8 * 2.0 ** -24 * float(v64.abs().max())
    v32 = a32.A.data.clone()
    a32.assemble_hessian(x)
    assert torch.equal(a32.A.data, v32)
    a64.assemble_hessian(x)
    assert torch.equal(a64.A.data, v64)  # bitwise reproducible (no atomics)
    e1 = a64.energy_fn(x)
    e2 = a64.energy_fn(x)
    assert e1 == e2


def test_element_eval_matches_oracle(require_gpu):
    d, adj = fixtures.gl(2)
    asm = GLReorderedElementAssembler(d, None, adj)
    u = np.random.default_rng(2).uniform(-1, 1, d["freedofs"].size)
    e, g, h = asm.element_eval(asm.create_vec(u[asm.layout.perm]))
    q = assembly.element_eval("gl", d, assembly.local_state(d, u))
    assert np.allclose(e.cpu().numpy(), q["e"], rtol=1e-14, atol=1e-17)
    assert np.allclose(g.cpu().numpy(), q["g"], rtol=1e-13, atol=1e-15)
    assert np.allclose(h.cpu().numpy(), q["h"], rtol=1e-13, atol=1e-13)


def test_zero_state_and_dirichlet_elements(require_gpu):
    """u = 0: p=3 Hessian is exactly 0 (limit), GL energy counts Dirichlet-only elements."""
    params, adj, u0 = build_plaplace_problem(16)
    asm = PLaplaceReorderedElementAssembler(params, None, adj)
    x = asm.create_vec(u0)
    asm.assemble_hessian(x)
    assert float(asm.A.data.abs().max()) == 0.0
    assert asm.energy_fn(x) == 0.0
    gp, gadj = build_gl_problem(8)
    gasm = GLReorderedElementAssembler(gp, None, gadj)
    E0 = gasm.energy_fn(gasm.create_vec(np.zeros(gp["freedofs"].size)))
    assert np.isclose(E0, 0.25 * gp["vol"].sum(), rtol=1e-14)  # (0-1)^2/4 * area


def test_spmv_matches_scipy(require_gpu):
    params, adj, _ = build_plaplace_problem(48)
    asm = PLaplaceReorderedElementAssembler(params, None, adj, hessian_precision="fp32")
    u = np.random.default_rng(3).random(params["freedofs"].size)
    asm.assemble_hessian(asm.create_vec(u))
    A = asm.A.to_scipy()
    xin = np.random.default_rng(4).standard_normal(A.shape[0])
    y = asm.create_vec(np.zeros(A.shape[0]))
    asm.A.mult(asm.create_vec(xin), y)
    ref = A @ xin
    assert np.allclose(y.array, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("amode", ["patch", "staged", "rowgather"])
def test_fused_energy_grad_hessian_matches_separate_calls(require_gpu, amode):
    """fne_energy_grad_hessian == fne_energy + fne_gradient + fne_hessian (GL has only-constrained
    corner elements, whose energy the fused pass must still count)."""
    gp, gadj = build_gl_problem(24)
    asm = GLReorderedElementAssembler(gp, None, gadj, hessian_precision="fp32", assembly_mode=amode)
    u = np.random.default_rng(8).uniform(-1, 1, gp["freedofs"].size)
    x = asm.create_vec(u[asm.layout.perm])
    g1, g2 = x.duplicate(), x.duplicate()
    e_sep = float(asm.energy_device(x.t).item())
    asm.gradient_device(x.t, g1.t)
    asm.hessian_device(x.t)
    h_sep = asm.A.data.clone()
    e_fused = float(asm.energy_grad_hessian_device(x.t, g2.t).item())
    assert abs(e_fused - e_sep) <= 1e-14 * abs(e_sep)
    assert torch.equal(g1.t, g2.t)
    assert torch.equal(h_sep, asm.A.data)
    ref = assembly.assemble("gl", gp, u)
    assert abs(e_fused - ref["e"]) <= 1e-13 * abs(ref["e"])


@pytest.mark.parametrize("amode", ["patch", "staged", "rowgather"])
def test_fused_energy_counts_the_load_term(require_gpu, amode):
    """p-Laplace with f = -10: the fused pass's energy includes -f.u (each free row once), equal to
    the separate energy call and the oracle; repeated fused calls give the same bits (the in-kernel
    finalisation merges the workgroup pairs in index order and re-arms its counter)."""
    d, adj, _ = build_plaplace_problem(40)
    asm = PLaplaceReorderedElementAssembler(d, None, adj, hessian_precision="fp64", assembly_mode=amode)
    u = np.random.default_rng(11).uniform(0.0, 1.0, d["freedofs"].size)
    x = asm.create_vec(u[asm.layout.perm])
    g = x.duplicate()
    e_sep = float(asm.energy_device(x.t).item())
    e1 = float(asm.energy_grad_hessian_device(x.t, g.t).item())
    e2 = float(asm.energy_grad_hessian_device(x.t, g.t).item())
    ref = assembly.assemble("plaplace", d, u)
    assert abs(e1 - ref["e"]) <= 1e-13 * abs(ref["e"])
    assert abs(e1 - e_sep) <= 1e-14 * abs(e_sep)
    assert e1 == e2


@pytest.mark.parametrize("family", ["plaplace", "gl"])
@pytest.mark.parametrize("amode", ["patch", "staged", "rowgather"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_node_geometry_bitwise_equal_to_geometry_arrays(require_gpu, family, amode, prec):
    """FNE_FLAG_NODE_GEO (element geometry from the vertex coordinates in the kernels) gives
    the same bits as reading the geometry arrays: energy, gradient, CSR values, the fused pass,
    element outputs and (staged plans) Hessian-vector products, on the reference's level-5 mesh."""
    d, adj = fixtures.plaplace(5) if family == "plaplace" else fixtures.gl(5)
    cls = PLaplaceReorderedElementAssembler if family == "plaplace" else GLReorderedElementAssembler
    a_arr = cls(d, None, adj, assembly_mode=amode, hessian_precision=prec, element_geometry="arrays")
    a_nod = cls(d, None, adj, assembly_mode=amode, hessian_precision=prec, element_geometry="nodes")
    assert not a_arr.node_geometry and a_nod.node_geometry
    assert "egeo" not in a_nod._dev and "pgeo" not in a_nod._dev
    rng = np.random.default_rng(21)
    u = rng.random(d["freedofs"].size) if family == "plaplace" else rng.uniform(-1, 1, d["freedofs"].size)
    xa = a_arr.create_vec(u[a_arr.layout.perm])
    xn = a_nod.create_vec(u[a_nod.layout.perm])
    assert torch.equal(xa.t, xn.t)
    assert a_arr.energy_fn(xa) == a_nod.energy_fn(xn)
    ga, gn = xa.duplicate(), xn.duplicate()
    a_arr.gradient_fn(xa, ga)
    a_nod.gradient_fn(xn, gn)
    assert torch.equal(ga.t, gn.t)
    a_arr.assemble_hessian(xa)
    a_nod.assemble_hessian(xn)
    assert torch.equal(a_arr.A.data, a_nod.A.data)
    ea = float(a_arr.energy_grad_hessian_device(xa.t, ga.t).item())
    en = float(a_nod.energy_grad_hessian_device(xn.t, gn.t).item())
    assert ea == en and torch.equal(ga.t, gn.t) and torch.equal(a_arr.A.data, a_nod.A.data)
    for ta, tn in zip(a_arr.element_eval(xa), a_nod.element_eval(xn)):
        assert torch.equal(ta, tn)
    if amode == "staged":
        v = torch.as_tensor(rng.standard_normal(xa.t.numel()), device=xa.t.device)
        ya, yn = torch.empty_like(v), torch.empty_like(v)
        a_arr.hvp_device(xa.t, v, ya)
        a_nod.hvp_device(xn.t, v, yn)
        assert torch.equal(ya, yn)
    a_arr.cleanup()
    a_nod.cleanup()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_fused_energy_reduction_is_exact_and_reproducible_over_many_calls(require_gpu, prec):
    """The fused staged pass: over 300 back-to-back calls on a grid with ~1,300 patches every
    energy (per-patch pairs merged in index order by the finalisation kernel) has the same bits,
    equal to the separate energy call to 1e-14, and the Hessian and gradient bits do not move
    either."""
    gp, gadj = build_gl_problem(512)
    asm = GLReorderedElementAssembler(gp, None, gadj, hessian_precision=prec, assembly_mode="staged")
    u = np.random.default_rng(21).uniform(-1, 1, gp["freedofs"].size)
    x = asm.create_vec(u[asm.layout.perm])
    g = x.duplicate()
    e_sep = float(asm.energy_device(x.t).item())
    outs = torch.empty(300, dtype=torch.float64, device=x.t.device)
    asm.energy_grad_hessian_device(x.t, g.t)
    h0, g0 = asm.A.data.clone(), g.t.clone()
    for k in range(300):
        asm.energy_grad_hessian_device(x.t, g.t, out=outs[k:k + 1])
    e = outs.cpu().numpy()
    assert np.all(e == e[0]), np.unique(e)
    assert abs(e[0] - e_sep) <= 1e-14 * abs(e_sep)
    assert torch.equal(h0, asm.A.data) and torch.equal(g0, g.t)
    asm.cleanup()
