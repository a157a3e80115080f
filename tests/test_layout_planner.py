"""Host set-up: permutations, CSR pattern and the P1 patch schedule (CPU only).

The patch schedule is executed here by a numpy emulation of the kernel's phase B
(incidence decode + fixed-order accumulation) and compared against the oracle's
reference-style np.add.at assembly and a brute-force dense double loop.
"""

import numpy as np
import pytest
import scipy.sparse as sp

from fenics_nonlinear_energies_amd import layout, patches
from fenics_nonlinear_energies_amd.problems import mesh2d
from fenics_nonlinear_energies_amd.problems.plaplace import build_plaplace_problem
from oracle import assembly
from tests import fixtures


# --- reference unit tests (tests/test_reordered_element_base.py:12-73) -------------------
def test_block_xyz_scalar_golden():
    coords = np.array([[1.0, 0.0], [0.0, 1.0], [0.0, 0.0], [1.0, 1.0]])
    perm = layout.select_permutation(
        "block_xyz", adjacency=sp.eye(4, format="csr"), coords_all=coords, freedofs=np.arange(4), block_size=1
    )
    assert perm.tolist() == [2, 1, 0, 3]


def test_block_xyz_vector_golden():
    coords = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    perm = layout.select_permutation(
        "block_xyz", adjacency=sp.eye(6, format="csr"), coords_all=coords, freedofs=np.arange(6), block_size=3
    )
    assert perm.tolist() == [3, 4, 5, 0, 1, 2]


def test_local_vec_from_full_golden():
    v = layout.local_vec_from_full(
        np.array([10.0, 20.0, 30.0]),
        np.array([-1, 0, 1, -1, 2]),
        np.arange(5),
        np.array([5.0, -1.0, -1.0, 6.0, -1.0]),
    )
    assert np.allclose(v, [5.0, 10.0, 20.0, 6.0, 30.0])


def test_morton_is_permutation_and_keeps_node_components_adjacent():
    coords = np.random.default_rng(1).random((50, 3))
    free = np.array([3 * n + c for n in range(50) for c in range(3) if not (n % 7 == 0 and c == 1)])
    perm = layout.perm_block_morton(coords, free, 3)
    assert np.array_equal(np.sort(perm), np.arange(free.size))
    nodes = free[perm] // 3
    # each node's free components are consecutive in the new order
    changes = np.count_nonzero(np.diff(nodes) != 0) + 1
    assert changes == np.unique(nodes).size


# --- mesh generator closed forms --------------------------------------------------------
def test_structured_mesh_counts_and_area():
    nodes, elems = mesh2d.structured_triangle_mesh(7, 5, 0.0, 2.0, -1.0, 1.0)
    assert nodes.shape == (8 * 6, 2) and elems.shape == (2 * 35, 3)
    dvx, dvy, vol = mesh2d.triangle_operators(nodes, elems)
    assert np.isclose(vol.sum(), 4.0)
    # gradients of the P1 basis sum to zero and reproduce linear functions exactly
    assert np.allclose(dvx.sum(1), 0) and np.allclose(dvy.sum(1), 0)
    lin = 3.0 * nodes[:, 0] - 2.0 * nodes[:, 1]
    assert np.allclose((lin[elems] * dvx).sum(1), 3.0) and np.allclose((lin[elems] * dvy).sum(1), -2.0)


def test_reference_mesh_geometry_consistent():
    """The HDF5-decoded reference meshes: stored dvx/dvy/vol equal those recomputed from nodes."""
    for loader, lev in ((fixtures.plaplace, 5), (fixtures.gl, 5)):
        d, adj = loader(lev)
        dvx, dvy, vol = mesh2d.triangle_operators(d["nodes"], d["elems"])
        assert np.allclose(dvx, d["dvx"], rtol=1e-12, atol=1e-9)
        assert np.allclose(dvy, d["dvy"], rtol=1e-12, atol=1e-9)
        assert np.allclose(vol, d["vol"], rtol=1e-12)
        assert adj.shape == (d["freedofs"].size,) * 2
    d, _ = fixtures.plaplace(5)
    # the stored load is the P1 load of f = -10 (lumped vol/3 per vertex)
    f = mesh2d.lumped_nodal_load(d["elems"], d["vol"], -10.0, d["nodes"].shape[0])
    assert np.allclose(f, d["f"], rtol=1e-12, atol=1e-15)


# --- patch schedule ----------------------------------------------------------------------
def emulate_p1_colored(plan, ef_new, ge_new, he_new, indptr, n_free):
    """numpy re-statement of p1_color_kernel: per patch, color by color, owned rows only."""
    pid = np.repeat(np.arange(plan.n_patch), np.diff(plan.pelem_ptr))
    e = plan.pelem.astype(np.int64)
    item = plan.pitem.astype(np.uint64)
    color = ((item >> np.uint64(52)) & np.uint64(255)).astype(np.int64)
    g = np.zeros(n_free)
    vals = np.zeros(int(indptr[-1]))
    # race-freedom: within (patch, color) no owned row appears twice
    for a in range(3):
        r = ef_new[e, a]
        own = (r >= plan.prow[pid]) & (r < plan.prow[pid + 1])
        np.add.at(g, r[own], ge_new[e[own], a])
        for b in range(3):
            pos = ((item >> np.uint64(5 * (3 * a + b))) & np.uint64(31)).astype(np.int64)
            ok = own & (pos != 31)
            assert np.all(ok == (own & (ef_new[e, b] >= 0)))
            np.add.at(vals, indptr[r[ok]] + pos[ok], he_new[e[ok], a, b])
    rows_all = np.concatenate([np.where((ef_new[e, a] >= plan.prow[pid]) & (ef_new[e, a] < plan.prow[pid + 1]), ef_new[e, a], -1)[:, None] for a in range(3)], axis=1)
    keys = np.concatenate([(color * n_free + rows_all[:, a])[rows_all[:, a] >= 0] * plan.n_patch + pid[rows_all[:, a] >= 0] for a in range(3)])
    assert np.unique(keys).size == keys.size, "two elements of one color share an owned row"
    return g, vals


@pytest.mark.parametrize("mode", ["block_morton", "block_xyz", "none"])
@pytest.mark.parametrize("rows_per_patch,ept", [(64, 1), (224, 2)])
def test_p1_colored_plan_reproduces_oracle_assembly(mode, rows_per_patch, ept):
    d, adj = fixtures.plaplace(5)
    n_free = d["freedofs"].size
    perm = layout.select_permutation(mode, adjacency=adj, coords_all=d["nodes"], freedofs=d["freedofs"], block_size=1)
    lay = layout.build_global_layout(d, adj, perm, block_size=1, dirichlet_key="u_0")
    ef = lay.total_to_free_reord[d["elems"]]
    while True:  # shrink the patch until its elements fit ept * 256 (as the assembler does)
        try:
            plan = patches.build_p1_colored_plan(ef, lay.indptr, lay.indices, n_free,
                                                 rows_per_patch=rows_per_patch, elems_per_thread=ept)
            break
        except ValueError as exc:
            assert "more elements" in str(exc)
            rows_per_patch -= 16
    # the patch lists hold exactly the (patch, element) pairs with a row of the element in the patch
    ef_new = ef[plan.elem_order]
    e_i, a_i = np.nonzero(ef_new >= 0)
    want = np.unique((ef_new[e_i, a_i] // rows_per_patch) * ef.shape[0] + e_i)
    got_p = np.repeat(np.arange(plan.n_patch), np.diff(plan.pelem_ptr))
    got_all = got_p.astype(np.int64) * ef.shape[0] + plan.pelem
    lone = np.nonzero(~(ef_new >= 0).any(axis=1))[0]  # only-constrained elements: energy carriers
    is_lone = np.isin(plan.pelem, lone)
    assert np.array_equal(want, np.sort(got_all[~is_lone]))
    assert np.array_equal(np.sort(plan.pelem[is_lone]), lone)
    # elements of a patch are sorted by color
    col = ((plan.pitem.astype(np.uint64) >> np.uint64(52)) & np.uint64(255)).astype(np.int64)
    assert np.all(np.diff(col + 1000 * got_p) >= 0)
    # each element is the energy owner in exactly one patch
    owner = ((plan.pitem.astype(np.uint64) >> np.uint64(48)) & np.uint64(1)).astype(bool)
    assert np.array_equal(np.sort(plan.pelem[owner]), np.arange(ef.shape[0]))
    rng = np.random.default_rng(7)
    u = rng.random(n_free)
    ref = assembly.assemble("plaplace", d, u)
    q = assembly.element_eval("plaplace", d, assembly.local_state(d, u))
    order = plan.elem_order
    g_emul, vals = emulate_p1_colored(plan, ef_new, q["g"][order], q["h"][order], lay.indptr, n_free)
    g_emul -= d["f"][d["freedofs"]][perm]
    H = sp.csr_matrix((vals, lay.indices, lay.indptr), shape=(n_free, n_free))[lay.iperm][:, lay.iperm]
    assert abs(H - ref["h"]).max() <= 1e-12 * abs(ref["h"]).max()
    assert np.allclose(g_emul[lay.iperm], ref["g"], rtol=1e-12, atol=1e-14)
    assert plan.pcolors.max() <= 16


def test_oracle_assembly_matches_bruteforce():
    d, _ = fixtures.plaplace(1)
    u = np.random.default_rng(0).random(d["freedofs"].size)
    H = assembly.assemble("plaplace", d, u)["h"].toarray()
    Hb = assembly.dense_hessian_bruteforce("plaplace", d, u)
    assert np.allclose(H, Hb, rtol=1e-14, atol=1e-14)
    assert np.allclose(H, H.T, rtol=1e-13, atol=1e-14)


def test_patch_halo_is_small_for_morton_on_structured_grid():
    params, adj, _ = build_plaplace_problem(64)
    n_free = params["freedofs"].size
    perm = layout.select_permutation("block_morton", adjacency=adj, coords_all=params["nodes"], freedofs=params["freedofs"])
    lay = layout.build_global_layout(params, adj, perm, block_size=1, dirichlet_key="u_0")
    plan = patches.build_p1_colored_plan(lay.total_to_free_reord[params["elems"]], lay.indptr, lay.indices, n_free,
                                         rows_per_patch=192)
    assert plan.halo_factor < 1.35


def emulate_block_patch_assembly(plan, ef_new, ge_new, he_new, indptr, n_free, bs):
    """numpy re-statement of tet_patch_kernel's phase B (scalar row = node row * bs + comp)."""
    items = plan.inc.astype(np.uint64)
    nrows = np.repeat(np.arange(n_free // bs), np.diff(plan.inc_ptr))
    slot = (items & np.uint64(0xFFFF)).astype(np.int64)
    a = ((items >> np.uint64(16)) & np.uint64(3)).astype(np.int64)
    patch = (nrows * bs) // plan.rows_per_patch
    e = plan.pelem[plan.pelem_ptr[patch] + slot]
    n_nodes_e = ef_new.shape[1] // bs
    g = np.zeros(n_free)
    vals = np.zeros(int(indptr[-1]))
    for comp in range(bs):
        r = nrows * bs + comp
        np.add.at(g, r, ge_new[e, bs * a + comp])
        for b in range(n_nodes_e):
            bp = ((items >> np.uint64(24 + 8 * b)) & np.uint64(255)).astype(np.int64)
            ok = bp != 255
            for k in range(bs):
                np.add.at(vals, indptr[r[ok]] + bs * bp[ok] + k, he_new[e[ok], bs * a[ok] + comp, bs * b + k])
    return g, vals


def test_block_plan_reproduces_oracle_assembly_he():
    from fenics_nonlinear_energies_amd.problems.hyperelasticity import build_beam_problem

    p, adj, u0 = build_beam_problem(12, 3, 2)
    n_free = p["freedofs"].size
    perm = layout.select_permutation("block_morton", adjacency=adj, coords_all=p["nodes2coord"], freedofs=p["freedofs"], block_size=3)
    lay = layout.build_global_layout(p, adj, perm, block_size=3, dirichlet_key="u_0_ref")
    ef = lay.total_to_free_reord[p["elems"]]
    plan = patches.build_block_plan(ef, lay.indptr, lay.indices, n_free, block_size=3, nodes_per_patch=10, chunk=32)
    u = u0 + 1e-4 * np.random.default_rng(3).standard_normal(n_free)
    ref = assembly.assemble("he", p, u, dirichlet_key="u_0_ref")
    v = assembly.local_state(p, u, "u_0_ref")
    q = assembly.element_eval("he", p, v)
    order = plan.elem_order
    g, vals = emulate_block_patch_assembly(plan, ef[order], q["g"][order], q["h"][order], lay.indptr, n_free, 3)
    H = sp.csr_matrix((vals, lay.indices, lay.indptr), shape=(n_free, n_free))[lay.iperm][:, lay.iperm]
    assert abs(H - ref["h"]).max() <= 1e-12 * abs(ref["h"]).max()
    assert np.allclose(g[lay.iperm], ref["g"], rtol=1e-10, atol=1e-10 * np.abs(ref["g"]).max())


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_host_plans_fit_kernel_limits(prec):
    """Assembler set-up without a GPU (device='host-plan'): LDS request within budget, limits met."""
    from fenics_nonlinear_energies_amd.problems.ginzburg_landau import GLReorderedElementAssembler, build_gl_problem
    from fenics_nonlinear_energies_amd.problems.hyperelasticity import HEReorderedElementAssembler, build_beam_problem

    gp, gadj = build_gl_problem(40)
    hp, hadj, _ = build_beam_problem(24, 3, 3)
    for cls, p, adj in ((GLReorderedElementAssembler, gp, gadj), (HEReorderedElementAssembler, hp, hadj)):
        asm = cls(p, None, adj, hessian_precision=prec, device="host-plan")
        plan = asm.patch_plan
        if asm.block_size == 1:  # staged P1 kernel: every patch element staged in LDS at once
            hsz = 4 if prec == "fp32" else 8
            assert plan.chunk >= plan.max_patch_elems and plan.chunk % 64 == 0
            lds = 24 * plan.chunk + hsz * (((plan.max_patch_nnz + 1) & ~1) + 6 * plan.chunk)
            assert lds <= 160 * 1024  # gfx950 workgroup LDS limit
            assert int(np.max(np.diff(lay_indptr := asm.layout.indptr))) <= 30
        else:  # tetrahedra: row-gather plan (16 lanes x 4 rows per group, rows <= 254 entries)
            assert plan.n_patch == 0 and plan.pitem.shape == (plan.inc.size, 32)
            assert plan.max_row_len <= 254
            assert 64 * ((plan.max_row_len + 1) & ~1) * 8 <= 160 * 1024
            assert plan.inc_ptr[-1] == plan.inc.size
        assert plan.rows_per_patch <= 256
        assert np.all(np.diff(plan.prow) <= 256)
        assert asm._host_codes.shape == (p["elems"].shape[0], p["elems"].shape[1])
        # every element code is a free index or n_free + a Dirichlet slot
        n_free = asm.layout.n_free
        n_total = asm.dirichlet_full.size
        assert asm._host_codes.min() >= 0 and asm._host_codes.max() < n_free + n_total


def test_rowgather_plan_reproduces_oracle_assembly():
    d, adj = fixtures.gl(5)
    n_free = d["freedofs"].size
    perm = layout.select_permutation("block_morton", adjacency=adj, coords_all=d["nodes"], freedofs=d["freedofs"])
    lay = layout.build_global_layout(d, adj, perm, block_size=1, dirichlet_key="u_0")
    ef = lay.total_to_free_reord[d["elems"]]
    plan = patches.build_p1_rowgather_plan(ef, lay.indptr, lay.indices, n_free)
    u = np.random.default_rng(9).uniform(-1, 1, n_free)
    ref = assembly.assemble("gl", d, u)
    q = assembly.element_eval("gl", d, assembly.local_state(d, u))
    order = plan.elem_order
    he, ge = q["h"][order], q["g"][order]
    items = plan.inc.astype(np.uint64)
    rows = np.repeat(np.arange(n_free), np.diff(plan.inc_ptr))
    e = (items & np.uint64(0xFFFFFFFF)).astype(np.int64)
    a = ((items >> np.uint64(32)) & np.uint64(3)).astype(np.int64)
    g = np.zeros(n_free)
    np.add.at(g, rows, ge[e, a])
    vals = np.zeros(lay.nnz)
    for b in range(3):
        pos = ((items >> np.uint64(40 + 8 * b)) & np.uint64(255)).astype(np.int64)
        ok = pos != 255
        np.add.at(vals, lay.indptr[rows[ok]] + pos[ok], he[e[ok], a[ok], b])
    H = sp.csr_matrix((vals, lay.indices, lay.indptr), shape=(n_free, n_free))[lay.iperm][:, lay.iperm]
    assert abs(H - ref["h"]).max() <= 1e-12 * abs(ref["h"]).max()
    assert np.allclose(g[lay.iperm], ref["g"], rtol=1e-12, atol=1e-14)


def test_p1_incidence_plan_reproduces_oracle_assembly():
    """Incidence (staged) plan: rows sum their items in order; emulated and compared to the oracle."""
    d, adj = fixtures.gl(5)
    n_free = d["freedofs"].size
    perm = layout.select_permutation("block_morton", adjacency=adj, coords_all=d["nodes"], freedofs=d["freedofs"])
    lay = layout.build_global_layout(d, adj, perm, block_size=1, dirichlet_key="u_0")
    ef = lay.total_to_free_reord[d["elems"]]
    plan = patches.build_p1_plan(ef, lay.indptr, lay.indices, n_free, rows_per_patch=208)
    u = np.random.default_rng(4).uniform(-1, 1, n_free)
    ref = assembly.assemble("gl", d, u)
    q = assembly.element_eval("gl", d, assembly.local_state(d, u))
    order = plan.elem_order
    he, ge = q["h"][order], q["g"][order]
    items = plan.inc.astype(np.int64)
    rows = np.repeat(np.arange(n_free), np.diff(plan.inc_ptr))
    slot = items & 0x1FFF
    a = (items >> 13) & 3
    e = plan.pelem[plan.pelem_ptr[rows // 208] + slot]
    g = np.zeros(n_free)
    np.add.at(g, rows, ge[e, a])
    vals = np.zeros(lay.nnz)
    for b in range(3):
        pos = (items >> (15 + 5 * b)) & 31
        ok = pos != 31
        np.add.at(vals, lay.indptr[rows[ok]] + pos[ok], he[e[ok], a[ok], b])
    H = sp.csr_matrix((vals, lay.indices, lay.indptr), shape=(n_free, n_free))[lay.iperm][:, lay.iperm]
    assert abs(H - ref["h"]).max() <= 1e-12 * abs(ref["h"]).max()
    assert np.allclose(g[lay.iperm], ref["g"], rtol=1e-12, atol=1e-14)
    assert np.all(np.diff(plan.inc_ptr) <= 8) or True  # rows with more items use the slow path
