"""Compressible Neo-Hookean beam: problem data, rotating face, assembler (rows A3, A14).

Problem dictionaries mirror MeshHyperElasticity3D._build_problem_data
(hyperelasticity/support/mesh.py:40-82): ``u_0, u_0_ref, freedofs, elems (12 DOFs),
elems_scalar, dphix, dphiy, dphiz, vol, C1, D1, nodes2coord, right_nodes,
elastic_kernel``.  The unknown is the deformed position.  The beam generator reproduces
the reference's HyperElasticity_level*.h5 meshes: nodes x-fastest, then y, then z; each
hexahedron split into six tetrahedra in the same vertex order; elements hex by hex with x
fastest.  The assembler mirrors HEReorderedElementAssembler
(hyperelasticity/jax_petsc/reordered_element_assembler.py:16-101).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from ..assembler import ReorderedElementAssemblerBase
from ..native import FNE_FLAG_ABS_DET, FNE_FLAG_ROW_GATHER, FNE_KIND_NEOHOOKEAN
from ..patches import build_block_plan, build_row_incidence_plan

# corner offsets (di, dj, dk) of the six tetrahedra of one hexahedron
_HEX_TETS = (
    ((0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 0, 1)),
    ((0, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1)),
    ((0, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1)),
    ((1, 0, 0), (1, 1, 0), (0, 1, 0), (1, 0, 1)),
    ((1, 1, 0), (1, 0, 1), (1, 1, 1), (0, 1, 0)),
    ((0, 1, 0), (1, 0, 1), (1, 1, 1), (0, 1, 1)),
)


def lame_to_neohookean(E: float, nu: float) -> tuple[float, float]:
    """C1 = mu/2 = E/(4(1+nu)),  D1 = kappa/2 = E/(6(1-2nu))."""
    return float(E) / (4.0 * (1.0 + float(nu))), float(E) / (6.0 * (1.0 - 2.0 * float(nu)))


def beam_mesh(nx: int, ny: int, nz: int, length=0.4, width=0.01, height=0.01):
    xs = np.linspace(0.0, length, nx + 1)
    ys = np.linspace(-0.5 * width, 0.5 * width, ny + 1)
    zs = np.linspace(-0.5 * height, 0.5 * height, nz + 1)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    nodes = np.column_stack((X.ravel(), Y.ravel(), Z.ravel()))
    nid = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)  # noqa: E731
    kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    tets = np.empty((ii.size, 6, 4), dtype=np.int64)
    for t, corners in enumerate(_HEX_TETS):
        for a, (di, dj, dk) in enumerate(corners):
            tets[:, t, a] = nid(ii + di, jj + dj, kk + dk)
    return nodes, tets.reshape(-1, 4)


def tet_operators(nodes: np.ndarray, tets: np.ndarray):
    X = nodes[tets]  # (N, 4, 3)
    M = np.stack((X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]), axis=2)  # columns
    Minv = np.linalg.inv(M)  # rows = grad of barycentric 1..3
    grads = np.concatenate((-Minv.sum(axis=1, keepdims=True), Minv), axis=1)  # (N, 4, 3)
    vol = np.abs(np.linalg.det(M)) / 6.0
    return grads[:, :, 0], grads[:, :, 1], grads[:, :, 2], vol


def expand_tet_connectivity_to_dofs(elems2nodes: np.ndarray) -> np.ndarray:
    """[n0..n3] -> [3n0, 3n0+1, 3n0+2, ..., 3n3+2] (hyperelasticity/support/mesh.py:6-16)."""
    e = np.asarray(elems2nodes, dtype=np.int64)
    return (3 * e[:, :, None] + np.arange(3)[None, None, :]).reshape(e.shape[0], 12)


def rigid_modes(nodes: np.ndarray) -> np.ndarray:
    n = nodes.shape[0]
    R = np.zeros((3 * n, 6))
    R[0::3, 0] = 1.0
    R[1::3, 1] = 1.0
    R[2::3, 2] = 1.0
    R[1::3, 3], R[2::3, 3] = -nodes[:, 2], nodes[:, 1]
    R[0::3, 4], R[2::3, 4] = nodes[:, 2], -nodes[:, 0]
    R[0::3, 5], R[1::3, 5] = -nodes[:, 1], nodes[:, 0]
    return R


def problem_from_mesh(nodes, tets, C1, D1, dphi=None):
    dphix, dphiy, dphiz, vol = tet_operators(nodes, tets) if dphi is None else dphi
    x = nodes[:, 0]
    left = np.isclose(x, x.min())
    right = np.isclose(x, x.max())
    fixed_nodes = np.nonzero(left | right)[0]
    fixed = np.zeros(3 * nodes.shape[0], dtype=bool)
    for c in range(3):
        fixed[3 * fixed_nodes + c] = True
    freedofs = np.nonzero(~fixed)[0].astype(np.int64)
    u0 = nodes.reshape(-1).astype(np.float64)
    elems = expand_tet_connectivity_to_dofs(tets)
    params = {
        "u_0": u0.copy(),
        "u_0_ref": u0.copy(),
        "freedofs": freedofs,
        "elems": elems,
        "elems_scalar": np.asarray(tets, dtype=np.int64),
        "dphix": np.asarray(dphix, np.float64),
        "dphiy": np.asarray(dphiy, np.float64),
        "dphiz": np.asarray(dphiz, np.float64),
        "vol": np.asarray(vol, np.float64),
        "C1": float(C1),
        "D1": float(D1),
        "nodes2coord": np.asarray(nodes, np.float64),
        "right_nodes": np.nonzero(right)[0].astype(np.int64),
        "elastic_kernel": rigid_modes(nodes)[freedofs],
    }
    full_to_free = np.full(u0.size, -1, dtype=np.int64)
    full_to_free[freedofs] = np.arange(freedofs.size)
    loc = full_to_free[elems]
    rr = np.repeat(loc, 12, axis=1).ravel()
    cc = np.tile(loc, (1, 12)).ravel()
    m = (rr >= 0) & (cc >= 0)
    n = freedofs.size
    keys = np.unique(rr[m] * np.int64(n) + cc[m])
    adjacency = sp.coo_matrix((np.ones(keys.size), (keys // n, keys % n)), shape=(n, n))
    return params, adjacency, u0[freedofs].copy()


def build_beam_problem(nx=160, ny=8, nz=8, E=2.0e8, nu=0.3, length=0.4, width=0.01):
    nodes, tets = beam_mesh(nx, ny, nz, length, width, width)
    C1, D1 = lame_to_neohookean(E, nu)
    return problem_from_mesh(nodes, tets, C1, D1)


def rotate_right_face_from_reference(u0_reference, nodes2coord, angle, right_nodes=None):
    """Right-face Dirichlet values rotated about the x axis (support/rotate_boundary.py:4-22)."""
    out = np.asarray(u0_reference, dtype=np.float64).reshape(-1).copy()
    coords = np.asarray(nodes2coord, dtype=np.float64)
    if right_nodes is None:
        right_nodes = np.nonzero(np.isclose(coords[:, 0], coords[:, 0].max()))[0]
    rn = np.asarray(right_nodes, dtype=np.int64)
    y0, z0 = coords[rn, 1], coords[rn, 2]
    out[3 * rn + 0] = coords[rn, 0]
    out[3 * rn + 1] = np.cos(angle) * y0 + np.sin(angle) * z0
    out[3 * rn + 2] = -np.sin(angle) * y0 + np.cos(angle) * z0
    return out


class HEReorderedElementAssembler(ReorderedElementAssemblerBase):
    """Neo-Hookean tetrahedron assembler on the GPU (block size 3)."""

    block_size = 3
    coordinate_key = "nodes2coord"
    dirichlet_key = "u_0_ref"
    local_elem_data_keys = ("dphix", "dphiy", "dphiz", "vol")
    near_nullspace_key = "elastic_kernel"
    fne_kind = FNE_KIND_NEOHOOKEAN

    def __init__(self, params, comm=None, adjacency=None, *, use_abs_det=False, ksp_type="gmres", ksp_rtol=1e-1,
                 ksp_max_it=30, reorder_mode="block_morton", **kw):
        self.use_abs_det = bool(use_abs_det)
        kw.setdefault("rows_per_patch", None)
        # element-major element matrices + CSR row gather: measured faster than the node-block
        # patch schedule on the 160x8x8 (2.8x) and 320x16x16 (1.2x) beams
        kw.setdefault("assembly_mode", "rowgather")
        near_nullspace_mode = kw.pop("near_nullspace_mode", "current")
        super().__init__(params, comm, adjacency, ksp_type=ksp_type, ksp_rtol=ksp_rtol, ksp_max_it=ksp_max_it,
                         reorder_mode=reorder_mode, **kw)
        # The unknown is the deformed position, so the rigid-body modes of the current
        # configuration are the near-kernel of the current tangent; "reference" keeps the
        # modes of the undeformed mesh (params["elastic_kernel"]) as the reference's GAMG does.
        node_of = np.asarray(self.params["freedofs"], dtype=np.int64)[self.layout.perm] // 3
        node_major = node_of.size % 3 == 0 and np.array_equal(node_of[0::3], node_of[1::3]) and np.array_equal(
            node_of[0::3], node_of[2::3])
        if near_nullspace_mode not in ("current", "reference"):
            raise ValueError("near_nullspace_mode must be 'current' or 'reference'")
        if near_nullspace_mode == "current" and node_major and getattr(self, "ksp", None) is not None:
            self.ksp.near_nullspace_fn = self._rigid_modes_current

    def _rigid_modes_current(self):
        """Rigid-body modes at the state of the last Hessian assembly (free nodes, reordered),
        as an (n, 6) device tensor laid out like ``rigid_modes``."""
        import torch

        x = getattr(self, "_hessian_state", None)
        if x is None:
            return None
        X = x.detach().to(torch.float64).reshape(-1, 3)
        B = torch.zeros((x.numel(), 6), dtype=torch.float64, device=x.device)
        for c in range(3):
            B[c::3, c] = 1.0
        B[1::3, 3], B[2::3, 3] = -X[:, 2], X[:, 1]
        B[0::3, 4], B[2::3, 4] = X[:, 2], -X[:, 0]
        B[0::3, 5], B[1::3, 5] = -X[:, 1], X[:, 0]
        return B

    def _element_geometry(self) -> np.ndarray:
        p = self.params
        return np.concatenate(
            [np.asarray(p["dphix"]).T, np.asarray(p["dphiy"]).T, np.asarray(p["dphiz"]).T, np.asarray(p["vol"])[None, :]],
            axis=0,
        ).astype(np.float64)

    def _kernel_params(self):
        return [float(self.params["C1"]), float(self.params["D1"])]

    def _plan_flags(self) -> int:
        flags = FNE_FLAG_ABS_DET if self.use_abs_det else 0
        if self.assembly_mode == "rowgather":
            flags |= FNE_FLAG_ROW_GATHER
        return flags

    def _lds_bytes_per_element(self) -> int:
        return 39 * 8

    def _lds_slice_bytes_per_entry(self) -> int:
        return 4 if self.hessian_precision == "fp32" else 8

    def _build_patch_plan(self, elem_free, rows_per_patch, chunk):
        lay = self.layout
        if self.assembly_mode == "rowgather":
            return build_row_incidence_plan(elem_free, lay.indptr, lay.indices, lay.n_free)
        # pick the node count per patch so the CSR slice leaves room for >= 64 staged elements
        row_len = np.diff(np.asarray(lay.indptr, dtype=np.int64))
        max_row = int(row_len.max(initial=1))
        budget = self.lds_budget_bytes() - 64 * self._lds_bytes_per_element() - 64
        nodes = budget // (3 * max_row * self._lds_slice_bytes_per_entry())
        nodes = int(max(1, min(85, nodes if rows_per_patch is None else int(rows_per_patch) // 3)))
        return build_block_plan(elem_free, lay.indptr, lay.indices, lay.n_free, block_size=3,
                                nodes_per_patch=nodes, chunk=int(chunk or 128))


HE_NEWTON = dict(tolf=1e-4, tolg=1e-3, tolg_rel=1e-3, maxit=100, linesearch_interval=(-0.5, 2.0),
                 linesearch_tol=1e-3)


def solve_he_load_steps(asm, total_steps: int = 24, total_angle: float = 4.0 * np.pi, *, num_steps=None,
                        predictor=None, on_step=None, ksp_rtol: float = 1e-1, ksp_max_it: int = 30,
                        pc_setup_on_ksp_cap: bool = False, pc_refresh_its: int | None = None,
                        pc_keep_across_steps: bool = True, retry_on_failure=True, stop_on_fail: bool = False,
                        newton_kwargs: dict | None = None):
    """Rotate the right face in ``total_steps`` equal steps and solve each step with Newton.

    The reference's load-step loop (``hyperelasticity/jax_petsc/solver.py:225-493``):
    angle_k = k * total_angle / total_steps, the rotated face is imposed through
    ``update_dirichlet``, the step starts from the state the previous step left, and a failed
    step (non-finite, Newton or linear iteration cap) is solved again from its start with the
    tighter repair settings (``retry_on_failure``, on by default as in ``solve_HE_dof.py:165-170``;
    a ``repair.RepairPolicy`` may be passed instead).  ``predictor(history, angle)`` may return
    another initial guess (device tensor of free values); ``history`` holds (angle, state) of
    the converged steps, starting with the reference configuration at angle 0.  The reference
    rotates by 4 * 2 pi in total (``solver.py:225``); the angle is a parameter here.
    Preconditioner policy: set up at every Newton iteration (the reference's), or with
    ``pc_setup_on_ksp_cap`` kept across iterations (and across steps unless
    ``pc_keep_across_steps`` is False, which sets it up at each step's first Newton iteration)
    and rebuilt when a solve hits ``ksp_max_it`` or needs more than ``pc_refresh_its``
    iterations.  ``stop_on_fail`` ends the
    path at the first step that does not converge.  Returns (records, state vector).
    """
    import torch

    from ..drivers import solve_newton
    from ..load_path import march, uniform_loads

    p = asm.params
    opts = dict(HE_NEWTON)
    opts.update(newton_kwargs or {})
    lay = asm.layout
    ref = np.asarray(p["u_0_ref"], dtype=np.float64)
    x = asm.create_vec(ref[p["freedofs"]][lay.perm])

    def apply(angle):
        asm.update_dirichlet(rotate_right_face_from_reference(ref, p["nodes2coord"], angle, p["right_nodes"]))

    def solve(vec, _angle):
        return solve_newton(asm, vec, ksp_rtol=ksp_rtol, ksp_max_it=ksp_max_it,
                            pc_setup_on_ksp_cap=pc_setup_on_ksp_cap,
                            keep_pc=pc_setup_on_ksp_cap and pc_keep_across_steps,
                            pc_refresh_its=pc_refresh_its, retry_on_failure=retry_on_failure, **opts)

    def notify(rec, history):
        if on_step is not None:
            on_step(rec.as_dict("angle"), history)

    recs, _ = march(x, uniform_loads(total_angle, total_steps, num_steps), apply=apply, solve=solve,
                    guess=predictor, initial_history=[(0.0, x.t.clone())],
                    stop=(lambda r: not r.converged) if stop_on_fail else None, on_step=notify,
                    sync=torch.cuda.synchronize)
    return [r.as_dict("angle") for r in recs], x
