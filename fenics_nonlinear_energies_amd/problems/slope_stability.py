"""Plane-strain Mohr-Coulomb slope problem (rows A4/A5): mesh, P_k data, assembler.

Problem data follow the reference's ``SlopeStabilityCaseData``
(``slope_stability/support/mesh.py:42-84``) and its same-mesh Lagrange builder
(``mesh.py:995-1091``):

* the homogeneous slope geometry is parametrised by (h, x1, x2, x3, y1, y2): a lower
  rectangle [0, x1+x2+x3] x [0, y1] and an upper band whose right edge runs from
  (x1+x2, y1) to (x1, y1+y2) (``mesh.py:334-521``); each grid cell is split into the
  triangles (i,j),(i+1,j),(i,j+1) and (i+1,j+1),(i,j+1),(i+1,j) (``mesh.py:382-444``);
* P_k nodes are the images of the equispaced reference lattice, merged by coordinates
  rounded to 12 decimals (``mesh.py:670-736``);
* x-displacement is fixed on x = 0, x = x_max and y = 0, y-displacement on y = 0
  (``mesh.py:674-684``); DOFs are interleaved (2n, 2n+1) (``mesh.py:524-527``);
* gravity load f = int phi_a (0, -gamma) (``mesh.py:609-628``).

Element geometry stored for the kernels: the inverse Jacobian K (4 values, the matrix
``inv`` of ``mesh.py:310-313``) and |det J|; the strain-displacement operator B of
``mesh.py:314-323`` is rebuilt on the device from K and the reference gradient table.
"""

from __future__ import annotations

import math

import numpy as np
from scipy import sparse

from ..assembler import ReorderedElementAssemblerBase
from ..native import FNE_KIND_MOHR_COULOMB
from ..patches import build_row_incidence_plan
from . import lagrange

COORD_DECIMALS = 12
LEVEL_TO_H = {1: 4.0, 2: 2.0, 3: 1.0, 4: 0.5, 5: 0.25, 6: 0.125, 7: 0.0625}
REFERENCE_GEOMETRY = dict(x1=15.0, x2=10.0, x3=15.0, y1=10.0, y2=10.0)
REFERENCE_MATERIAL = dict(E=40000.0, nu=0.30, c0=6.0, phi_deg=45.0, psi_deg=0.0, gamma=20.0)


def davis_b_reduction(c0: float, phi_deg: float, psi_deg: float, lam: float) -> tuple[float, float]:
    """Davis-B strength reduction -> (reduced cohesion c_lambda, reduced friction angle in degrees).

    Same definition as ``slope_stability/support/reduction.py:8-29``: c01 = c0/lam,
    phi1 = atan(tan(phi)/lam), psi1 = atan(tan(psi)/lam),
    beta = cos(phi1) cos(psi1) / (1 - sin(phi1) sin(psi1)), c_lambda = beta c01,
    phi_lambda = atan(beta tan(phi1)).
    """
    if lam <= 0.0:
        raise ValueError("lambda must be positive")
    phi = math.radians(float(phi_deg))
    psi = math.radians(float(psi_deg))
    c01 = float(c0) / float(lam)
    phi1 = math.atan(math.tan(phi) / float(lam))
    psi1 = math.atan(math.tan(psi) / float(lam))
    beta = math.cos(phi1) * math.cos(psi1) / (1.0 - math.sin(phi1) * math.sin(psi1))
    return beta * c01, math.degrees(math.atan(beta * math.tan(phi1)))


def slope_macro_mesh(h: float, x1: float, x2: float, x3: float, y1: float, y2: float):
    """Vertex coordinates (n, 2) and CCW triangles (m, 3) of the homogeneous slope."""
    nx12 = int(round((x1 + x2) / h))
    nx3 = int(round(x3 / h))
    nx = nx12 + nx3
    ny1 = int(round(y1 / h))
    ny2 = int(round(y2 / h))
    xs = np.concatenate((np.linspace(0.0, x1 + x2, nx12 + 1), np.linspace(x1 + x2, x1 + x2 + x3, nx3 + 1)[1:]))
    ys = np.concatenate((np.linspace(0.0, y1, ny1 + 1), np.linspace(y1, y1 + y2, ny2 + 1)[1:]))
    # lower block: (nx+1) x (ny1+1) vertices, row-major in j
    X, Y = np.meshgrid(xs, ys[: ny1 + 1], indexing="xy")
    lower = np.stack((X.ravel(), Y.ravel()), axis=1)
    upper = []
    for j in range(ny1 + 1, ny1 + ny2 + 1):
        x_max = x1 + x2 * (y1 + y2 - ys[j]) / y2
        upper.append(np.stack((np.linspace(0.0, x_max, nx12 + 1), np.full(nx12 + 1, ys[j])), axis=1))
    nodes = np.concatenate([lower] + upper) if upper else lower

    def vid(i, j):
        i = np.asarray(i)
        j = np.asarray(j)
        low = j * (nx + 1) + i
        up = (ny1 + 1) * (nx + 1) + (j - ny1 - 1) * (nx12 + 1) + i
        return np.where(j <= ny1, low, up)

    tris = []
    for j in range(ny1 + ny2):
        ncol = nx if j < ny1 else nx12
        i = np.arange(ncol)
        t1 = np.stack((vid(i, j), vid(i + 1, j), vid(i, j + 1)), axis=1)
        t2 = np.stack((vid(i + 1, j + 1), vid(i, j + 1), vid(i + 1, j)), axis=1)
        tris.append(np.stack((t1, t2), axis=1).reshape(-1, 3))
    return nodes.astype(np.float64), np.concatenate(tris).astype(np.int64)


def lagrange_mesh(macro_nodes: np.ndarray, macro_tris: np.ndarray, degree: int):
    """P_k nodes (n, 2) and scalar connectivity (m, n_p) on the same triangles."""
    ref = lagrange.lattice_points(degree)
    v = macro_nodes[macro_tris]  # (m, 3, 2)
    lam0 = 1.0 - ref[:, 0] - ref[:, 1]
    pts = (lam0[None, :, None] * v[:, None, 0, :] + ref[None, :, 0, None] * v[:, None, 1, :]
           + ref[None, :, 1, None] * v[:, None, 2, :])  # (m, n_p, 2)
    flat = pts.reshape(-1, 2)
    keys = np.round(flat, COORD_DECIMALS)
    # first-appearance numbering, like the reference's insertion-ordered lookup
    uniq, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    nodes = flat[first[order]]
    return nodes.astype(np.float64), rank[inv.ravel()].reshape(pts.shape[0], pts.shape[1]).astype(np.int64)


def q_mask_from_nodes(nodes: np.ndarray) -> np.ndarray:
    x_max = float(np.max(nodes[:, 0])) - 1.0e-9
    q = np.zeros((nodes.shape[0], 2), dtype=bool)
    q[:, 0] = (nodes[:, 0] > 0.0) & (nodes[:, 1] > 0.0) & (nodes[:, 0] < x_max)
    q[:, 1] = nodes[:, 1] > 0.0
    return q


def element_geometry(nodes: np.ndarray, elems_scalar: np.ndarray):
    """Inverse Jacobian K (m, 2, 2) and det J (m,) of the affine map (vertices first)."""
    v = nodes[elems_scalar[:, :3]]
    j11 = v[:, 1, 0] - v[:, 0, 0]
    j12 = v[:, 1, 1] - v[:, 0, 1]
    j21 = v[:, 2, 0] - v[:, 0, 0]
    j22 = v[:, 2, 1] - v[:, 0, 1]
    det = j11 * j22 - j12 * j21
    K = np.stack((np.stack((j22, -j12), 1), np.stack((-j21, j11), 1)), 1) / det[:, None, None]
    return K, det


def reference_tables(degree: int):
    """(quadrature points, weights, phi (n_p, n_q), dphi (n_q, n_p, 2)) for degree k."""
    xq, wq = lagrange.quadrature(lagrange.quadrature_degree_for_lagrange(degree))
    phi, d1, d2 = lagrange.tabulate(degree, xq)
    dhat = np.stack((d1.T, d2.T), axis=2)  # (n_q, n_p, 2)
    return xq, wq, phi, dhat


def build_elem_B(params: dict) -> np.ndarray:
    """Strain-displacement operators (m, n_q, 3, 2 n_p) as the reference stores them (``mesh.py:314-323``)."""
    K, _ = params["K"], None
    dhat = params["dhat"]  # (n_q, n_p, 2)
    gx = np.einsum("qa,m->mqa", dhat[:, :, 0], K[:, 0, 0]) + np.einsum("qa,m->mqa", dhat[:, :, 1], K[:, 0, 1])
    gy = np.einsum("qa,m->mqa", dhat[:, :, 0], K[:, 1, 0]) + np.einsum("qa,m->mqa", dhat[:, :, 1], K[:, 1, 1])
    m, nq, npn = gx.shape
    B = np.zeros((m, nq, 3, 2 * npn))
    B[:, :, 0, 0::2] = gx
    B[:, :, 1, 1::2] = gy
    B[:, :, 2, 0::2] = gy
    B[:, :, 2, 1::2] = gx
    return B


def dof_adjacency(elems: np.ndarray, freedofs: np.ndarray, n_dofs: int) -> sparse.coo_matrix:
    full_to_free = np.full(n_dofs, -1, dtype=np.int64)
    full_to_free[freedofs] = np.arange(freedofs.size)
    ef = full_to_free[elems]
    nl = ef.shape[1]
    r = np.repeat(ef, nl, axis=1).ravel()
    c = np.tile(ef, (1, nl)).ravel()
    ok = (r >= 0) & (c >= 0)
    keys = np.unique(r[ok] * np.int64(freedofs.size) + c[ok])
    rr, cc = np.divmod(keys, np.int64(freedofs.size))
    return sparse.coo_matrix((np.ones(keys.size), (rr, cc)), shape=(freedofs.size, freedofs.size))


def build_slope_problem(h: float = 0.5, degree: int = 4, *, geometry: dict | None = None,
                        material: dict | None = None, with_adjacency: bool = True):
    """Homogeneous slope problem on P_k triangles.

    Returns ``(params, adjacency, u_init)``; ``u_init`` is zero (the reference solves for an
    elastic initial guess separately, ``mesh.py:785-799``).  ``params`` holds the
    reference's keys (nodes, elems_scalar, elems, freedofs, q_mask, force, u_0, E, nu, c0,
    phi_deg, psi_deg, gamma) plus the kernel data K, detJ, wq, dhat, phi.
    """
    geo = dict(REFERENCE_GEOMETRY if geometry is None else geometry)
    mat = dict(REFERENCE_MATERIAL if material is None else material)
    vnodes, vtris = slope_macro_mesh(h, **geo)
    nodes, elems_scalar = lagrange_mesh(vnodes, vtris, degree) if degree > 1 else (vnodes, vtris)
    q_mask = q_mask_from_nodes(nodes)
    n_nodes = nodes.shape[0]
    elems = (2 * elems_scalar[:, :, None] + np.arange(2)[None, None, :]).reshape(elems_scalar.shape[0], -1)
    freedofs = np.arange(2 * n_nodes, dtype=np.int64).reshape(n_nodes, 2)[q_mask].astype(np.int64)
    K, det = element_geometry(nodes, elems_scalar)
    xq, wq, phi, dhat = reference_tables(degree)
    # gravity: f_a = sum_q |det J| w_q phi_a(q) (0, -gamma)
    fa = np.abs(det)[:, None] * (phi @ wq)[None, :]  # (m, n_p)
    force = np.zeros(2 * n_nodes)
    np.add.at(force, 2 * elems_scalar.ravel() + 1, -float(mat["gamma"]) * fa.ravel())
    params = {
        "nodes": nodes,
        "elems_scalar": elems_scalar,
        "elems": elems,
        "freedofs": freedofs,
        "q_mask": q_mask,
        "force": force,
        "u_0": np.zeros(2 * n_nodes),
        "K": K,
        "detJ": det,
        "xq": xq,
        "wq": wq,
        "phi": phi,
        "dhat": dhat,
        "degree": int(degree),
        "h": float(h),
        **geo,
        **mat,
        "cohesion": float(mat["c0"]),
        "reg": 1.0e-12,
    }
    adjacency = dof_adjacency(elems, freedofs, 2 * n_nodes) if with_adjacency else None
    return params, adjacency, np.zeros(freedofs.size)


def set_strength_reduction(params: dict, lam: float) -> None:
    """Apply Davis-B reduction at factor ``lam`` (``slope_stability/jax_petsc/solver.py:600-603``)."""
    c, phi = davis_b_reduction(params["c0"], params["phi_deg_raw"] if "phi_deg_raw" in params else params["phi_deg"],
                               params["psi_deg"], lam)
    params.setdefault("phi_deg_raw", params["phi_deg"])
    params["cohesion"] = float(c)
    params["phi_deg"] = float(phi)
    params["lambda"] = float(lam)


class SlopeStabilityReorderedElementAssembler(ReorderedElementAssemblerBase):
    """Mohr-Coulomb P_k assembler (energy, gradient, CSR Hessian) on the GPU.

    Mirrors ``SlopeStabilityReorderedElementAssembler``
    (``slope_stability/jax_petsc/reordered_element_assembler.py:152-311``): kernel
    parameters (E, nu, phi_deg, cohesion, reg) and the constitutive mode switch of
    ``assemble_hessian_with_mode`` (:305-311).
    """

    block_size = 2
    coordinate_key = "nodes"
    dirichlet_key = "u_0"
    local_elem_data_keys = ("K", "detJ")
    near_nullspace_key = None
    fne_kind = FNE_KIND_MOHR_COULOMB

    def __init__(self, params, comm=None, adjacency=None, *, ksp_type="cg", ksp_rtol=1e-3, ksp_max_it=10000,
                 reorder_mode="block_morton", **kw):
        kw.setdefault("hessian_precision", "fp32")
        super().__init__(params, comm, adjacency, ksp_type=ksp_type, ksp_rtol=ksp_rtol, ksp_max_it=ksp_max_it,
                         reorder_mode=reorder_mode, **kw)
        self._constitutive_mode = 0

    def _near_nullspace(self) -> np.ndarray:
        """Plane rigid-body modes on the free DOFs (``build_near_nullspace_modes``, ``mesh.py:802-813``)."""
        X = np.asarray(self.params["nodes"], dtype=np.float64)
        n = X.shape[0]
        R = np.zeros((2 * n, 3))
        R[0::2, 0] = 1.0
        R[1::2, 1] = 1.0
        c = X.mean(axis=0)
        R[0::2, 2] = -(X[:, 1] - c[1])
        R[1::2, 2] = X[:, 0] - c[0]
        return R[np.asarray(self.params["freedofs"], dtype=np.int64)]

    def _configure_ksp(self, ksp) -> None:
        """pc_type "mg": the same-mesh p-multigrid levels P_k -> ... -> P1 (halving the degree,
        4 -> 2 -> 1; ``slope_stability/jax_petsc/multigrid.py:1291-1835`` with its
        ``same_mesh_p4_p2_p1`` hierarchy), smoothed aggregation below P1.  Three smoothing
        steps per level and side unless ``mg_levels_ksp_max_it`` says otherwise: the step count
        of the reference's p-multigrid smoothers (``multigrid.py:1410``); on the 1.57M-DOF slope
        continuation 3 Chebyshev steps took 17.0 s against 18.9 s with 2 and 18.7 s with 4."""
        if ksp.pc_type != "mg":
            return
        if not getattr(self, "_pc_sweeps_given", False):
            ksp.mg_sweeps = 3
        ksp.mg_prolongators, ksp.mg_coarse_near_nullspace = self.pmg_transfers()

    def pmg_transfers(self):
        """Host CSR transfers between the free DOFs of the same-mesh levels (finest first, each
        coarse level in its own block Morton order) and the plane rigid-body modes on the
        coarsest of them."""
        from .. import pmg
        from ..layout import perm_block_morton

        p = self.params
        degree = int(p.get("degree", 4))
        degrees = [degree]
        while degrees[-1] > 1:
            degrees.append(max(1, degrees[-1] // 2))
        geo = {k: float(p[k]) for k in ("x1", "x2", "x3", "y1", "y2")}
        vnodes, vtris = slope_macro_mesh(float(p["h"]), **geo)
        meshes = {degree: (np.asarray(p["nodes"]), np.asarray(p["elems_scalar"]))}
        for k in degrees[1:]:
            meshes[k] = lagrange_mesh(vnodes, vtris, k) if k > 1 else (vnodes, vtris)
        order_f = pmg.free_order(p["freedofs"], self.layout.perm)
        transfers = []
        nodes_c = None
        for kf, kc in zip(degrees[:-1], degrees[1:]):
            (nf, ef), (nc, ec) = meshes[kf], meshes[kc]
            Ps = pmg.scalar_transfer(ef, kf, ec, kc, nf.shape[0], nc.shape[0])
            free_c = np.flatnonzero(q_mask_from_nodes(nc).ravel())
            order_c = pmg.free_order(free_c, perm_block_morton(nc, free_c, 2))
            transfers.append(pmg.vector_free_transfer(Ps, 2, order_f, order_c))
            order_f, nodes_c = order_c, nc
        if nodes_c is None:
            return [], None
        R = np.zeros((2 * nodes_c.shape[0], 3))
        R[0::2, 0] = 1.0
        R[1::2, 1] = 1.0
        c = np.asarray(p["nodes"], dtype=np.float64).mean(axis=0)
        R[0::2, 2] = -(nodes_c[:, 1] - c[1])
        R[1::2, 2] = nodes_c[:, 0] - c[0]
        return transfers, R[order_f]

    def _element_geometry(self) -> np.ndarray:
        K = np.asarray(self.params["K"], dtype=np.float64)
        det = np.asarray(self.params["detJ"], dtype=np.float64)
        return np.stack((K[:, 0, 0], K[:, 0, 1], K[:, 1, 0], K[:, 1, 1], np.abs(det)), axis=0)

    def _quadrature_table(self) -> np.ndarray:
        wq = np.asarray(self.params["wq"], dtype=np.float64)
        dhat = np.asarray(self.params["dhat"], dtype=np.float64)
        return np.concatenate((wq, dhat.ravel()))

    def _quadrature_points(self) -> int:
        return int(np.asarray(self.params["wq"]).size)

    def _kernel_params(self):
        p = self.params
        return [float(p["E"]), float(p["nu"]), float(p["phi_deg"]), float(p["cohesion"]), float(p.get("reg", 1e-12))]

    def _plan_flags(self) -> int:
        return 0

    def _build_patch_plan(self, elem_free, rows_per_patch, chunk):
        lay = self.layout
        return build_row_incidence_plan(elem_free, lay.indptr, lay.indices, lay.n_free)

    def _build_rhs_owned(self) -> np.ndarray:
        f = np.asarray(self.params["force"], dtype=np.float64)
        return np.asarray(f[np.asarray(self.params["freedofs"], dtype=np.int64)][self.layout.perm], dtype=np.float64)

    def set_strength(self, cohesion: float, phi_deg: float) -> None:
        """Update the reduced strength parameters between continuation steps."""
        self.params["cohesion"] = float(cohesion)
        self.params["phi_deg"] = float(phi_deg)
        self.set_kernel_params(self._kernel_params())

    def set_constitutive_mode(self, mode: int) -> None:
        from .. import native

        native.check(native.lib().fne_plan_set_mode(self._plan, int(mode)), "fne_plan_set_mode")
        self._constitutive_mode = int(mode)

    def assemble_hessian_with_mode(self, u_owned, *, constitutive_mode: str = "plastic"):
        """``constitutive_mode`` "plastic" (full MC) or "elastic" (elastic tangent)."""
        mode = {"plastic": 0, "elastic": 1}[str(constitutive_mode)]
        prev = self._constitutive_mode
        self.set_constitutive_mode(mode)
        try:
            return self.assemble_hessian(u_owned)
        finally:
            self.set_constitutive_mode(prev)

    # matrix-free Hessian operator (slope_stability/jax_petsc/reordered_element_assembler.py:313-562)
    MATRIX_FREE_MODES = ("matfree_element", "matfree_overlap")

    def get_matrix_free_operator(self, mode: str):
        """The shell operator of ``mode`` (one per mode, created on first use; ``:329-333``)."""
        from ..linalg import MatrixFreeOperator

        mode = str(mode)
        ops = self.__dict__.setdefault("_matrix_free_ops", {})
        if mode not in ops:
            op = MatrixFreeOperator(
                self.layout.n_free,
                lambda x, y, m=mode: self._matrix_free_apply(m, x, y),
                lambda d, m=mode: self._matrix_free_diag_into(m, d),
                block_size=self.block_size,
                device=self.device,
                name=mode,
            )
            op.setNearNullSpace(self._near_nullspace_reordered())
            ops[mode] = op
        return ops[mode]

    def prepare_matrix_free_operator(self, u_owned, *, mode: str):
        """Fix the linearisation point of the matrix-free Hessian and return its operator
        (``:335-388``).  Both of the reference's modes ("matfree_element": element tangents
        applied element by element; "matfree_overlap": the linearised gradient of the local
        subdomain) are the same operator on one device, H(u) v summed element by element in
        ``fne_hvp``; the mode name is kept so that operators, caches and statistics stay
        per mode as in the reference."""
        import time

        import torch

        mode = str(mode)
        if mode not in self.MATRIX_FREE_MODES:
            raise ValueError(f"Unsupported matrix-free operator mode {mode!r}")
        t0 = time.perf_counter()
        x = self._as_device(u_owned)
        self._mf_x = torch.empty(self.layout.n_free, dtype=torch.float64, device=self.device)
        self._mf_x.copy_(x)
        self._mf_mode = mode
        self._mf_state_id = getattr(self, "_mf_state_id", 0) + 1
        self._mf_diag = None
        self._mf_pc_cache = {}
        self._mf_stats = {
            "mode": mode,
            "prepare_total": time.perf_counter() - t0,
            "diagonal_source": "not_requested",
            "diagonal_prepare_total": 0.0,
            "python_pc_variant": "none",
            "python_pc_prepare_total": 0.0,
            "python_pc_apply_calls": 0,
            "python_pc_apply_total": 0.0,
            "mult_calls": 0,
            "mult_apply": 0.0,
            "mult_total": 0.0,
        }
        op = self.get_matrix_free_operator(mode)
        op.stateIncrease()
        return op

    def matrix_free_summary(self) -> dict:
        return dict(getattr(self, "_mf_stats", {}))

    def _require_matrix_free_state(self, mode: str | None = None) -> None:
        if getattr(self, "_mf_x", None) is None:
            raise RuntimeError("Matrix-free state has not been prepared")
        if mode is not None and self._mf_mode != str(mode):
            raise RuntimeError(f"Matrix-free state {self._mf_mode!r} does not match requested mode {mode!r}")

    def _matrix_free_apply(self, mode: str, x, y) -> None:
        import time

        t0 = time.perf_counter()
        if getattr(self, "_mf_x", None) is None or self._mf_mode != str(mode):
            raise RuntimeError(f"Matrix-free operator {mode!r} is missing a matching linearization point")
        self.hvp_device(self._mf_x, x, y)
        st = self._mf_stats
        st["mult_calls"] += 1
        dt = time.perf_counter() - t0
        st["mult_apply"] += dt
        st["mult_total"] += dt

    def matrix_free_mult(self, mode: str, x, y) -> None:
        """y = H(u) x at the prepared point (``:494-562``); x, y: assembler vectors."""
        self._matrix_free_apply(mode, x.t, y.t)

    def _matrix_free_diag_into(self, mode: str, d) -> None:
        import time

        import torch

        self._require_matrix_free_state(mode)
        if self._mf_diag is None or self._mf_diag[0] != self._mf_state_id:
            t0 = time.perf_counter()
            diag = torch.empty(self.layout.n_free, dtype=torch.float64, device=self.device)
            self.hessian_diagonal_device(self._mf_x, diag)
            self._mf_diag = (self._mf_state_id, diag)
            self._mf_stats["diagonal_source"] = "element_hessian_diagonal"
            self._mf_stats["diagonal_prepare_total"] = time.perf_counter() - t0
        d.copy_(self._mf_diag[1])

    def matrix_free_get_diagonal(self, mode: str, diag_vec) -> None:
        """The diagonal of H(u) at the prepared point, cached per state (``:418-427``)."""
        self._matrix_free_diag_into(mode, diag_vec.t)

    def make_matrix_free_python_pc_context(self, variant: str) -> "MatrixFreePythonPC":
        return MatrixFreePythonPC(self, variant)

    def prepare_matrix_free_python_pc(self, variant: str) -> None:
        """Factor the free-DOF Hessian at the prepared point (variant "overlap_lu", ``:451-469``).

        On one device the overlap-local operator is the whole free-DOF operator.  It is
        assembled in FP64 by ``fne_hessian`` and factored on the host by SuperLU, as the
        reference factors its local operator with scipy's ``splu``."""
        import time

        import torch
        from scipy.sparse.linalg import splu

        from .. import native

        variant = str(variant)
        self._require_matrix_free_state()
        cache = self._mf_pc_cache.get(variant)
        if cache is not None and cache["state_id"] == self._mf_state_id:
            return
        if variant != "overlap_lu":
            raise ValueError(f"Unsupported matrix-free python PC variant {variant!r}")
        t0 = time.perf_counter()
        vals = torch.empty(self.A.nnz, dtype=torch.float64, device=self.device)
        native.check(native.lib().fne_hessian(self._plan, native.ptr(self._mf_x), native.ptr(self.dvals), None,
                                              None, native.ptr(vals), native.FNE_FP64, native.stream_ptr()),
                     "fne_hessian")
        H = sparse.csr_matrix((vals.cpu().numpy(), self.A.indices.cpu().numpy(), self.A.indptr.cpu().numpy()),
                              shape=(self.layout.n_free, self.layout.n_free))
        self._mf_pc_cache[variant] = {"state_id": self._mf_state_id, "factor": splu(H.tocsc())}
        self._mf_stats["python_pc_variant"] = variant
        self._mf_stats["python_pc_prepare_total"] = time.perf_counter() - t0

    def apply_matrix_free_python_pc(self, variant: str, b, x) -> None:
        """x = LU solve of b with the prepared factor (``:471-492``)."""
        import time

        import torch

        variant = str(variant)
        self.prepare_matrix_free_python_pc(variant)
        t0 = time.perf_counter()
        from ..vec import DeviceVec

        rhs = b.t if isinstance(b, DeviceVec) else b
        sol = self._mf_pc_cache[variant]["factor"].solve(rhs.double().cpu().numpy())
        (x.t if isinstance(x, DeviceVec) else x).copy_(torch.as_tensor(sol))
        self._mf_stats["python_pc_apply_calls"] += 1
        self._mf_stats["python_pc_apply_total"] += time.perf_counter() - t0

    def cleanup(self) -> None:
        for op in getattr(self, "_matrix_free_ops", {}).values():
            op.destroy()
        self.__dict__.pop("_matrix_free_ops", None)
        self._mf_x = None
        self._mf_pc_cache = {}
        self._mf_stats = {}
        super().cleanup()

    def hessian_diagonal_device(self, x, diag) -> None:
        """diag = diagonal of H(x) without assembling H (``fne_hessian_diag``)."""
        from .. import native

        native.check(native.lib().fne_hessian_diag(self._plan, native.ptr(x), native.ptr(self.dvals),
                                                   native.ptr(diag), native.stream_ptr()), "fne_hessian_diag")


class MatrixFreePythonPC:
    """Preconditioner context over the assembler's factored operator (the reference's
    ``_MatrixFreePythonPCContext``, ``slope_stability/jax_petsc/reordered_element_assembler.py:53-72``):
    ``setUp`` factors, ``apply`` solves, ``view`` names the variant."""

    def __init__(self, assembler: SlopeStabilityReorderedElementAssembler, variant: str):
        self.assembler = assembler
        self.variant = str(variant)

    def setUp(self, pc=None) -> None:
        del pc
        self.assembler.prepare_matrix_free_python_pc(self.variant)

    def apply(self, pc, b, x) -> None:
        del pc
        self.assembler.apply_matrix_free_python_pc(self.variant, b, x)

    def view(self, pc=None, viewer=None) -> str:
        del pc
        text = f"Matrix-free Python PC variant: {self.variant}\n"
        if viewer is not None:
            viewer.printfASCII(text)
        return text
