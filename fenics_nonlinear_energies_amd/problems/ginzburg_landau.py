"""Ginzburg-Landau problem data + assembler (row A2).

Data keys mirror ``GL_level*.h5`` (ginzburg_landau/jax/mesh.py:10-43):
``nodes, elems, u_0, freedofs, dvx, dvy, vol, ip, w, eps``.  The 3-point rule
``ip[a][q] = phi_a(xi_q)`` at (1/6,1/6), (2/3,1/6), (1/6,2/3) with weights 1/3 is the one
stored in the reference's GL files.  The assembler mirrors ``GLReorderedElementAssembler``
(ginzburg_landau/jax_petsc/reordered_element_assembler.py:27-97): on one device every
element, including those with only Dirichlet nodes, is in the energy set, so the
reference's separately added Dirichlet-only energy is part of the kernel sum.
"""

from __future__ import annotations

import numpy as np

from ..assembler import ReorderedElementAssemblerBase
from .. import native
from ..native import FNE_KIND_GINZBURG_LANDAU
from . import mesh2d

GL_IP = np.array(
    [
        [1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0],
        [2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0],
        [1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0],
    ],
    dtype=np.float64,
)
GL_W = np.array([1.0 / 3.0, 1.0 / 3.0, 1.0 / 3.0], dtype=np.float64)


def build_gl_problem(n_cells: int = 1024, eps: float = 0.01, bbox=(-1.0, 1.0, -1.0, 1.0)):
    """Structured GL problem on ``bbox``, u = 0 on the boundary.  Returns (params, adjacency)."""
    x0, x1, y0, y1 = map(float, bbox)
    nodes, elems = mesh2d.structured_triangle_mesh(n_cells, n_cells, x0, x1, y0, y1)
    dvx, dvy, vol = mesh2d.triangle_operators(nodes, elems)
    bnd = mesh2d.rectangle_boundary_mask(nodes, x0, x1, y0, y1)
    freedofs = np.nonzero(~bnd)[0].astype(np.int64)
    params = {
        "nodes": nodes,
        "elems": elems,
        "u_0": np.zeros(nodes.shape[0], dtype=np.float64),
        "freedofs": freedofs,
        "dvx": dvx,
        "dvy": dvy,
        "vol": vol,
        "ip": GL_IP.copy(),
        "w": GL_W.copy(),
        "eps": float(eps),
    }
    return params, mesh2d.free_dof_adjacency(elems, freedofs, nodes.shape[0])


def sine_noise_init(params: dict, seed: int = 0, amplitude: float = 0.5, noise: float = 1e-3) -> np.ndarray:
    """u0 = amplitude*sin(pi x)*sin(pi y) + U[-noise, noise] on the free nodes (seeded)."""
    nodes = np.asarray(params["nodes"], dtype=np.float64)
    free = np.asarray(params["freedofs"], dtype=np.int64)
    u = amplitude * np.sin(np.pi * nodes[free, 0]) * np.sin(np.pi * nodes[free, 1])
    rng = np.random.default_rng(seed)
    return u + rng.uniform(-noise, noise, size=free.size)


def reference_sine_init(params: dict) -> np.ndarray:
    """The reference's GL initial guess (ginzburg_landau/jax/mesh.py:20-23)."""
    nodes = np.asarray(params["nodes"], dtype=np.float64)
    u = np.sin(np.pi * (nodes[:, 0] - 1) / 2) * np.sin(np.pi * (nodes[:, 1] - 1) / 2)
    return u[np.asarray(params["freedofs"], dtype=np.int64)]


class GLReorderedElementAssembler(ReorderedElementAssemblerBase):
    """Ginzburg-Landau element assembler on the GPU."""

    block_size = 1
    coordinate_key = "nodes"
    dirichlet_key = "u_0"
    local_elem_data_keys = ("dvx", "dvy", "vol")
    fne_kind = FNE_KIND_GINZBURG_LANDAU

    def _element_geometry(self) -> np.ndarray:
        p = self.params
        return np.concatenate(
            [np.asarray(p["dvx"], np.float64).T, np.asarray(p["dvy"], np.float64).T, np.asarray(p["vol"], np.float64)[None, :]],
            axis=0,
        )

    def _geometry_from_nodes(self) -> np.ndarray | None:
        return mesh2d.triangle_geometry_rows(self.params["nodes"], self.params["elems"])

    def _kernel_params(self) -> list[float]:
        return [float(self.params["eps"])]

    def _quadrature_table(self) -> np.ndarray | None:
        """ip[3][nq], w[nq], then the products w_q ip[a][q] ip[b][q] for ab = 00 11 22 01 02 12
        (FNE_FLAG_GL_PRODUCTS): the Hessian's quadrature weights, one per entry and point."""
        ip = np.asarray(self.params["ip"], dtype=np.float64)
        w = np.asarray(self.params["w"], dtype=np.float64)
        pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
        prod = np.stack([w * ip[a] * ip[b] for a, b in pairs])
        return np.concatenate([ip.reshape(-1), w, prod.reshape(-1)])

    def _plan_flags(self) -> int:
        return super()._plan_flags() | native.FNE_FLAG_GL_PRODUCTS

    def _build_rhs_owned(self) -> np.ndarray:
        return np.zeros(self.layout.n_free, dtype=np.float64)
