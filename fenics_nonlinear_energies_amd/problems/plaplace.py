"""p-Laplace problem data + assembler (row A1).

Problem data mirror the reference's ``pLaplace_level*.h5`` dictionaries
(plaplace/support/mesh.py:10-41): ``nodes, elems, u_0, freedofs, dvx, dvy, vol, f, p``.
The assembler mirrors ``PLaplaceReorderedElementAssembler``
(plaplace/jax_petsc/reordered_element_assembler.py:22-72).
"""

from __future__ import annotations

import numpy as np

from ..assembler import ReorderedElementAssemblerBase
from ..native import FNE_KIND_PLAPLACE
from . import mesh2d


def build_plaplace_problem(n_cells: int = 128, p: float = 3.0, f: float = -10.0, bbox=(0.0, 1.0, 0.0, 1.0)):
    """Structured p-Laplace problem on a rectangle, u = 0 on the boundary.

    Returns ``(params, adjacency, u_init)`` with a zero initial guess on the free DOFs.
    """
    x0, x1, y0, y1 = map(float, bbox)
    nodes, elems = mesh2d.structured_triangle_mesh(n_cells, n_cells, x0, x1, y0, y1)
    dvx, dvy, vol = mesh2d.triangle_operators(nodes, elems)
    bnd = mesh2d.rectangle_boundary_mask(nodes, x0, x1, y0, y1)
    freedofs = np.nonzero(~bnd)[0].astype(np.int64)
    n_nodes = nodes.shape[0]
    params = {
        "nodes": nodes,
        "elems": elems,
        "u_0": np.zeros(n_nodes, dtype=np.float64),
        "freedofs": freedofs,
        "dvx": dvx,
        "dvy": dvy,
        "vol": vol,
        "f": mesh2d.lumped_nodal_load(elems, vol, f, n_nodes),
        "p": float(p),
    }
    adjacency = mesh2d.free_dof_adjacency(elems, freedofs, n_nodes)
    return params, adjacency, np.zeros(freedofs.size, dtype=np.float64)


def reference_random_init(n_free: int) -> np.ndarray:
    """The reference's initial guess for p-Laplace meshes (plaplace/support/mesh.py:19-21)."""
    rng = np.random.RandomState(0)
    return rng.rand(n_free)


class PLaplaceReorderedElementAssembler(ReorderedElementAssemblerBase):
    """p-Laplace element assembler on the GPU (energy, gradient, CSR Hessian)."""

    block_size = 1
    coordinate_key = "nodes"
    dirichlet_key = "u_0"
    local_elem_data_keys = ("dvx", "dvy", "vol")
    fne_kind = FNE_KIND_PLAPLACE

    def _element_geometry(self) -> np.ndarray:
        p = self.params
        return np.concatenate(
            [np.asarray(p["dvx"], np.float64).T, np.asarray(p["dvy"], np.float64).T, np.asarray(p["vol"], np.float64)[None, :]],
            axis=0,
        )

    def _geometry_from_nodes(self) -> np.ndarray | None:
        return mesh2d.triangle_geometry_rows(self.params["nodes"], self.params["elems"])

    def _kernel_params(self) -> list[float]:
        return [float(self.params["p"])]

    def assemble_first_step_hessian(self, u_owned):
        """Assemble the p = 2 (Laplacian) Hessian into ``A``: first-step policy for u = 0 starts."""
        p = float(self.params["p"])
        self.set_kernel_params([2.0])
        try:
            return self.assemble_hessian(u_owned)
        finally:
            self.set_kernel_params([p])

    def _build_rhs_owned(self) -> np.ndarray:
        rhs = np.asarray(self.params["f"], dtype=np.float64)
        freedofs = np.asarray(self.params["freedofs"], dtype=np.int64)
        return np.asarray(rhs[freedofs][self.layout.perm], dtype=np.float64)
