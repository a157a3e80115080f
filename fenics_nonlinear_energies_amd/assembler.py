"""GPU element assembler with the reference's ReorderedElementAssemblerBase surface.

Mirrors core/petsc/reordered_element_base.py:643-1783 (single device: lo = 0, hi = n_free,
no MPI exchange): the constructor builds the permutation, the layout and the CSR pattern,
then -- instead of JAX closures and a numpy scatter plan -- uploads the element data and
an owner-computes patch schedule to the GPU and creates a ``libfne_hip`` plan.

Callbacks keep their reference signatures:

* ``energy_fn(vec) -> float``                 (:1579-1611)
* ``gradient_fn(vec, g) -> None``             (:1613-1633)
* ``assemble_hessian(u_owned) -> timings``    (:1635-1641, :1733-1775; writes ``self.A``)
* ``update_dirichlet(u_0_new)``, ``create_vec``, ``cleanup``, ``setup_summary``,
  ``memory_summary``, ``callback_summary``.

Vectors are :class:`DeviceVec` (FP64 on the GPU); ``assemble_hessian`` also accepts a host
array.  There is no CPU path: a missing or failing HIP library raises.
"""

from __future__ import annotations

import time
from types import SimpleNamespace

import numpy as np
import torch

from . import native
from .layout import build_global_layout, select_permutation, validate_permutation
from .linalg import KSP, DeviceCSR
from .patches import PatchPlan, build_p1_colored_plan, build_p1_plan, build_p1_rowgather_plan
from .vec import DeviceVec


class _SelfComm:
    """Stand-in for a one-rank communicator (the single-GPU case of MPI.COMM_WORLD)."""

    rank = 0
    size = 1

    def Get_rank(self):
        return 0

    def Get_size(self):
        return 1

    def allreduce(self, v, op=None):
        return v

    def allgather(self, v):
        return [v]


class ReorderedElementAssemblerBase:
    """Generic GPU assembler; subclasses provide the energy family and element data."""

    distribution_strategy = "single_device"
    block_size = 1
    coordinate_key = "nodes"
    dirichlet_key = "u_0"
    local_elem_data_keys: tuple[str, ...] = ()
    near_nullspace_key: str | None = None
    fne_kind: int = 0

    def __init__(
        self,
        params,
        comm=None,
        adjacency=None,
        *,
        ksp_rtol=1e-3,
        ksp_type="cg",
        pc_type="jacobi",
        ksp_max_it=10000,
        pc_options=None,
        reorder_mode="block_morton",
        local_hessian_mode="element",
        use_near_nullspace=False,
        perm_override=None,
        distribution_strategy=None,
        reuse_hessian_value_buffers=True,
        assembly_backend="coo",
        petsc_log_events=False,
        memory_guard_total_gib=None,
        hessian_precision="fp64",
        pc_precision=None,
        assembly_mode="staged",
        rows_per_patch=None,
        chunk=None,
        lds_budget=None,
        element_geometry="arrays",
        device="cuda",
    ):
        if comm is not None and int(comm.Get_size()) != 1:
            raise ValueError("the GPU assembler runs one rank per device; pass a one-rank communicator")
        if local_hessian_mode not in ("element", "sfd_local", "sfd_local_vmap"):
            raise ValueError(f"Unsupported local_hessian_mode={local_hessian_mode!r} "
                             "(element | sfd_local | sfd_local_vmap)")
        if assembly_backend not in {"coo", "coo_local", "blocked_local"}:
            raise ValueError(f"Unsupported assembly backend {assembly_backend!r}")
        host_only = str(device) == "host-plan"
        if not host_only:
            if not torch.cuda.is_available():
                raise RuntimeError("ReorderedElementAssemblerBase needs a GPU (torch.cuda is unavailable)")
            native.lib()  # fail loudly before any set-up work
        self.comm = comm if comm is not None else _SelfComm()
        self.rank = 0
        self.size = 1
        self.params = params
        self.device = None if host_only else torch.device(device)
        self.reorder_mode = str(reorder_mode)
        self.local_hessian_mode = str(local_hessian_mode)
        self.use_near_nullspace = bool(use_near_nullspace)
        self.assembly_backend = str(assembly_backend)
        self.hessian_precision = str(hessian_precision)
        self.assembly_mode = str(assembly_mode)
        if element_geometry not in ("auto", "arrays", "nodes"):
            raise ValueError(f"element_geometry must be 'auto', 'arrays' or 'nodes', not {element_geometry!r}")
        self.element_geometry = str(element_geometry)
        self.node_geometry = False
        self._lds_budget = None if lds_budget is None else int(lds_budget)
        self.memory_guard_total_gib = memory_guard_total_gib
        self.iter_timings: list[dict] = []
        self._hvp_eval_mode = "element_patch_lds"
        self._setup_timings: dict[str, float] = {}
        self._callback_stats = {
            "energy": {"calls": 0, "kernel": 0.0, "total": 0.0},
            "gradient": {"calls": 0, "kernel": 0.0, "total": 0.0},
            "hessian": {"calls": 0, "hvp_compute": 0.0, "total": 0.0},
        }
        t_total = time.perf_counter()

        freedofs = np.asarray(params["freedofs"], dtype=np.int64).ravel()
        t0 = time.perf_counter()
        if perm_override is None:
            perm = select_permutation(
                self.reorder_mode,
                adjacency=adjacency,
                coords_all=np.asarray(params[self.coordinate_key], dtype=np.float64),
                freedofs=freedofs,
                n_parts=1,
                block_size=int(self.block_size),
            )
        else:
            perm = validate_permutation(perm_override, freedofs.size)
        self._setup_timings["permutation"] = time.perf_counter() - t0

        t0 = time.perf_counter()
        self.layout = build_global_layout(
            params, adjacency, perm, block_size=int(self.block_size), dirichlet_key=self.dirichlet_key
        )
        self._setup_timings["global_layout"] = time.perf_counter() - t0
        self.part = SimpleNamespace(
            perm=self.layout.perm,
            iperm=self.layout.iperm,
            lo=0,
            hi=self.layout.n_free,
            n_free=self.layout.n_free,
            n_owned=self.layout.n_free,
        )
        self.dirichlet_full = np.asarray(params[self.dirichlet_key], dtype=np.float64).ravel().copy()

        t0 = time.perf_counter()
        self._build_host_plan(rows_per_patch=rows_per_patch, chunk=chunk)
        self._setup_timings["scatter_build"] = time.perf_counter() - t0
        if host_only:
            # CPU-side planning only (tests, plan inspection): no device state is created
            return
        t0 = time.perf_counter()
        self._upload_plan()
        self._setup_timings["device_upload"] = time.perf_counter() - t0

        t0 = time.perf_counter()
        self._f_owned = np.asarray(self._build_rhs_owned(), dtype=np.float64)
        self._f_dev = (
            torch.as_tensor(self._f_owned).to(self.device) if np.any(self._f_owned != 0.0) else None
        )
        self._setup_timings["rhs_build"] = time.perf_counter() - t0

        t0 = time.perf_counter()
        self.A = DeviceCSR(self.layout.indptr, self.layout.indices, self.layout.n_free, self.hessian_precision, self.device)
        self.ksp = KSP(ksp_type=ksp_type, pc_type=pc_type, rtol=ksp_rtol, max_it=ksp_max_it)
        self.ksp.near_nullspace = self._near_nullspace_reordered()
        # precision of the multigrid hierarchy (None: the Hessian's); "fp32" with an FP64
        # Hessian keeps the Krylov operator exact and stores only the preconditioner in FP32
        self.ksp.pc_precision = pc_precision
        self._apply_pc_options(pc_options)
        self._configure_ksp(self.ksp)
        self._setup_timings["matrix_create"] = time.perf_counter() - t0
        # multigrid symbolic phase (patterns of the Galerkin products): mesh preprocessing,
        # done once here like the layout and the assembly plans
        self._setup_timings["pc_symbolic"] = self.ksp.prepare_pc(self.A)
        if self.local_hessian_mode != "element":
            t0 = time.perf_counter()
            self._setup_sfd()
            self._setup_timings["sfd_coloring"] = time.perf_counter() - t0
        self._energy_out = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._setup_timings["total"] = time.perf_counter() - t_total

    # ------------------------------------------------------------------------------------
    # problem hooks
    # ------------------------------------------------------------------------------------
    def _apply_pc_options(self, pc_options) -> None:
        """The multigrid level-smoother options among ``pc_options`` (PETSc option names, as the
        reference passes them): ``mg_levels_ksp_type`` ("chebyshev", the default as in PCGAMG,
        or "richardson" = damped Jacobi sweeps) and ``mg_levels_ksp_max_it`` (steps per level and
        side, 2 by default); and this build's ``mg_coarse_factor`` ("host": LAPACK, the default,
        or "device": the GPU's dense solvers) for the coarsest level's inverse.  Other options
        have no counterpart on this solver stack and are ignored."""
        opts = {str(k).lstrip("-"): v for k, v in dict(pc_options or {}).items()}
        kind = str(opts.get("mg_levels_ksp_type", "chebyshev")).lower()
        if kind not in ("richardson", "jacobi", "chebyshev"):
            raise ValueError(f"unsupported mg_levels_ksp_type {kind!r} (richardson | chebyshev)")
        self.ksp.mg_smoother = "chebyshev" if kind == "chebyshev" else "jacobi"
        coarse = str(opts.get("mg_coarse_factor", "host")).lower()
        if coarse not in ("host", "device"):
            raise ValueError(f"unsupported mg_coarse_factor {coarse!r} (host | device)")
        self.ksp.mg_coarse_factor = coarse
        self._pc_sweeps_given = "mg_levels_ksp_max_it" in opts
        if self._pc_sweeps_given:
            self.ksp.mg_sweeps = int(opts["mg_levels_ksp_max_it"])

    def _configure_ksp(self, ksp) -> None:
        """Problem-specific KSP set-up before the multigrid's symbolic phase (e.g. the level
        transfers of pc_type "mg"); nothing by default."""

    def _near_nullspace(self) -> np.ndarray | None:
        """Near-nullspace vectors on the free DOFs (original order) for the AMG preconditioner:
        ``params[near_nullspace_key]`` when present (``reordered_element_base.py:1062-1090``),
        else constants for scalar problems."""
        key = getattr(self, "near_nullspace_key", None)
        if key and key in self.params:
            return np.asarray(self.params[key], dtype=np.float64)
        if self.block_size == 1:
            return np.ones((self.layout.n_free, 1))
        return None

    def _near_nullspace_reordered(self) -> np.ndarray | None:
        B = self._near_nullspace()
        return None if B is None else np.ascontiguousarray(B[self.layout.perm])

    def _element_geometry(self) -> np.ndarray:
        """(n_geo, n_elem) FP64 element data in the kernel's order."""
        raise NotImplementedError

    def _geometry_from_nodes(self) -> np.ndarray | None:
        """The element geometry as the kernels evaluate it from the vertex coordinates
        (FNE_FLAG_NODE_GEO), same shape as ``_element_geometry``; None where the family has no
        such form."""
        return None

    def _choose_node_geometry(self) -> bool:
        """Whether the kernels evaluate the element geometry from the vertex coordinates:
        "auto" when the coordinates reproduce the given geometry arrays bit for bit (then
        both forms give identical results and the node form stores 16 bytes per vertex instead
        of n_geo doubles per element and patch element), "nodes" requires it, "arrays" (the
        default) never.  On the 1024^2 Ginzburg-Landau grid both forms assemble in the same
        time (the staged kernel is bound by latency, not bytes), so the node form is a memory
        saving: 255 MB there."""
        if self.element_geometry == "arrays":
            return False
        geo_nodes = self._geometry_from_nodes()
        same = geo_nodes is not None and np.array_equal(geo_nodes, self._element_geometry())
        if self.element_geometry == "nodes" and not same:
            raise ValueError("element_geometry='nodes': the vertex coordinates do not reproduce the element "
                             "geometry arrays bit for bit (or the family has no node form)")
        return bool(same)

    def _dof_coordinates(self) -> np.ndarray:
        """(n_free + n_total, 2) vertex coordinates per DOF code: free DOFs in the reordered
        numbering, then every node (the Dirichlet codes n_free + node)."""
        lay = self.layout
        nodes = np.asarray(self.params[self.coordinate_key], dtype=np.float64)
        t2f = lay.total_to_free_reord
        node_of_free = np.empty(lay.n_free, dtype=np.int64)
        free_nodes = np.flatnonzero(t2f >= 0)
        node_of_free[t2f[free_nodes]] = free_nodes
        return np.ascontiguousarray(np.concatenate((nodes[node_of_free], nodes), axis=0)[:, :2])

    def _kernel_params(self) -> list[float]:
        return []

    def _quadrature_table(self) -> np.ndarray | None:
        return None

    def _plan_flags(self) -> int:
        if self.block_size != 1:
            return 0
        if self.assembly_mode == "rowgather":
            return native.FNE_FLAG_ROW_GATHER
        return native.FNE_FLAG_STAGED if self.assembly_mode == "staged" else native.FNE_FLAG_COLORED

    def _build_rhs_owned(self) -> np.ndarray:
        return np.zeros(self.layout.n_free, dtype=np.float64)

    def _lds_bytes_per_element(self) -> int:
        """LDS staged per element in the patch kernel (P1: 6 Hessian + 3 gradient FP64)."""
        return 72

    def _single_pass_staging(self) -> bool:
        """True when the P1 kernel holds every patch element at once (colored: registers, staged: LDS)."""
        return self.assembly_mode in ("patch", "staged") and self.block_size == 1

    def _lds_slice_bytes_per_entry(self) -> int:
        """LDS bytes per CSR entry of the patch slice (P1 accumulates in FP64)."""
        return 8

    def lds_budget_bytes(self) -> int:
        """Dynamic LDS one patch workgroup may use (default 64 KiB: two or more workgroups per CU)."""
        return self._lds_budget if getattr(self, "_lds_budget", None) else 64 * 1024

    def _build_patch_plan(self, elem_free, rows_per_patch, chunk) -> PatchPlan:
        if self.assembly_mode == "rowgather":
            return build_p1_rowgather_plan(elem_free, self.layout.indptr, self.layout.indices, self.layout.n_free)
        if self.assembly_mode == "staged":
            plan = build_p1_plan(
                elem_free, self.layout.indptr, self.layout.indices, self.layout.n_free,
                rows_per_patch=int(rows_per_patch or 208), chunk=64,
            )
            plan.chunk = int(-(-max(plan.max_patch_elems, 1) // 64) * 64)  # stage every patch element
            return plan
        if self.assembly_mode != "patch":
            raise ValueError(f"unknown assembly_mode {self.assembly_mode!r} (patch | staged | rowgather)")
        # colored patch plan: largest row patch whose elements fit elems_per_thread * 256
        ept = int(chunk or 2)
        rpp = int(rows_per_patch or 256)
        while True:
            try:
                return build_p1_colored_plan(
                    elem_free, self.layout.indptr, self.layout.indices, self.layout.n_free,
                    rows_per_patch=rpp, elems_per_thread=ept,
                )
            except ValueError as exc:
                if "more elements" not in str(exc) or rpp <= 16:
                    raise
                rpp -= 16

    # ------------------------------------------------------------------------------------
    # device plan
    # ------------------------------------------------------------------------------------
    def _t(self, a, dtype) -> torch.Tensor:
        return torch.as_tensor(np.ascontiguousarray(a, dtype=dtype)).to(self.device)

    def _build_host_plan(self, rows_per_patch, chunk) -> None:
        lay = self.layout
        elems = np.asarray(self.params["elems"], dtype=np.int64)
        elem_free = lay.total_to_free_reord[elems]
        plan = self._build_patch_plan(elem_free, rows_per_patch, chunk)
        # LDS budget of one workgroup: CSR slice + staged element data per chunk
        budget = self.lds_budget_bytes()
        if plan.n_patch == 0 or (self._single_pass_staging() and self.assembly_mode == "patch"):
            budget = 1 << 30  # no LDS element staging
        per_elem = self._lds_bytes_per_element()
        p1_staged = self.assembly_mode == "staged" and self.block_size == 1
        slice_bytes = 0 if p1_staged else self._lds_slice_bytes_per_entry() * int(plan.max_patch_nnz)
        avail = budget - slice_bytes - 64
        if plan.chunk * per_elem > avail:
            if self._single_pass_staging():
                raise ValueError("patch elements do not fit the LDS budget; reduce rows_per_patch")
            if avail < 64 * per_elem:
                raise ValueError("patch CSR slice too large for LDS; reduce rows_per_patch")
            plan.chunk = int(max(64, (avail // per_elem) // 64 * 64))
        self.patch_plan = plan
        order = plan.elem_order
        n_free = lay.n_free
        self._host_codes = np.where(elem_free >= 0, elem_free, n_free + elems)[order]
        self.node_geometry = self._choose_node_geometry()
        self._host_geo = self._element_geometry()[:, order]
        self.n_elem = int(elems.shape[0])
        self.n_loc = int(elems.shape[1])

    def _upload_plan(self) -> None:
        lay = self.layout
        plan = self.patch_plan
        n_free = lay.n_free
        codes, geo = self._host_codes, self._host_geo
        node_geo = bool(self.node_geometry)
        self._dev = {
            "edof": self._t(codes.T, np.int32),
            "indptr": self._t(lay.indptr, np.int32),
            "indices": self._t(lay.indices, np.int32),
            "prow": self._t(plan.prow, np.int32),
            "pelem_ptr": self._t(plan.pelem_ptr, np.int32),
            "pelem": self._t(plan.pelem, np.int32),
            "inc_ptr": self._t(plan.inc_ptr, np.int32),
            "inc": self._t(plan.inc.view(np.int32), np.int32),
        }
        if node_geo:
            self._dev["dof_xy"] = self._t(self._dof_coordinates(), np.float64)
        else:
            self._dev["egeo"] = self._t(geo, np.float64)
        if plan.inc_aux is not None:
            self._dev["inc_aux"] = self._t(plan.inc_aux, np.int32)
        if plan.n_patch > 0:
            pe = plan.pelem.astype(np.int64)
            self._dev["pedof"] = self._t(codes[pe].T, np.int32)
            if not node_geo:
                self._dev["pgeo"] = self._t(geo[:, pe], np.float64)
            self._dev["pown"] = self._t(plan.pown, np.int32)
        if getattr(plan, "pitem", None) is not None:
            self._dev["pitem"] = self._t(np.ascontiguousarray(plan.pitem).view(np.int64), np.int64)
        if getattr(plan, "pcolors", None) is not None:
            self._dev["pcolors"] = self._t(plan.pcolors, np.int32)
        q = self._quadrature_table()
        if q is not None:
            self._dev["qdata"] = self._t(q, np.float64)
        self.dvals = self._t(self.dirichlet_full, np.float64)

        d = native.FnePlanDesc()
        d.kind = int(self.fne_kind)
        d.n_elem = self.n_elem
        d.n_loc = self.n_loc
        d.n_geo = int(geo.shape[0])
        d.n_free = int(n_free)
        d.n_total = int(self.dirichlet_full.size)
        d.nnz = int(lay.nnz)
        d.edof = native.ptr(self._dev["edof"])
        d.egeo = native.ptr(self._dev.get("egeo"))
        d.dof_xy = native.ptr(self._dev.get("dof_xy"))
        d.eweight = None
        d.indptr = native.ptr(self._dev["indptr"])
        d.indices = native.ptr(self._dev["indices"])
        d.n_patch = int(plan.n_patch)
        d.chunk = int(plan.chunk)
        d.max_patch_nnz = int(plan.max_patch_nnz)
        d.max_row_len = int(plan.max_row_len)
        d.prow = native.ptr(self._dev["prow"])
        d.pelem_ptr = native.ptr(self._dev["pelem_ptr"])
        d.pelem = native.ptr(self._dev["pelem"])
        d.inc_ptr = native.ptr(self._dev["inc_ptr"])
        d.inc = native.ptr(self._dev["inc"])
        d.inc_aux = native.ptr(self._dev.get("inc_aux"))
        d.n_pe = int(plan.pelem.size) if plan.n_patch > 0 else 0
        d.pedof = native.ptr(self._dev.get("pedof"))
        d.pgeo = native.ptr(self._dev.get("pgeo"))
        d.pown = native.ptr(self._dev.get("pown"))
        d.pitem = native.ptr(self._dev.get("pitem"))
        d.pcolors = native.ptr(self._dev.get("pcolors"))
        prm = list(self._kernel_params()) + [0.0] * 8
        for i in range(8):
            d.param[i] = float(prm[i])
        d.qdata = native.ptr(self._dev.get("qdata"))
        d.nq = int(0 if q is None else self._quadrature_points())
        d.flags = int(self._plan_flags()) | (native.FNE_FLAG_NODE_GEO if node_geo else 0)
        self._desc = d
        import ctypes

        handle = ctypes.c_void_p()
        native.check(native.lib().fne_plan_create(ctypes.byref(d), ctypes.byref(handle)), "fne_plan_create")
        self._plan = handle

    def _quadrature_points(self) -> int:
        return int(np.asarray(self.params["w"]).size)

    def set_kernel_params(self, values) -> None:
        import ctypes

        arr = (ctypes.c_double * 8)(*(list(map(float, values)) + [0.0] * 8)[:8])
        native.check(native.lib().fne_plan_set_params(self._plan, arr), "fne_plan_set_params")

    # ------------------------------------------------------------------------------------
    # vectors
    # ------------------------------------------------------------------------------------
    def create_vec(self, full_array_reordered=None) -> DeviceVec:
        if full_array_reordered is None:
            return DeviceVec.zeros(self.layout.n_free, self.device)
        arr = np.asarray(full_array_reordered, dtype=np.float64)
        return DeviceVec(torch.as_tensor(arr).to(self.device))

    def _as_device(self, u) -> torch.Tensor:
        if isinstance(u, DeviceVec):
            return u.t
        if isinstance(u, torch.Tensor):
            return u.to(device=self.device, dtype=torch.float64)
        return torch.as_tensor(np.asarray(u, dtype=np.float64)).to(self.device)

    # ------------------------------------------------------------------------------------
    # callbacks
    # ------------------------------------------------------------------------------------
    def energy_device(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        out = self._energy_out if out is None else out
        native.check(
            native.lib().fne_energy(
                self._plan,
                native.ptr(x),
                native.ptr(self.dvals),
                native.ptr(self._f_dev),
                native.ptr(out),
                native.stream_ptr(),
            ),
            "fne_energy",
        )
        return out

    def energy_fn(self, vec) -> float:
        t0 = time.perf_counter()
        val = float(self.energy_device(self._as_device(vec)).item())
        dt = time.perf_counter() - t0
        self._record_callback("energy", kernel=dt, total=dt)
        return val

    def gradient_device(self, x: torch.Tensor, g: torch.Tensor) -> None:
        native.check(
            native.lib().fne_gradient(
                self._plan,
                native.ptr(x),
                native.ptr(self.dvals),
                native.ptr(self._f_dev),
                native.ptr(g),
                native.stream_ptr(),
            ),
            "fne_gradient",
        )

    def gradient_fn(self, vec, g) -> None:
        t0 = time.perf_counter()
        if isinstance(g, DeviceVec):
            self.gradient_device(self._as_device(vec), g.t)
        else:
            gt = torch.empty(self.layout.n_free, dtype=torch.float64, device=self.device)
            self.gradient_device(self._as_device(vec), gt)
            g[:] = gt.cpu().numpy()
        dt = time.perf_counter() - t0
        self._record_callback("gradient", kernel=dt, total=dt)

    def hessian_device(self, x: torch.Tensor, g: torch.Tensor | None = None) -> None:
        self._hessian_state = x  # read by state-dependent preconditioner set-ups
        native.check(
            native.lib().fne_hessian(
                self._plan,
                native.ptr(x),
                native.ptr(self.dvals),
                native.ptr(self._f_dev),
                native.ptr(g),
                native.ptr(self.A.data),
                self.A.fne_precision,
                native.stream_ptr(),
            ),
            "fne_hessian",
        )

    def energy_grad_hessian_device(self, x: torch.Tensor, g: torch.Tensor | None = None,
                                   out: torch.Tensor | None = None) -> torch.Tensor:
        """One fused pass: energy (device scalar), gradient (optional) and Hessian into ``A``."""
        out = self._energy_out if out is None else out
        self._hessian_state = x
        native.check(
            native.lib().fne_energy_grad_hessian(
                self._plan,
                native.ptr(x),
                native.ptr(self.dvals),
                native.ptr(self._f_dev),
                native.ptr(out),
                native.ptr(g),
                native.ptr(self.A.data),
                self.A.fne_precision,
                native.stream_ptr(),
            ),
            "fne_energy_grad_hessian",
        )
        return out

    def assemble_hessian(self, u_owned, variant=2):
        del variant
        if self.local_hessian_mode != "element":
            return self.assemble_hessian_sfd(u_owned)
        return self.assemble_hessian_element(u_owned)

    # ------------------------------------------------------------------------------------
    # Hessian-vector products and the SFD Hessian mode
    # ------------------------------------------------------------------------------------
    def hvp_device(self, x, v, y) -> None:
        """y = H(x) v without assembling H (``fne_hvp``); device FP64 tensors on the free DOFs."""
        native.check(native.lib().fne_hvp(self._plan, native.ptr(x), native.ptr(self.dvals), native.ptr(v),
                                          native.ptr(y), native.stream_ptr()), "fne_hvp")

    def _setup_sfd(self) -> None:
        """Colouring of the SFD mode (``_setup_local_sfd``, ``reordered_element_base.py:1193-1302``):
        a distance-2 colouring of the free-DOF pattern (``fne_color_distance2``), then per colour
        the CSR positions whose column has that colour and their rows.  One Hessian-vector
        product with the colour's indicator vector then gives every such entry exactly: the
        columns of one colour never meet in a row."""
        import ctypes as C

        lay = self.layout
        n = int(lay.n_free)
        indptr = np.ascontiguousarray(lay.indptr, dtype=np.int32)
        indices = np.ascontiguousarray(lay.indices, dtype=np.int32)
        colors = np.empty(n, dtype=np.int32)
        i32 = C.POINTER(C.c_int32)
        k = native.lib().fne_color_distance2(n, indptr.ctypes.data_as(i32), indices.ctypes.data_as(i32),
                                             colors.ctypes.data_as(i32))
        native.check(min(int(k), 0), "fne_color_distance2")
        col_color = colors[indices]
        order = np.argsort(col_color, kind="stable")
        bounds = np.searchsorted(col_color[order], np.arange(int(k) + 1))
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(indptr.astype(np.int64)))
        self._sfd_n_colors = int(k)
        self._sfd_colors = torch.as_tensor(colors.astype(np.int64)).to(self.device)
        self._sfd_pos = [torch.as_tensor(order[bounds[c]:bounds[c + 1]]).to(self.device) for c in range(int(k))]
        self._sfd_rows = [torch.as_tensor(rows[order[bounds[c]:bounds[c + 1]]]).to(self.device) for c in range(int(k))]

    def assemble_hessian_sfd(self, u_owned):
        """Hessian values from one Hessian-vector product per colour (the reference's
        ``local_hessian_mode`` "sfd_local" / "sfd_local_vmap", ``reordered_element_base.py:
        1669-1731``, with JAX's jvp of the gradient replaced by ``fne_hvp``).  Both modes run
        the same products here; the values equal the element assembly's up to rounding."""
        t_total = time.perf_counter()
        x = self._as_device(u_owned)
        self._hessian_state = x
        n = self.layout.n_free
        ind = torch.empty(n, dtype=torch.float64, device=self.device)
        y = torch.empty(n, dtype=torch.float64, device=self.device)
        vals = torch.empty(self.A.nnz, dtype=torch.float64, device=self.device)
        t0 = time.perf_counter()
        for c in range(self._sfd_n_colors):
            ind.copy_(self._sfd_colors == c)
            self.hvp_device(x, ind, y)
            vals[self._sfd_pos[c]] = y[self._sfd_rows[c]]
        self.A.data.copy_(vals)
        t_k = time.perf_counter() - t0
        timings = {
            "elem_hessian_compute": 0.0,
            "scatter": 0.0,
            "coo_assembly": 0.0,
            "allgatherv": 0.0,
            "ghost_exchange": 0.0,
            "build_v_local": 0.0,
            "p2p_exchange": 0.0,
            "hvp_compute": t_k,
            "extraction": 0.0,
            "n_hvps": int(self._sfd_n_colors),
            "assembly_mode": "sfd_overlap_local" if self.local_hessian_mode == "sfd_local"
            else "sfd_overlap_local_vmap_hvpjit",
            "total": time.perf_counter() - t_total,
        }
        self.iter_timings.append(timings)
        self._record_callback("hessian", hvp_compute=t_k, total=timings["total"])
        return timings

    def assemble_hessian_element(self, u_owned):
        t_total = time.perf_counter()
        x = self._as_device(u_owned)
        t0 = time.perf_counter()
        self.hessian_device(x)
        t_k = time.perf_counter() - t0
        timings = {
            "elem_hessian_compute": t_k,
            "scatter": 0.0,
            "coo_assembly": 0.0,
            "allgatherv": 0.0,
            "ghost_exchange": 0.0,
            "build_v_local": 0.0,
            "p2p_exchange": 0.0,
            "hvp_compute": t_k,
            "extraction": 0.0,
            "n_hvps": 0,
            "assembly_mode": "element_patch_lds",
            "total": time.perf_counter() - t_total,
        }
        self.iter_timings.append(timings)
        self._record_callback("hessian", hvp_compute=t_k, total=timings["total"])
        return timings

    def element_eval(self, u_free, want_energy=True, want_grad=True, want_hess=True):
        """Per-element energy (n_elem,), gradient (n_elem, n_loc), Hessian (n_elem, n_loc, n_loc).

        Elements are returned in the mesh's original order.
        """
        x = self._as_device(u_free)
        ne, nl = self.n_elem, self.n_loc
        e = torch.empty(ne, dtype=torch.float64, device=self.device) if want_energy else None
        g = torch.empty(nl * ne, dtype=torch.float64, device=self.device) if want_grad else None
        h = torch.empty(nl * nl * ne, dtype=torch.float64, device=self.device) if want_hess else None
        native.check(
            native.lib().fne_element_eval(
                self._plan,
                native.ptr(x),
                native.ptr(self.dvals),
                native.ptr(e),
                native.ptr(g),
                native.ptr(h),
                native.stream_ptr(),
            ),
            "fne_element_eval",
        )
        inv = np.empty_like(self.patch_plan.elem_order)
        inv[self.patch_plan.elem_order] = np.arange(inv.size)
        inv_t = torch.as_tensor(inv).to(self.device)
        out = []
        if want_energy:
            out.append(e[inv_t])
        if want_grad:
            out.append(g.view(nl, ne).t()[inv_t])
        if want_hess:
            out.append(h.view(nl, nl, ne).permute(2, 0, 1)[inv_t])
        return tuple(out)

    def update_dirichlet(self, u_0_new) -> None:
        self.dirichlet_full = np.asarray(u_0_new, dtype=np.float64).ravel().copy()
        self.dvals.copy_(torch.as_tensor(self.dirichlet_full))

    # ------------------------------------------------------------------------------------
    # bookkeeping (same bucket names as the reference)
    # ------------------------------------------------------------------------------------
    def _record_callback(self, phase: str, **timings) -> None:
        stats = self._callback_stats[phase]
        stats["calls"] = int(stats.get("calls", 0)) + 1
        for k, v in timings.items():
            stats[k] = float(stats.get(k, 0.0)) + float(v)

    def setup_summary(self) -> dict[str, float]:
        return {str(k): float(v) for k, v in self._setup_timings.items()}

    def callback_summary(self) -> dict:
        return {ph: dict(st) for ph, st in self._callback_stats.items()}

    def memory_summary(self) -> dict:
        dev_bytes = sum(int(t.numel() * t.element_size()) for t in self._dev.values())
        csr_bytes = int(self.A.data.numel() * self.A.data.element_size()) + 4 * (self.A.nnz + self.A.n + 1)
        return {
            "device_plan_bytes": dev_bytes,
            "csr_bytes": csr_bytes,
            "owned_nnz": int(self.layout.nnz),
            "local_elements": int(self.n_elem),
            "patch_count": int(self.patch_plan.n_patch),
            "patch_halo_factor": float(self.patch_plan.halo_factor),
            "assembly_backend": "patch_lds",
            "matrix_type": self.A.getType(),
        }

    def cleanup(self) -> None:
        if getattr(self, "_plan", None) is not None and self._plan.value:
            native.lib().fne_plan_destroy(self._plan)
            self._plan = None
        self.ksp.destroy()
        self.A.destroy()
        self._dev = {}

    def __del__(self):
        try:
            if getattr(self, "_plan", None) is not None and self._plan.value:
                native.lib().fne_plan_destroy(self._plan)
                self._plan = None
        except Exception:
            pass
