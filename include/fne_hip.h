/*
 * fne_hip.h -- C ABI of the MI355X Newton-step assembly engine (libfne_hip.so).
 *
 * The reference (Beremi/fenics_nonlinear_energies) evaluates per-element energies,
 * gradients and Hessians with JAX closures and scatters them into a PETSc COO/CSR
 * matrix with numpy.  This library replaces that path with hand-written gfx950
 * kernels.  Every function is plain C: device pointers and sizes in, int status out.
 *
 * Conventions (mirroring the reference's only native FFI, core/coloring/coloring_custom.py:41-115
 * and core/petsc/trust_ksp.py:55-73):
 *   - return 0 on success, a negative FNE_E* code on failure; fne_last_error() gives text;
 *   - buffers are caller-owned device memory (hipMalloc / torch); the hot-path calls
 *     allocate nothing; plan creation allocates a small reduction scratch once;
 *   - numerical failure (NaN/inf in an energy) is NOT an error: it propagates to the
 *     caller exactly as the reference's Newton loop expects (core/petsc/minimizers.py:280-291);
 *   - all work is enqueued on the given HIP stream (NULL = default stream);
 *   - nothing here is thread-safe for concurrent calls on one plan.
 *
 * Free-DOF numbering: free DOFs are numbered 0..n_free-1 in the reordered order
 * (GlobalLayout.perm of core/petsc/reordered_element_base.py:24-39).  An element DOF
 * entry edof[a*n_elem + e] holds the reordered free index when the DOF is free and
 * n_free + t (t = total DOF index into the Dirichlet vector) when it is constrained,
 * which replaces local_vec_from_full (reordered_element_base.py:224-235).
 */
#ifndef FNE_HIP_H
#define FNE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* fne_stream_t; /* == hipStream_t */

/* status codes */
#define FNE_OK 0
#define FNE_EINVAL (-1)
#define FNE_EHIP (-2)
#define FNE_ENOMEM (-3)
#define FNE_EUNSUPPORTED (-4)

/* energy families (rows A1-A4 of the hot path) */
#define FNE_KIND_PLAPLACE 1       /* plaplace/jax_petsc/reordered_element_assembler.py:16-19 */
#define FNE_KIND_GINZBURG_LANDAU 2 /* ginzburg_landau/jax_petsc/reordered_element_assembler.py:18-24 */
#define FNE_KIND_NEOHOOKEAN 3     /* hyperelasticity/jax_petsc/parallel_hessian_dof.py:17-52 */
#define FNE_KIND_MOHR_COULOMB 4   /* slope_stability/jax/jax_energy.py:45-108 */

/* Hessian / SpMV value precision */
#define FNE_FP64 0
#define FNE_FP32 1

/*
 * Plan descriptor.  Every pointer is a DEVICE pointer owned by the caller and must stay
 * valid until fne_plan_destroy.  Layouts are structure-of-arrays so that a wavefront
 * reading element e..e+63 issues coalesced loads.
 */
typedef struct fne_plan_desc {
  int32_t kind;      /* FNE_KIND_* */
  int32_t n_elem;    /* elements in the energy set (all elements of the mesh) */
  int32_t n_loc;     /* DOFs per element: 3 (P1 scalar), 12 (P1 tet, 3 comps), 30 (P4, 2 comps) */
  int32_t n_geo;     /* geometry doubles per element */
  int32_t n_free;    /* free DOFs = CSR rows = CSR cols */
  int32_t n_total;   /* length of the Dirichlet value vector */
  int64_t nnz;       /* CSR nonzeros */
  const int32_t* edof;    /* [n_loc][n_elem] element DOF map (see header comment) */
  const double* egeo;     /* [n_geo][n_elem] element geometry, FP64 */
  const double* eweight;  /* [n_elem] energy weights or NULL (all 1) */
  const int32_t* indptr;  /* [n_free+1] CSR row pointer */
  const int32_t* indices; /* [nnz] CSR column indices, ascending per row */
  /* patch schedule for the owner-computes gradient/Hessian assembly */
  int32_t n_patch;        /* number of row patches (one workgroup each) */
  int32_t chunk;          /* elements staged in LDS per phase */
  int32_t max_patch_nnz;  /* max CSR entries of one patch (LDS slice size) */
  int32_t max_row_len;    /* max CSR row length */
  const int32_t* prow;       /* [n_patch+1] first row of each patch */
  const int32_t* pelem_ptr;  /* [n_patch+1] offsets into pelem */
  const int32_t* pelem;      /* patch element lists (element ids, owned first) */
  const int32_t* inc_ptr;    /* [n_inc_rows+1] incidence list offsets, one list per assembly row */
  const uint32_t* inc;       /* incidence items (packed slot/local index/column positions) */
  const int32_t* inc_aux;    /* family-specific auxiliary table (NULL for P1) */
  /* patch-major copies of the element data (owned elements of a patch first, then its halo),
     so a patch workgroup streams its elements without an indirection */
  int64_t n_pe;              /* total patch elements = pelem_ptr[n_patch] */
  const int32_t* pedof;      /* [n_loc][n_pe] */
  const double* pgeo;        /* [n_geo][n_pe] */
  const int32_t* pown;       /* [n_patch] owned elements per patch (energy is summed over these) */
  const uint64_t* pitem;     /* [n_pe] colored-patch item per patch element (FNE_FLAG_COLORED) */
  const int32_t* pcolors;    /* [n_patch] number of colors of each patch (FNE_FLAG_COLORED) */
  /* problem parameters */
  double param[8];        /* p | eps | C1,D1 | E,nu,phi_deg,cohesion,reg */
  const double* qdata;    /* quadrature / tabulation table (GL: ip[3][nq] then w[nq]; with
                             FNE_FLAG_GL_PRODUCTS then w[q] ip[a][q] ip[b][q] for ab = 00 11 22 01 02 12,
                             [6][nq]: 10 nq doubles in all, which the caller must provide) */
  int32_t nq;             /* quadrature points */
  int32_t flags;          /* FNE_FLAG_* */
  /* P1 with FNE_FLAG_NODE_GEO: vertex coordinates (x, y) of every DOF code, [n_free + n_total][2]
     (code c < n_free: free DOF c; else Dirichlet entry c - n_free); egeo/pgeo are not read */
  const double* dof_xy;
} fne_plan_desc;

#define FNE_FLAG_ABS_DET 1 /* Neo-Hookean: |det F| (serial path hyperelasticity/jax/jax_energy.py:26) */
#define FNE_FLAG_COLORED 4 /* P1: colored patch assembly.  pitem = pos_ab[5(3a+b) : +5) (31 = none)
                               | owned-row mask [45:48) | energy owner [48] | color [52:60) */
#define FNE_FLAG_STAGED 8 /* P1: staged patch assembly.  Elements of a patch are evaluated into LDS,
                              then one thread per row sums its incidence list (inc_ptr/inc: slot[0:13)
                              | a[13:15) | pos_b[15+5b:20+5b), 31 = none) in registers */
#define FNE_FLAG_NODE_GEO 16 /* P1: element geometry (dvx, dvy, vol) evaluated in the kernels from the
                                vertex coordinates dof_xy with the host's operation order
                                (det = (x1-x0)(y2-y0) - (x2-x0)(y1-y0), dvx = (y1-y2, y2-y0, y0-y1)/det,
                                dvy = (x2-x1, x0-x2, x1-x0)/det, vol = |det|/2), in place of reading
                                n_geo = 7 doubles per element */
#define FNE_FLAG_ROW_GATHER 2 /* P1: one thread per CSR row gathers its incident elements (no LDS
                                 staging); inc holds uint64 items e[0:32) a[32:34) pos_b[40+8b:48+8b) */
#define FNE_FLAG_GL_PRODUCTS 32 /* Ginzburg-Landau: qdata carries the quadrature products after w; with
                                   nq == 3 the kernels evaluate energy, gradient and Hessian in one
                                   fused-multiply-add pass over them */

typedef struct fne_plan fne_plan;

/* Replaces ReorderedElementAssemblerBase.__init__'s kernel build + scatter build
 * (core/petsc/reordered_element_base.py:803-819). */
int fne_plan_create(const fne_plan_desc* desc, fne_plan** out);
int fne_plan_destroy(fne_plan* plan);
/* Replaces update of model parameters between continuation steps (e.g. Davis-B reduced
 * cohesion/friction, slope_stability/jax_petsc/solver.py:600-603). */
int fne_plan_set_params(fne_plan* plan, const double* param8);
/* Mohr-Coulomb only: 0 = full MC energy, 1 = elastic tangent/energy
 * (slope_stability/jax_petsc/reordered_element_assembler.py:305-311). */
int fne_plan_set_mode(fne_plan* plan, int32_t constitutive_mode);

/*
 * Energy: out[0] = sum_e w_e E_e(v) - sum_i f[i] x[i]  (FP64, compensated, deterministic).
 * Replaces energy_fn (core/petsc/reordered_element_base.py:1579-1611).
 * x: [n_free] free values, dvals: [n_total] Dirichlet values, f: [n_free] load or NULL,
 * out: device FP64 scalar.
 */
int fne_energy(fne_plan* plan, const double* x, const double* dvals, const double* f,
               double* out, fne_stream_t stream);

/*
 * Gradient: g[i] = sum over incident elements of dE_e/dv - f[i].
 * Replaces gradient_fn (core/petsc/reordered_element_base.py:1613-1633).
 */
int fne_gradient(fne_plan* plan, const double* x, const double* dvals, const double* f,
                 double* g, fne_stream_t stream);

/*
 * Hessian values into the plan's fixed CSR pattern: vals[k] for k < nnz, FP64 or FP32.
 * Replaces assemble_hessian_element's JAX element Hessian + np.add.at scatter + setValuesCOO
 * (core/petsc/reordered_element_base.py:1733-1775).  g may be non-NULL to receive the
 * gradient from the same pass (fused gradient + Hessian).
 */
int fne_hessian(fne_plan* plan, const double* x, const double* dvals, const double* f,
                double* g, void* vals, int32_t precision, fne_stream_t stream);

/*
 * Fused Newton-step assembly: energy (out, FP64 scalar, may be NULL), gradient (g, may be
 * NULL) and Hessian values (vals) in one pass over the elements.  Same definitions as
 * fne_energy / fne_gradient / fne_hessian.
 */
int fne_energy_grad_hessian(fne_plan* plan, const double* x, const double* dvals, const double* f,
                            double* out, double* g, void* vals, int32_t precision,
                            fne_stream_t stream);

/*
 * Per-element energy / gradient / Hessian (dense, for the autograd boundary and tests):
 * e_out[n_elem] (may be NULL), ge_out[n_loc][n_elem] (may be NULL),
 * he_out[n_loc*n_loc][n_elem] (may be NULL).  Replaces the _make_local_element_kernels
 * triple (reordered_element_base.py:922-923).
 */
int fne_element_eval(fne_plan* plan, const double* x, const double* dvals, double* e_out,
                     double* ge_out, double* he_out, fne_stream_t stream);

/*
 * Matrix-free Hessian products (FP64), element by element without forming H:
 * fne_hvp: y = H(x) v over the free DOFs (v is taken as zero on constrained DOFs), for the
 * Mohr-Coulomb P2/P4 plans (FNE_FLAG_ROW_GATHER; H is the return-mapping tangent at x in the
 * plan's constitutive mode), the P1 families with staged patch plans (FNE_FLAG_STAGED) and the
 * tetrahedral plans with FNE_FLAG_ROW_GATHER; other plan forms: FNE_EUNSUPPORTED.
 * fne_hessian_diag: diag = diagonal of H(x), Mohr-Coulomb plans only.
 * Replace the matrix-free operator and its diagonal
 * (slope_stability/jax_petsc/reordered_element_assembler.py:313-562) and the jvp of the
 * gradient behind the SFD Hessian mode (core/petsc/reordered_element_base.py:1295-1302).
 * y and v must not alias.
 */
int fne_hvp(fne_plan* plan, const double* x, const double* dvals, const double* v, double* y, fne_stream_t stream);
int fne_hessian_diag(fne_plan* plan, const double* x, const double* dvals, double* diag, fne_stream_t stream);

/*
 * CSR SpMV y = A x, A values FP64 or FP32, x/y FP64.  Replaces PETSc MatMult
 * (hessian_matvec_fn = A.mult, core/petsc/scalar_problem_driver.py:294).
 */
int fne_spmv(int32_t n_rows, const int32_t* indptr, const int32_t* indices, const void* vals,
             int32_t precision, const double* x, double* y, fne_stream_t stream);

/* CSR diagonal (Jacobi preconditioner set-up): diag[i] = A[i,i] (FP64 out). */
int fne_csr_diag(int32_t n_rows, const int32_t* indptr, const int32_t* indices, const void* vals,
                 int32_t precision, double* diag, fne_stream_t stream);

/*
 * Numeric sparse product C = A B into C's fixed CSR pattern (c_indptr/c_indices, sorted
 * columns; it must contain every column the row's products reach).  A values FP64 or FP32
 * (a_precision), B and C FP64; A's column indices must be < B's row count.  Deterministic:
 * each entry sums its products in the order of A's row.  Used by the multigrid set-up for
 * the smoothed prolongator and the Galerkin products P^T A P, which PETSc GAMG recomputes at
 * every PCSetUp (pc_type "gamg", hyperelasticity/jax_petsc/solver.py:138-143; the PC is set
 * up per Newton iteration through ksp.setOperators/solve in core/petsc/scalar_problem_driver.py).
 */
/*
 * Golden-section line search on the device (replaces golden_section_search,
 * core/petsc/minimizers.py:14-64, with the same points and comparisons).  state: device FP64
 * array of fne_golden_state_len() slots, initialised by the caller to zero except
 * state[0] = lo, state[1] = hi, state[10] = lo (first point), state[13] = tol.  Per point:
 * fne_golden_trial writes trial = x + alpha p for the state's next alpha (no-op once finished),
 * the caller evaluates the energy of trial into f_new (e.g. fne_energy), and fne_golden_step
 * consumes it and advances the state.  Finished when state[11] >= 7; then state[8] is the chosen
 * alpha, state[9] its energy (non-finite energies count as +inf) and state[12] the number of
 * energies used.  Extra steps after the finish are no-ops.
 */
int fne_golden_state_len(void);
int fne_golden_trial(int32_t n, const double* x, const double* p, const double* state, double* trial,
                     fne_stream_t stream);
int fne_golden_step(double* state, const double* f_new, fne_stream_t stream);

/*
 * Tentative prolongator of the multigrid set-up: thin QR of every aggregate's rows of the
 * near-nullspace B (n x k row-major FP64, k in {1, 2, 3, 6}); aggregate a holds rows
 * [a*agg, min(n, (a+1)*agg)), k <= agg <= 256.  q_vals receives the Q blocks in CSR order
 * (aggregate a's m x min(m, k) block, row-major, at offset a*agg*k), bc the R factors (the
 * coarse near-nullspace, min(m, k) x k row-major at offset a*k*k).  R has a non-negative
 * diagonal.  Stands in for the near-nullspace interpolation of PETSc GAMG
 * (MatSetNearNullSpace + pc_type "gamg", hyperelasticity/jax_petsc/solver.py:138-217).
 */
int fne_aggregate_qr(int32_t n, int32_t k, int32_t agg, const double* B, double* q_vals, double* bc,
                     fne_stream_t stream);

/*
 * Fused multigrid operator, one CSR pass (precision: FNE_FP64/FNE_FP32 | lanes_per_row << 8):
 *   op 0: y = w.r + w.(r - A (w.r))   first damped-Jacobi sweep from zero (x unused)
 *   op 1: y = x + w.(r - A x)          damped-Jacobi sweep
 *   op 2: y = r - A x                  residual
 *   op 3: y = y_in + A x               prolongation added to the iterate (y may equal y_in)
 * x, r, w, y_in, y: FP64 device vectors; y must not alias x, r or w.  The smoother, residual
 * and prolongation of the multigrid V-cycle (stands in for the level smoothers of PETSc
 * GAMG, pc_type "gamg", hyperelasticity/jax_petsc/solver.py:138-143).
 */
int fne_vcycle_op(int32_t op, int32_t n_rows, const int32_t* indptr, const int32_t* indices, const void* vals,
                  int32_t precision, const double* x, const double* r, const double* w, const double* y_in,
                  double* y, fne_stream_t stream);

/*
 * Chebyshev smoothing step of the multigrid levels, one CSR pass:
 *   y = x + c (x - x_prev) + w.(r - A x)      (x_prev may be NULL when c = 0)
 * With w = (2 rho_k / delta) D^-1 and c = rho_k rho_{k-1} it is step k of the Chebyshev iteration
 * for D^-1 A on [a, b]; step 0 is w = D^-1 / theta, c = 0 (theta, delta: centre and half-width of
 * [a, b]).  Stands in for PETSc's KSPCHEBYSHEV level smoother with Jacobi, the default smoother
 * of PCGAMG (pc_type "gamg", hyperelasticity/jax_petsc/solver.py:138-143).  y may alias x_prev,
 * not x, r or w.  precision as in fne_vcycle_op.
 */
int fne_vcycle_cheb(int32_t n_rows, const int32_t* indptr, const int32_t* indices, const void* vals,
                    int32_t precision, const double* x, const double* x_prev, const double* r, const double* w,
                    double c, double* y, fne_stream_t stream);

int fne_spgemm_numeric(int32_t n_rows, const int32_t* a_indptr, const int32_t* a_indices, const void* a_vals,
                       int32_t a_precision, const int32_t* b_indptr, const int32_t* b_indices,
                       const double* b_vals, const int32_t* c_indptr, const int32_t* c_indices, double* c_vals,
                       fne_stream_t stream);

/*
 * Preconditioned CG on the device (KSPCG semantics: zero initial guess, point-Jacobi left
 * preconditioning via inv_diag (NULL = none), convergence when the preconditioned residual
 * norm <= max(rtol*initial, atol), divergence at dtol*initial, indefinite checks).  Replaces
 * ksp.solve with ksp_type "cg" (core/petsc/scalar_problem_driver.py:196).  Scalars stay on
 * the device; the host reads the stop flag every check_every iterations.  work: device
 * buffer of fne_pcg_workspace_bytes(n, nnz) bytes.  Outputs: iterations, PETSc-style
 * converged reason (2 rtol, 3 atol, -3 max_it, -4 dtol, -8 indefinite PC, -9 nan/inf,
 * -10 indefinite matrix) and the final preconditioned residual norm (host pointers).
 */
int64_t fne_pcg_workspace_bytes(int32_t n, int64_t nnz);
int fne_pcg(int32_t n, const int32_t* indptr, const int32_t* indices, const void* vals, int32_t precision,
            const double* inv_diag, const double* b, double* x, double rtol, double atol, double dtol,
            int32_t max_it, int32_t check_every, void* work, int64_t work_bytes, int32_t* its_out,
            int32_t* reason_out, double* rnorm_out, fne_stream_t stream);

/*
 * The same CG with a preconditioner given as a recorded HIP graph (e.g. the multigrid
 * V-cycle, pc_type "gamg"): pc_graph_exec is a hipGraphExec_t whose launch computes
 * pc_out = M pc_in, reading nothing else that the solve writes.  The residual lives in pc_in
 * and the preconditioned residual in pc_out, so the solve overwrites pc_in.  Same semantics,
 * workspace and outputs as fne_pcg; the iteration runs on the device, the host reads the
 * stop flag every check_every iterations (iterations enqueued after the stop are no-ops
 * except for the graph launch).  Replaces ksp.solve with ksp_type "cg" and pc_type "gamg"
 * (hyperelasticity/jax_petsc/solver.py:138-143; core/petsc/scalar_problem_driver.py:196).
 */
int fne_pcg_graph(int32_t n, const int32_t* indptr, const int32_t* indices, const void* vals, int32_t precision,
                  void* pc_graph_exec, double* pc_in, const double* pc_out, const double* b, double* x, double rtol,
                  double atol, double dtol, int32_t max_it, int32_t check_every, void* work, int64_t work_bytes,
                  int32_t* its_out, int32_t* reason_out, double* rnorm_out, fne_stream_t stream);

/*
 * Steihaug-Toint CG (PETSc KSPSTCG, ksp_type "stcg"), the trust-region subproblem solve:
 * approximately minimise -b.x + x.A x / 2 over ||x||_2 <= radius (radius >= 0, INFINITY =
 * no bound).  Replaces ksp.solve after KSPCGSetRadius (core/petsc/trust_ksp.py:55-72,
 * core/petsc/scalar_problem_driver.py:143-170; the reference's serial restatement is
 * core/serial/sparse_solvers.py:20-93).  Same arguments, workspace and outputs as fne_pcg /
 * fne_pcg_graph plus the radius; differences: convergence on the unpreconditioned residual
 * norm ||r|| (also the reported norm); curvature p.Ap <= 0 moves x along p to the sphere
 * and returns reason 5 (with radius = INFINITY x stays at the last iterate); a step that
 * would leave the ball moves x to the sphere and returns reason 6.
 */
int fne_pstcg(int32_t n, const int32_t* indptr, const int32_t* indices, const void* vals, int32_t precision,
              const double* inv_diag, const double* b, double* x, double radius, double rtol, double atol,
              double dtol, int32_t max_it, int32_t check_every, void* work, int64_t work_bytes, int32_t* its_out,
              int32_t* reason_out, double* rnorm_out, fne_stream_t stream);
int fne_pstcg_graph(int32_t n, const int32_t* indptr, const int32_t* indices, const void* vals, int32_t precision,
                    void* pc_graph_exec, double* pc_in, const double* pc_out, const double* b, double* x,
                    double radius, double rtol, double atol, double dtol, int32_t max_it, int32_t check_every,
                    void* work, int64_t work_bytes, int32_t* its_out, int32_t* reason_out, double* rnorm_out,
                    fne_stream_t stream);

/*
 * Replaces ksp.solve with ksp_type="gmres" (left preconditioning; the reference's
 * Ginzburg-Landau runs, ginzburg_landau/jax_petsc/solver.py:31-32, through
 * scalar_problem_driver.py:174-196) and ksp_type="fgmres" (flexible = 1: right preconditioning
 * with the preconditioned basis kept; the slope p-multigrid runs, slope_stability/jax_petsc/
 * solver.py:1275).  PETSc KSPGMRES / KSPFGMRES semantics: zero initial guess, restart length
 * `restart` (1..512), classical Gram-Schmidt, Givens rotations; GMRES tests the preconditioned
 * residual estimate against max(rtol ||B b||, atol), FGMRES the unpreconditioned one; dtol
 * divergence, happy breakdown.  Preconditioner: pc_graph_exec (a recorded graph mapping pc_in
 * to pc_out, as fne_pcg_graph) when given, else inv_diag (Jacobi), else none.  The host reads
 * the device state every check_every Arnoldi steps (0: only at the end of a cycle).  Workspace:
 * fne_gmres_workspace_bytes(n, restart, flexible) bytes of device memory.  Reasons: 2 rtol,
 * 3 atol, -3 iteration cap, -4 dtol, -9 NaN/inf.
 */
int64_t fne_gmres_workspace_bytes(int32_t n, int32_t restart, int32_t flexible);
int fne_gmres(int32_t n, const int32_t* indptr, const int32_t* indices, const void* vals, int32_t precision,
              const double* inv_diag, void* pc_graph_exec, double* pc_in, const double* pc_out, const double* b,
              double* x, double rtol, double atol, double dtol, int32_t max_it, int32_t restart, int32_t flexible,
              int32_t check_every, void* work, int64_t work_bytes, int32_t* its_out, int32_t* reason_out,
              double* rnorm_out, fne_stream_t stream);

/*
 * fne_pcg without a workspace argument, for callers that cannot allocate device memory (code
 * that must not call hipMalloc, such as tight solve loops or harnesses that forbid it).  No
 * counterpart in the reference.  Same arguments, semantics and outputs as fne_pcg minus
 * work / work_bytes; the solve runs in a static device array of FNE_PCG_NOALLOC_MAX_BYTES
 * (16 MiB per GPU, held from library load to process exit): solves needing more than
 * fne_pcg_workspace_bytes(n, nnz) <= FNE_PCG_NOALLOC_MAX_BYTES (about 300k rows) are rejected
 * with FNE_EINVAL.  All calls on one GPU share that array: calls that overlap in time on one
 * GPU (other streams or threads) race.  Allocates nothing.
 */
#define FNE_PCG_NOALLOC_MAX_BYTES (16ll << 20)
int fne_pcg_noalloc(int32_t n, const int32_t* indptr, const int32_t* indices, const void* vals, int32_t precision,
                    const double* inv_diag, const double* b, double* x, double rtol, double atol, double dtol,
                    int32_t max_it, int32_t check_every, int32_t* its_out, int32_t* reason_out, double* rnorm_out,
                    fne_stream_t stream);

/*
 * Deterministic compensated FP64 dot product: out[0] = sum_i a[i]*b[i] (the PETSc VecDot /
 * VecNorm of the reference's Newton loop, core/petsc/minimizers.py:286, 329, 1003).
 * work: caller-owned device scratch of fne_dot_work_len() (sum, compensation) pairs
 * (16 bytes each), written by the call; calls in flight at the same time need distinct work
 * buffers.  Allocates nothing.
 */
int32_t fne_dot_work_len(void);
int fne_dot(int64_t n, const double* a, const double* b, double* out, void* work, fne_stream_t stream);

/*
 * Host (CPU) helper for plan construction: greedy coloring of patch elements so that no two
 * elements of one color share an owned row of their patch.  rows[3*i+a] is the owned row of
 * local vertex a of patch element i (-1 if not owned); patch elements are grouped by patch via
 * pelem_ptr.  colors[i] receives the color (0..63).  Returns the largest color count of a patch,
 * or FNE_EINVAL if some patch needs more than 64 colors.  Replaces nothing in the reference
 * (its scatter is serial np.add.at); the coloring is what makes the GPU reduction race-free.
 */
int fne_color_patch_elements(int32_t n_patch, const int64_t* pelem_ptr, const int64_t* rows,
                             int32_t n_loc, int64_t n_rows, int32_t* colors);

/*
 * Host (CPU) helper of the SFD Hessian mode: greedy distance-2 colouring of a structurally
 * symmetric CSR pattern (n rows, column indices < n).  Two rows that share a column, or whose
 * columns meet (distance <= 2 in the pattern's graph), get different colours, so the columns
 * of one colour are structurally orthogonal and one Hessian-vector product per colour
 * recovers every entry.  Rows are coloured in order with the smallest colour free of their
 * distance-2 neighbourhood.  colors[n] receives 0..k-1; returns k, or FNE_EINVAL.  Replaces
 * the colouring of the reference's SFD set-up (igraph's greedy vertex colouring of the
 * pattern's square, core/petsc/reordered_element_base.py:1193-1240; the C colouring
 * core/coloring/custom_coloring.c:37-155); the colouring decides only how many products are
 * taken, not the Hessian's values.
 */
int fne_color_distance2(int32_t n, const int32_t* indptr, const int32_t* indices, int32_t* colors);

/*
 * Learned warm start: per-node MLP (BF16 MFMA products, FP32 activations and optimiser state).
 * No counterpart in the reference, which warm-starts every load / continuation step from the
 * previous converged state (core/petsc/load_step_driver.py:58-110,
 * hyperelasticity/jax_petsc/solver.py:351-359); that state stays the fallback.
 * Layer l maps dims[l] -> dims[l+1]: y = W_l x + b_l, W_l [dims[l+1]][dims[l]] BF16 bits,
 * b_l FP32, tanh after every layer but the last.  Row-major node arrays: x [n][dims[0]].
 */
#define FNE_MLP_MAX_LAYERS 8
typedef struct fne_mlp_desc {
  int32_t n_layers;
  int32_t dims[FNE_MLP_MAX_LAYERS + 1];
  const uint16_t* w[FNE_MLP_MAX_LAYERS];
  const float* b[FNE_MLP_MAX_LAYERS];
} fne_mlp_desc;

/* y [n][dims[L]]; acts (optional unless training) receives the hidden activations of layers
 * 0..L-2 back to back ([n][dims[1]], [n][dims[2]], ...). */
int fne_mlp_forward(const fne_mlp_desc* net, int64_t n, const float* x, float* y, float* acts,
                    fne_stream_t stream);
/* Bytes of the workspace fne_mlp_backward needs for n nodes (negative on bad arguments). */
int64_t fne_mlp_workspace_bytes(const fne_mlp_desc* net, int64_t n);
/* Gradients of sum_i dy[i] . y[i]: dw[l] [dims[l+1]][dims[l]] and db[l] [dims[l+1]] (FP32),
 * from the inputs x and the activations saved by fne_mlp_forward.  Deterministic. */
int fne_mlp_backward(const fne_mlp_desc* net, int64_t n, const float* x, const float* acts,
                     const float* dy, float* const* dw, float* const* db, void* workspace,
                     int64_t ws_bytes, fne_stream_t stream);
/* One Adam step (bias-corrected; weight_decay adds weight_decay * w to the gradient) on n FP32
 * master weights; w_bf16 (may be NULL) receives the rounded copy used by the products. */
int fne_mlp_adam(int64_t n, float* w, uint16_t* w_bf16, const float* grad, float* m, float* v,
                 float lr, float beta1, float beta2, float eps, float weight_decay, int32_t step,
                 fne_stream_t stream);
/* Round FP32 weights to BF16 (nearest even). */
int fne_mlp_to_bf16(int64_t n, const float* w, uint16_t* w_bf16, fne_stream_t stream);

const char* fne_last_error(void);
/* Library/ABI version string. */
const char* fne_version(void);

#ifdef __cplusplus
}
#endif
#endif /* FNE_HIP_H */
