// Device-resident GMRES(m) and FGMRES(m) around the CSR SpMV.
//
// PETSc's KSPGMRES as the reference's Ginzburg-Landau runs drive it (ksp_type "gmres",
// ginzburg_landau/jax_petsc/solver.py:31-32; scalar_problem_driver.py:174-196): zero initial
// guess, left preconditioning, restart length m, classical Gram-Schmidt without refinement
// (PETSc's default orthogonalisation), Givens rotations on the Hessenberg matrix, convergence on
// the preconditioned residual estimate |g_{k+1}| <= max(rtol ||B b||, atol), divergence at
// dtol ||B b||, happy breakdown at h_{k+1,k} = 0.  KSPFGMRES (the outer solver of the slope
// p-multigrid runs, slope_stability/jax_petsc/solver.py:1275): right preconditioning with the
// preconditioned vectors z_k = B v_k kept, x += Z y, residual norms of b - A x.
//
// Every scalar lives in device memory: the Hessenberg column, the rotations, g, the norms, the
// iteration count and the flags.  One Arnoldi step is a fixed sequence of launches -- product
// (SpMV in the CSR-stream or lanes-per-row form, and the preconditioner: Jacobi scaling or a
// recorded HIP graph between two copies), the k + 1 projections as fixed-order block partials,
// their reduction, the update w -= V h with ||w||^2 partials, and a one-workgroup step that
// appends the column, applies the rotations and tests convergence -- and the host reads the
// flags only every check_every steps and at the end of a cycle.  Steps enqueued after the end
// of a cycle are no-ops.  All sums are compensated FP64 sums in a fixed order: the solve is
// bitwise reproducible.
#include "fne_common.h"

#include <algorithm>
#include <cmath>
#include <string>

namespace {

// state header slots (doubles); the header is followed by H[(m+1) m] (column-major,
// H[i + (m+1) k]), g[m+1], cs[m], sn[m], y[m], hcol[m+1]
enum : int { gRnorm0 = 0, gRnorm, gIts, gReason, gDone, gEnd, gKUsed, gInv, gFirst, gHdr = 16 };

constexpr int kConvRtol = 2, kConvAtol = 3, kDivDtol = -4, kDivNan = -9;
constexpr int kGmMaxRestart = 512;
constexpr int kGmBlocks = 512;  // workgroups of the projection and update passes (fixed: fixed order)

struct GmParams {
  int n;
  int m;
  int max_it;
  double rtol, atol, dtol;
};

__host__ __device__ __forceinline__ int64_t gm_state_len(int m) {
  return gHdr + int64_t(m + 1) * m + (m + 1) + 3 * int64_t(m) + (m + 1);
}
__device__ __forceinline__ double* gm_H(double* st) { return st + gHdr; }
__device__ __forceinline__ double* gm_g(double* st, int m) { return gm_H(st) + (m + 1) * m; }
__device__ __forceinline__ double* gm_cs(double* st, int m) { return gm_g(st, m) + (m + 1); }
__device__ __forceinline__ double* gm_sn(double* st, int m) { return gm_cs(st, m) + m; }
__device__ __forceinline__ double* gm_y(double* st, int m) { return gm_sn(st, m) + m; }
__device__ __forceinline__ double* gm_hcol(double* st, int m) { return gm_y(st, m) + m; }

__device__ __forceinline__ int gm_converged(int it, double rn, double rn0, const GmParams& P) {
  if (!isfinite(rn)) return kDivNan;
  if (rn <= fmax(P.rtol * rn0, P.atol)) return rn > P.atol ? kConvRtol : kConvAtol;
  if (it > 0 && rn >= P.dtol * rn0) return kDivDtol;
  return 0;
}

__device__ __forceinline__ void gm_merge(const double2* part, int nb, double& s, double& c) {
  s = 0.0;
  c = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) acc_merge(s, c, part[i].x, part[i].y);
  block_reduce_pair<1024>(s, c);
}

// rows [lo, hi) of workgroup b in the fixed partition of the projection / update passes
__device__ __forceinline__ void gm_rows(int n, int& lo, int& hi) {
  const int chunk = (n + int(gridDim.x) - 1) / int(gridDim.x);
  lo = min(n, int(blockIdx.x) * chunk);
  hi = min(n, lo + chunk);
}

// out = b - A x (flag slot: the kernel is a no-op once st[slot] != 0)
template <typename VT, int LPR, int FORM>
__global__ __launch_bounds__(256) void gm_residual_kernel(int n, const int32_t* __restrict__ indptr,
                                                          const int32_t* __restrict__ indices,
                                                          const VT* __restrict__ vals, const double* __restrict__ x,
                                                          const double* __restrict__ b, double* __restrict__ out,
                                                          const double* __restrict__ st, int slot) {
  if (st[slot] != 0.0) return;  // uniform over the grid
  __shared__ double prod[FORM == 2 ? kStreamCap : 1];
  int row;
  bool leader;
  const double acc =
      csr_form_dot<VT, LPR, FORM>(n, indptr, indices, vals, [&](int c) { return x[c]; }, prod, row, leader);
  if (leader) out[row] = b[row] - acc;
}

// out = A in
template <typename VT, int LPR, int FORM>
__global__ __launch_bounds__(256) void gm_spmv_kernel(int n, const int32_t* __restrict__ indptr,
                                                      const int32_t* __restrict__ indices,
                                                      const VT* __restrict__ vals, const double* __restrict__ in,
                                                      double* __restrict__ out, const double* __restrict__ st,
                                                      int slot) {
  if (st[slot] != 0.0) return;
  __shared__ double prod[FORM == 2 ? kStreamCap : 1];
  int row;
  bool leader;
  const double acc =
      csr_form_dot<VT, LPR, FORM>(n, indptr, indices, vals, [&](int c) { return in[c]; }, prod, row, leader);
  if (leader) out[row] = acc;
}

// out = D^-1 in (inv_diag null: out = in)
__global__ __launch_bounds__(kBlock) void gm_scale_diag_kernel(int n, const double* __restrict__ inv_diag,
                                                               const double* __restrict__ in, double* __restrict__ out,
                                                               const double* __restrict__ st, int slot) {
  if (st[slot] != 0.0) return;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = inv_diag ? inv_diag[i] * in[i] : in[i];
}

// out = st[gInv] * in
__global__ __launch_bounds__(kBlock) void gm_normalize_kernel(int n, const double* in, double* out,
                                                              const double* __restrict__ st) {
  if (st[gEnd] != 0.0) return;
  const double s = st[gInv];
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = s * in[i];
}

// part[j * gridDim + b] = sum over the workgroup's rows of V[j][i] w[i], j < k1 (w may be V[0])
__global__ __launch_bounds__(kBlock) void gm_dots_kernel(int n, const double* V, int k1, const double* w,
                                                         double2* __restrict__ part,
                                                         const double* __restrict__ st, int slot) {
  if (st[slot] != 0.0) return;
  int lo, hi;
  gm_rows(n, lo, hi);
  for (int j = 0; j < k1; ++j) {
    const double* v = V + size_t(j) * n;
    double s = 0.0, c = 0.0;
    for (int i = lo + int(threadIdx.x); i < hi; i += blockDim.x) acc_add(s, c, v[i] * w[i]);
    block_reduce_pair(s, c);
    if (threadIdx.x == 0) part[size_t(j) * gridDim.x + blockIdx.x] = make_double2(s, c);
    __syncthreads();
  }
}

// cycle start: beta = ||v0|| from the partials; the first cycle also sets rnorm0 and tests
// convergence at iteration 0
__global__ __launch_bounds__(1024) void gm_start_finalize(GmParams P, const double2* __restrict__ part, int nb,
                                                          double* __restrict__ st) {
  if (st[gDone] != 0.0) return;
  double s, c;
  gm_merge(part, nb, s, c);
  if (threadIdx.x != 0) return;
  const double beta = sqrt(fmax(s + c, 0.0));
  const int m = P.m;
  double* g = gm_g(st, m);
  for (int i = 0; i <= m; ++i) g[i] = 0.0;
  g[0] = beta;
  st[gKUsed] = 0.0;
  int reason = 0;
  if (st[gFirst] != 0.0) {
    st[gFirst] = 0.0;
    st[gRnorm0] = beta;
    st[gRnorm] = beta;
    reason = gm_converged(0, beta, beta, P);
  }
  if (!reason && beta == 0.0) reason = kConvAtol;
  st[gReason] = reason;
  st[gDone] = reason ? 1.0 : 0.0;
  st[gEnd] = reason ? 1.0 : 0.0;
  st[gInv] = beta > 0.0 ? 1.0 / beta : 0.0;
}

// hcol[j] = projection j, j < k1
__global__ __launch_bounds__(1024) void gm_hcol_finalize(GmParams P, const double2* __restrict__ part, int nb, int k1,
                                                         double* __restrict__ st) {
  if (st[gEnd] != 0.0) return;
  double* hcol = gm_hcol(st, P.m);
  for (int j = 0; j < k1; ++j) {
    double s, c;
    gm_merge(part + size_t(j) * nb, nb, s, c);
    if (threadIdx.x == 0) hcol[j] = s + c;
    __syncthreads();
  }
}

// w -= sum_j hcol[j] V[j] (j ascending); partials of ||w||^2
__global__ __launch_bounds__(kBlock) void gm_update_kernel(int n, const double* __restrict__ V, int k1,
                                                           double* __restrict__ w, double2* __restrict__ part,
                                                           const double* __restrict__ st, int m) {
  if (st[gEnd] != 0.0) return;
  const double* hcol = gm_hcol(const_cast<double*>(st), m);
  int lo, hi;
  gm_rows(n, lo, hi);
  double s = 0.0, c = 0.0;
  for (int i = lo + int(threadIdx.x); i < hi; i += blockDim.x) {
    double wi = w[i];
    for (int j = 0; j < k1; ++j) wi -= hcol[j] * V[size_t(j) * n + i];
    w[i] = wi;
    acc_add(s, c, wi * wi);
  }
  block_reduce_pair(s, c);
  if (threadIdx.x == 0) part[blockIdx.x] = make_double2(s, c);
}

// append column k: h_{k+1,k} = ||w||, previous rotations, a new rotation, g, the residual
// estimate |g_{k+1}| and the stopping tests
__global__ __launch_bounds__(1024) void gm_step_finalize(GmParams P, const double2* __restrict__ part, int nb, int k,
                                                         double* __restrict__ st) {
  if (st[gEnd] != 0.0) return;
  double s, c;
  gm_merge(part, nb, s, c);
  if (threadIdx.x != 0) return;
  const int m = P.m;
  double* H = gm_H(st) + size_t(m + 1) * k;  // column k
  double* g = gm_g(st, m);
  double* cs = gm_cs(st, m);
  double* sn = gm_sn(st, m);
  const double* hcol = gm_hcol(st, m);
  const double hn = sqrt(fmax(s + c, 0.0));
  for (int j = 0; j <= k; ++j) H[j] = hcol[j];
  H[k + 1] = hn;
  for (int j = 0; j < k; ++j) {
    const double t0 = cs[j] * H[j] + sn[j] * H[j + 1];
    H[j + 1] = -sn[j] * H[j] + cs[j] * H[j + 1];
    H[j] = t0;
  }
  const double den = hypot(H[k], H[k + 1]);
  if (den == 0.0) {
    cs[k] = 1.0;
    sn[k] = 0.0;
  } else {
    cs[k] = H[k] / den;
    sn[k] = H[k + 1] / den;
  }
  H[k] = cs[k] * H[k] + sn[k] * H[k + 1];
  H[k + 1] = 0.0;
  g[k + 1] = -sn[k] * g[k];
  g[k] = cs[k] * g[k];
  const int its = int(st[gIts]) + 1;
  st[gIts] = its;
  const double rn = fabs(g[k + 1]);
  st[gRnorm] = rn;
  st[gKUsed] = k + 1;
  const int reason = gm_converged(its, rn, st[gRnorm0], P);
  st[gReason] = reason;
  if (reason) st[gDone] = 1.0;
  if (reason || its >= P.max_it || hn == 0.0) st[gEnd] = 1.0;
  st[gInv] = hn > 0.0 ? 1.0 / hn : 0.0;
}

// H y = g (upper triangular, k_used columns) by back substitution
__global__ __launch_bounds__(64) void gm_solve_kernel(GmParams P, double* __restrict__ st) {
  if (threadIdx.x != 0) return;
  const int m = P.m;
  const int ku = int(st[gKUsed]);
  const double* H = gm_H(st);
  const double* g = gm_g(st, m);
  double* y = gm_y(st, m);
  for (int i = ku - 1; i >= 0; --i) {
    double s = g[i];
    for (int j = i + 1; j < ku; ++j) s -= H[i + size_t(m + 1) * j] * y[j];
    const double d = H[i + size_t(m + 1) * i];
    y[i] = d != 0.0 ? s / d : 0.0;
  }
}

// x += sum_j y[j] basis[j] (j ascending, j < k_used)
__global__ __launch_bounds__(kBlock) void gm_xupdate_kernel(int n, const double* __restrict__ basis,
                                                            double* __restrict__ x, const double* __restrict__ st,
                                                            int m) {
  const int ku = int(st[gKUsed]);
  if (ku == 0) return;
  const double* y = gm_y(const_cast<double*>(st), m);
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double xi = x[i];
    for (int j = 0; j < ku; ++j) xi += y[j] * basis[size_t(j) * n + i];
    x[i] = xi;
  }
}

__global__ void gm_init_state(double* __restrict__ st, int64_t len, double* __restrict__ x, int n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < len) st[i] = 0.0;
  if (i == gFirst) st[i] = 1.0;
  if (i < n) x[i] = 0.0;
}

// the CSR pass of one product in the kernels' form
struct GmShape {
  int form;
  int lpr;
  int blocks;
};

GmShape gm_shape(int n, int64_t nnz, const void* indices, const void* vals) {
  const double avg = n > 0 ? double(nnz) / n : 1.0;
  GmShape sh;
  if (csr_stream_ok(indices, vals)) {
    sh.form = 2;
    sh.lpr = csr_stream_lanes(avg);
  } else {
    sh.lpr = csr_lanes_per_row(avg);
    sh.form = csr_use_vec(indices, vals, sh.lpr) ? 1 : 0;
  }
  sh.blocks = int((int64_t(n) * sh.lpr + 255) / 256);
  return sh;
}

template <typename VT, int LPR, bool RES>
void gm_row_launch(int form, int nb, hipStream_t s, int n, const int32_t* indptr, const int32_t* indices,
                   const VT* vals, const double* in, const double* b, double* out, const double* st, int slot) {
  if (RES) {
    if (form == 2)
      hipLaunchKernelGGL((gm_residual_kernel<VT, LPR, 2>), dim3(nb), dim3(256), 0, s, n, indptr, indices, vals, in, b,
                         out, st, slot);
    else if (form == 1)
      hipLaunchKernelGGL((gm_residual_kernel<VT, LPR, 1>), dim3(nb), dim3(256), 0, s, n, indptr, indices, vals, in, b,
                         out, st, slot);
    else
      hipLaunchKernelGGL((gm_residual_kernel<VT, LPR, 0>), dim3(nb), dim3(256), 0, s, n, indptr, indices, vals, in, b,
                         out, st, slot);
  } else {
    if (form == 2)
      hipLaunchKernelGGL((gm_spmv_kernel<VT, LPR, 2>), dim3(nb), dim3(256), 0, s, n, indptr, indices, vals, in, out,
                         st, slot);
    else if (form == 1)
      hipLaunchKernelGGL((gm_spmv_kernel<VT, LPR, 1>), dim3(nb), dim3(256), 0, s, n, indptr, indices, vals, in, out,
                         st, slot);
    else
      hipLaunchKernelGGL((gm_spmv_kernel<VT, LPR, 0>), dim3(nb), dim3(256), 0, s, n, indptr, indices, vals, in, out,
                         st, slot);
  }
}

// out = A in (RES false) or out = b - A in (RES true)
template <typename VT, bool RES>
void gm_row_pass(const GmShape& sh, hipStream_t s, int n, const int32_t* indptr, const int32_t* indices,
                 const VT* vals, const double* in, const double* b, double* out, const double* st, int slot) {
  const int f = sh.form, nb = sh.blocks;
  switch (sh.lpr) {
    case 1: gm_row_launch<VT, 1, RES>(f, nb, s, n, indptr, indices, vals, in, b, out, st, slot); break;
    case 2: gm_row_launch<VT, 2, RES>(f, nb, s, n, indptr, indices, vals, in, b, out, st, slot); break;
    case 4: gm_row_launch<VT, 4, RES>(f, nb, s, n, indptr, indices, vals, in, b, out, st, slot); break;
    case 8: gm_row_launch<VT, 8, RES>(f, nb, s, n, indptr, indices, vals, in, b, out, st, slot); break;
    case 16: gm_row_launch<VT, 16, RES>(f, nb, s, n, indptr, indices, vals, in, b, out, st, slot); break;
    case 32: gm_row_launch<VT, 32, RES>(f, nb, s, n, indptr, indices, vals, in, b, out, st, slot); break;
    default: gm_row_launch<VT, 64, RES>(f, nb, s, n, indptr, indices, vals, in, b, out, st, slot); break;
  }
}

int gm_blocks(int n) { return std::max(1, std::min(kGmBlocks, (n + kBlock - 1) / kBlock)); }
int gm_vec_blocks(int n) { return std::max(1, std::min(2048, (n + kBlock - 1) / kBlock)); }

}  // namespace

extern "C" {

int64_t fne_gmres_workspace_bytes(int32_t n, int32_t restart, int32_t flexible) {
  if (n < 0 || restart < 1 || restart > kGmMaxRestart) return FNE_EINVAL;
  const int64_t m = restart;
  const int64_t vecs = (m + 1) + (flexible ? m : 0) + 2;  // V, Z, t, w
  const int64_t nparts = int64_t(gm_blocks(n)) * (m + 1);
  return int64_t(sizeof(double)) * (vecs * int64_t(n) + gm_state_len(int(m))) + int64_t(sizeof(double2)) * nparts;
}

int fne_gmres(int32_t n, const int32_t* indptr, const int32_t* indices, const void* vals, int32_t precision,
              const double* inv_diag, void* pc_graph_exec, double* pc_in, const double* pc_out, const double* b,
              double* x, double rtol, double atol, double dtol, int32_t max_it, int32_t restart, int32_t flexible,
              int32_t check_every, void* work, int64_t work_bytes, int32_t* its_out, int32_t* reason_out,
              double* rnorm_out, fne_stream_t stream) {
  const char* name = flexible ? "fne_gmres (flexible)" : "fne_gmres";
  if (n < 0 || !indptr || !indices || !vals || !b || !x || !work || max_it < 0 || restart < 1 ||
      restart > kGmMaxRestart || (pc_graph_exec && (!pc_in || !pc_out || pc_in == pc_out))) {
    fne_set_error(std::string(name) + ": invalid arguments (restart 1.." + std::to_string(kGmMaxRestart) + ")");
    return FNE_EINVAL;
  }
  const int prec = precision & 0xFF;
  if (prec != FNE_FP32 && prec != FNE_FP64) {
    fne_set_error(std::string(name) + ": precision must be FNE_FP64 or FNE_FP32");
    return FNE_EINVAL;
  }
  if (work_bytes < fne_gmres_workspace_bytes(n, restart, flexible)) {
    fne_set_error(std::string(name) + ": workspace too small");
    return FNE_EINVAL;
  }
  if (n == 0) {  // nothing to solve: converged at iteration 0
    if (its_out) *its_out = 0;
    if (reason_out) *reason_out = kConvAtol;
    if (rnorm_out) *rnorm_out = 0.0;
    return FNE_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  int64_t nnz = 0;
  {
    int32_t last = 0;
    FNE_HIP_CHECK(hipMemcpyAsync(&last, indptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FNE_HIP_CHECK(hipStreamSynchronize(s));
    nnz = last;
  }
  const int m = restart;
  double* V = reinterpret_cast<double*>(work);
  double* Z = V + size_t(m + 1) * n;
  double* t = Z + (flexible ? size_t(m) * n : 0);
  double* w = t + n;
  double* st = w + n;
  const int64_t slen = gm_state_len(m);
  double2* part = reinterpret_cast<double2*>(st + slen);
  const GmShape sh = gm_shape(n, nnz, indices, vals);
  const int nb = gm_blocks(n);
  const int nbv = gm_vec_blocks(n);
  GmParams P{n, m, max_it, rtol, atol, dtol};
  hipGraphExec_t pc = (hipGraphExec_t)pc_graph_exec;

  auto product = [&](const double* in, double* out, int slot) {
    if (prec == FNE_FP32)
      gm_row_pass<float, false>(sh, s, n, indptr, indices, (const float*)vals, in, nullptr, out, st, slot);
    else
      gm_row_pass<double, false>(sh, s, n, indptr, indices, (const double*)vals, in, nullptr, out, st, slot);
  };
  auto residual = [&](double* out) {
    if (prec == FNE_FP32)
      gm_row_pass<float, true>(sh, s, n, indptr, indices, (const float*)vals, x, b, out, st, gDone);
    else
      gm_row_pass<double, true>(sh, s, n, indptr, indices, (const double*)vals, x, b, out, st, gDone);
  };
  // out = B in: Jacobi / identity, or the recorded graph between two guarded copies
  auto precondition = [&](const double* in, double* out, int slot) -> int {
    if (!pc) {
      hipLaunchKernelGGL(gm_scale_diag_kernel, dim3(nbv), dim3(kBlock), 0, s, n, inv_diag, in, out, st, slot);
      return FNE_OK;
    }
    hipLaunchKernelGGL(gm_scale_diag_kernel, dim3(nbv), dim3(kBlock), 0, s, n, (const double*)nullptr, in, pc_in, st,
                       slot);
    FNE_HIP_CHECK(hipGraphLaunch(pc, s));
    hipLaunchKernelGGL(gm_scale_diag_kernel, dim3(nbv), dim3(kBlock), 0, s, n, (const double*)nullptr, pc_out, out, st,
                       slot);
    return FNE_OK;
  };

  {
    const int64_t cover = std::max<int64_t>(slen, n);
    hipLaunchKernelGGL(gm_init_state, dim3(int((cover + 255) / 256)), dim3(256), 0, s, st, slen, x, n);
    FNE_LAUNCH_CHECK();
  }
  double hst[gHdr];
  bool first = true;
  int rc = FNE_OK;
  while (true) {
    // cycle start: v0 = B b (GMRES) or b (FGMRES) on the first cycle (x = 0), else from b - A x
    double* r0 = flexible ? V : t;
    if (first)
      hipLaunchKernelGGL(gm_scale_diag_kernel, dim3(nbv), dim3(kBlock), 0, s, n, (const double*)nullptr, b, r0, st,
                         int(gDone));
    else
      residual(r0);
    if (!flexible && (rc = precondition(t, V, gDone)) != FNE_OK) return rc;
    hipLaunchKernelGGL(gm_dots_kernel, dim3(nb), dim3(kBlock), 0, s, n, V, 1, V, part, st, int(gDone));
    hipLaunchKernelGGL(gm_start_finalize, dim3(1), dim3(1024), 0, s, P, part, nb, st);
    hipLaunchKernelGGL(gm_normalize_kernel, dim3(nbv), dim3(kBlock), 0, s, n, V, V, st);
    FNE_LAUNCH_CHECK();
    first = false;
    for (int k = 0; k < m; ++k) {
      if (k > 0 && check_every > 0 && k % check_every == 0) {
        FNE_HIP_CHECK(hipMemcpyAsync(hst, st, sizeof(hst), hipMemcpyDeviceToHost, s));
        FNE_HIP_CHECK(hipStreamSynchronize(s));
        if (hst[gEnd] != 0.0) break;
      }
      double* vk = V + size_t(k) * n;
      if (flexible) {
        double* zk = Z + size_t(k) * n;
        if ((rc = precondition(vk, zk, gEnd)) != FNE_OK) return rc;
        product(zk, w, gEnd);
      } else {
        product(vk, t, gEnd);
        if ((rc = precondition(t, w, gEnd)) != FNE_OK) return rc;
      }
      hipLaunchKernelGGL(gm_dots_kernel, dim3(nb), dim3(kBlock), 0, s, n, V, k + 1, w, part, st, int(gEnd));
      hipLaunchKernelGGL(gm_hcol_finalize, dim3(1), dim3(1024), 0, s, P, part, nb, k + 1, st);
      hipLaunchKernelGGL(gm_update_kernel, dim3(nb), dim3(kBlock), 0, s, n, V, k + 1, w, part, st, m);
      hipLaunchKernelGGL(gm_step_finalize, dim3(1), dim3(1024), 0, s, P, part, nb, k, st);
      hipLaunchKernelGGL(gm_normalize_kernel, dim3(nbv), dim3(kBlock), 0, s, n, w, V + size_t(k + 1) * n, st);
      FNE_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(gm_solve_kernel, dim3(1), dim3(64), 0, s, P, st);
    hipLaunchKernelGGL(gm_xupdate_kernel, dim3(nbv), dim3(kBlock), 0, s, n, flexible ? Z : V, x, st, m);
    FNE_LAUNCH_CHECK();
    FNE_HIP_CHECK(hipMemcpyAsync(hst, st, sizeof(hst), hipMemcpyDeviceToHost, s));
    FNE_HIP_CHECK(hipStreamSynchronize(s));
    if (hst[gDone] != 0.0 || int(hst[gIts]) >= max_it) break;
  }
  int reason = int(hst[gReason]);
  if (hst[gDone] == 0.0) reason = -3;  // KSP_DIVERGED_ITS
  if (its_out) *its_out = int(hst[gIts]);
  if (reason_out) *reason_out = reason;
  if (rnorm_out) *rnorm_out = hst[gRnorm];
  return FNE_OK;
}

}  // extern "C"
