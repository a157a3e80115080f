// Device-resident preconditioned CG (KSPCG semantics) around the CSR SpMV.
//
// PETSc's KSPCG as the reference drives it (scalar_problem_driver.py:154-330): zero initial
// guess, point-Jacobi left preconditioning (zero diagonal entries -> 1), convergence on the
// preconditioned residual norm ||z_k|| <= max(rtol ||z_0||, atol), divergence at
// dtol ||z_0||, indefinite-matrix / indefinite-PC breakdown checks, iteration cap.  Every
// scalar (alpha, beta, norms, the stop flag) lives in device memory: an iteration is four
// launches (fused p-update + SpMV + p.w partials; alpha; x/r/z update + z.z and z.r partials;
// beta and the convergence test) and the host only reads the stop flag every
// ``check_every`` iterations.  Iterations enqueued after the stop are no-ops, so the result
// and the iteration count are exactly those of the iteration that met the test.  All dot
// products are compensated FP64 sums in a fixed order: the solve is bitwise reproducible.
//
// Trust mode (fne_pstcg, fne_pstcg_graph): PETSc's KSPSTCG, the Steihaug-Toint CG the
// trust-region Newton solves its subproblems with (core/petsc/trust_ksp.py:55-72,
// core/petsc/scalar_problem_driver.py:143-170; restated for the serial path in
// core/serial/sparse_solvers.py:20-93).  The same four launches carry two more dot products
// (p.p and x.p, fused into the SpMV pass) and ||x||^2 by recurrence; a step that would leave
// the ball ||x|| <= radius, or curvature p.Ap <= 0, moves x along p to the sphere and ends
// the solve (reasons 6 and 5); the stopping test uses the unpreconditioned residual ||r||.
#include "fne_common.h"

#include <algorithm>
#include <cmath>
#include <mutex>

namespace {

// state slots (doubles)
enum : int {
  kBeta = 0,
  kBetaOld,
  kDpiOld,
  kRnorm0,
  kRnorm,
  kAlpha,
  kRatio,
  kDone,
  kIts,
  kReason,
  kXX,     // trust mode: ||x||^2
  kFinal,  // trust mode: the pending update is the last one (boundary step)
  kStateLen = 16
};

constexpr int kConvRtol = 2, kConvAtol = 3, kConvNegCurve = 5, kConvStepLength = 6, kDivDtol = -4, kDivIndefPc = -8,
              kDivNan = -9, kDivIndefMat = -10;

struct CgParams {
  int n;
  const int32_t* indptr;
  const int32_t* indices;
  const double* inv_diag;  // null -> identity
  double rtol, atol, dtol;
  int trust;      // Steihaug-Toint mode
  double radius;  // trust mode: radius of the ball (inf: none)
};

__device__ __forceinline__ int cg_converged(int it, double rn, double rn0, const CgParams& P) {
  if (!isfinite(rn)) return kDivNan;
  if (rn <= fmax(P.rtol * rn0, P.atol)) return rn > P.atol ? kConvRtol : kConvAtol;
  if (it > 0 && rn >= P.dtol * rn0) return kDivDtol;
  return 0;
}

__device__ __forceinline__ void merge_partials(const double2* part, int nb, double& s, double& c) {
  s = 0.0;
  c = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) acc_merge(s, c, part[i].x, part[i].y);
  block_reduce_pair<1024>(s, c);
}

// r = b, z = D^-1 r, x = 0, p_old = 0; partials of z.z and z.r
__global__ __launch_bounds__(kBlock) void cg_init_kernel(CgParams P, const double* __restrict__ b,
                                                         double* __restrict__ x, double* __restrict__ r,
                                                         double* __restrict__ z, double* __restrict__ p0,
                                                         double2* __restrict__ part) {
  double s1 = 0.0, c1 = 0.0, s2 = 0.0, c2 = 0.0;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += stride) {
    const double ri = b[i];
    const double zi = P.inv_diag ? ri * P.inv_diag[i] : ri;
    r[i] = ri;
    z[i] = zi;
    x[i] = 0.0;
    p0[i] = 0.0;
    acc_add(s1, c1, P.trust ? ri * ri : zi * zi);
    acc_add(s2, c2, zi * ri);
  }
  block_reduce_pair(s1, c1);
  __syncthreads();
  block_reduce_pair(s2, c2);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = make_double2(s1, c1);
    part[gridDim.x + blockIdx.x] = make_double2(s2, c2);
  }
}

__global__ __launch_bounds__(1024) void cg_init_finalize(CgParams P, const double2* __restrict__ part, int nb,
                                                         double* __restrict__ st) {
  double zz, czz, zr, czr;
  merge_partials(part, nb, zz, czz);
  __syncthreads();
  merge_partials(part + nb, nb, zr, czr);
  if (threadIdx.x != 0) return;
  const double dp = sqrt(fmax(zz + czz, 0.0));
  st[kRnorm0] = dp;
  st[kRnorm] = dp;
  st[kIts] = 0.0;
  st[kDpiOld] = NAN;
  st[kRatio] = 0.0;
  st[kDone] = 0.0;
  st[kReason] = 0.0;
  st[kXX] = 0.0;
  st[kFinal] = 0.0;
  int reason = cg_converged(0, dp, dp, P);
  const double beta = zr + czr;
  if (!reason && beta < 0.0) reason = kDivIndefPc;
  st[kBeta] = beta;
  st[kBetaOld] = beta;
  if (reason) {
    st[kReason] = reason;
    st[kDone] = 1.0;
  }
}

// p_new = z + ratio p_old; w = A p_new (on the fly from z and p_old); partials of p_new.w
// (trust mode also p_new.p_new and x.p_new, at part + gridDim.x and part + 2 gridDim.x).
// FORM: 0 lanes per row, 1 lanes per row with grouped loads, 2 CSR-stream (256 / LPR rows per
// workgroup)
template <typename VT, int LPR, int FORM>
__global__ __launch_bounds__(kBlock) void cg_spmv_kernel(CgParams P, const VT* __restrict__ vals,
                                                         const double* __restrict__ z,
                                                         const double* __restrict__ p_old,
                                                         double* __restrict__ p_new, double* __restrict__ w,
                                                         const double* __restrict__ x,
                                                         const double* __restrict__ st,
                                                         double2* __restrict__ part) {
  if (st[kDone] != 0.0) return;  // uniform over the grid
  const double ratio = st[kRatio];
  __shared__ double prod[FORM == 2 ? kStreamCap : 1];
  int row;
  bool leader;
  const double acc = csr_form_dot<VT, LPR, FORM>(
      P.n, P.indptr, P.indices, vals, [&](int c) { return z[c] + ratio * p_old[c]; }, prod, row, leader);
  double s = 0.0, cc = 0.0, s2 = 0.0, c2 = 0.0, s3 = 0.0, c3 = 0.0;
  if (leader) {
    const double pn = z[row] + ratio * p_old[row];
    p_new[row] = pn;
    w[row] = acc;
    acc_add(s, cc, pn * acc);
    if (P.trust) {
      acc_add(s2, c2, pn * pn);
      acc_add(s3, c3, x[row] * pn);
    }
  }
  block_reduce_pair(s, cc);
  if (P.trust) {  // uniform over the grid
    __syncthreads();
    block_reduce_pair(s2, c2);
    __syncthreads();
    block_reduce_pair(s3, c3);
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = make_double2(s, cc);
    if (P.trust) {
      part[gridDim.x + blockIdx.x] = make_double2(s2, c2);
      part[2 * gridDim.x + blockIdx.x] = make_double2(s3, c3);
    }
  }
}

__global__ __launch_bounds__(1024) void cg_alpha_finalize(CgParams P, const double2* __restrict__ part, int nb,
                                                          double* __restrict__ st) {
  if (st[kDone] != 0.0) return;
  double s, c;
  merge_partials(part, nb, s, c);
  if (P.trust) {
    double pp, cpp, xp, cxp;
    __syncthreads();
    merge_partials(part + nb, nb, pp, cpp);
    __syncthreads();
    merge_partials(part + 2 * nb, nb, xp, cxp);
    if (threadIdx.x != 0) return;
    const double kappa = s + c;
    pp += cpp;
    xp += cxp;
    if (!isfinite(kappa)) {
      st[kIts] += 1.0;
      st[kReason] = kDivNan;
      st[kDone] = 1.0;
      return;
    }
    const double delta = P.radius, xx = st[kXX];
    // positive root tau of ||x + tau p||^2 = delta^2
    const double tau = (-xp + sqrt(fmax(xp * xp + pp * (delta * delta - xx), 0.0))) / pp;
    if (kappa <= 0.0) {
      st[kReason] = kConvNegCurve;
      if (isfinite(delta)) {
        st[kAlpha] = tau;
        st[kFinal] = 1.0;
      } else {  // no boundary to move to: keep the iterate
        st[kIts] += 1.0;
        st[kDone] = 1.0;
      }
      return;
    }
    const double alpha = st[kBeta] / kappa;
    const double xx_new = xx + 2.0 * alpha * xp + alpha * alpha * pp;
    if (xx_new >= delta * delta) {
      st[kReason] = kConvStepLength;
      st[kAlpha] = tau;
      st[kFinal] = 1.0;
      return;
    }
    st[kXX] = xx_new;
    st[kAlpha] = alpha;
    return;
  }
  if (threadIdx.x != 0) return;
  const double dpi = s + c;
  const double old = st[kDpiOld];
  if (dpi == 0.0 || (isfinite(old) && dpi * old <= 0.0) || !isfinite(dpi)) {
    st[kReason] = isfinite(dpi) ? kDivIndefMat : kDivNan;
    st[kDone] = 1.0;
    return;
  }
  st[kDpiOld] = dpi;
  st[kAlpha] = st[kBeta] / dpi;
}

// x += alpha p, r -= alpha w, z = D^-1 r; partials of z.z and z.r
__global__ __launch_bounds__(kBlock) void cg_update_kernel(CgParams P, const double* __restrict__ p,
                                                           const double* __restrict__ w, double* __restrict__ x,
                                                           double* __restrict__ r, double* __restrict__ z,
                                                           const double* __restrict__ st,
                                                           double2* __restrict__ part) {
  if (st[kDone] != 0.0) return;
  const double a = st[kAlpha];
  const int stride = gridDim.x * blockDim.x;
  if (st[kFinal] != 0.0) {  // trust mode: the boundary step ends the solve
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += stride) x[i] += a * p[i];
    return;
  }
  double s1 = 0.0, c1 = 0.0, s2 = 0.0, c2 = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += stride) {
    x[i] += a * p[i];
    const double ri = r[i] - a * w[i];
    r[i] = ri;
    const double zi = P.inv_diag ? ri * P.inv_diag[i] : ri;
    z[i] = zi;
    acc_add(s1, c1, P.trust ? ri * ri : zi * zi);
    acc_add(s2, c2, zi * ri);
  }
  block_reduce_pair(s1, c1);
  __syncthreads();
  block_reduce_pair(s2, c2);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = make_double2(s1, c1);
    part[gridDim.x + blockIdx.x] = make_double2(s2, c2);
  }
}

__global__ __launch_bounds__(1024) void cg_beta_finalize(CgParams P, const double2* __restrict__ part, int nb,
                                                         double* __restrict__ st) {
  if (st[kDone] != 0.0) return;
  if (st[kFinal] != 0.0) {  // trust mode: boundary step taken, reason already set
    if (threadIdx.x == 0) {
      st[kIts] += 1.0;
      st[kDone] = 1.0;
    }
    return;
  }
  double zz, czz, zr, czr;
  merge_partials(part, nb, zz, czz);
  __syncthreads();
  merge_partials(part + nb, nb, zr, czr);
  if (threadIdx.x != 0) return;
  const double dp = sqrt(fmax(zz + czz, 0.0));  // ||z||, or ||r|| in trust mode
  const int its = int(st[kIts]) + 1;
  st[kIts] = its;
  st[kRnorm] = dp;
  int reason = cg_converged(its, dp, st[kRnorm0], P);
  if (reason) {
    st[kReason] = reason;
    st[kDone] = 1.0;
    return;
  }
  const double beta_old = st[kBeta];
  const double beta = zr + czr;
  st[kBetaOld] = beta_old;
  st[kBeta] = beta;
  if (beta < 0.0) {
    st[kReason] = kDivIndefPc;
    st[kDone] = 1.0;
    return;
  }
  st[kRatio] = beta / beta_old;
}

// graph-preconditioned variant: x += alpha p, r -= alpha w (z comes from the preconditioner)
__global__ __launch_bounds__(kBlock) void cg_update_xr_kernel(int n, const double* __restrict__ p,
                                                              const double* __restrict__ w, double* __restrict__ x,
                                                              double* __restrict__ r, const double* __restrict__ st) {
  if (st[kDone] != 0.0) return;
  const double a = st[kAlpha];
  const bool last = st[kFinal] != 0.0;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    x[i] += a * p[i];
    if (!last) r[i] -= a * w[i];
  }
}

// partials of z.z (r.r in trust mode) and z.r (z = M r from the preconditioner)
__global__ __launch_bounds__(kBlock) void cg_dots_kernel(int n, int trust, const double* __restrict__ z,
                                                         const double* __restrict__ r, const double* __restrict__ st,
                                                         double2* __restrict__ part) {
  if (st != nullptr && (st[kDone] != 0.0 || st[kFinal] != 0.0)) return;
  double s1 = 0.0, c1 = 0.0, s2 = 0.0, c2 = 0.0;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double zi = z[i], ri = r[i];
    acc_add(s1, c1, trust ? ri * ri : zi * zi);
    acc_add(s2, c2, zi * ri);
  }
  block_reduce_pair(s1, c1);
  __syncthreads();
  block_reduce_pair(s2, c2);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = make_double2(s1, c1);
    part[gridDim.x + blockIdx.x] = make_double2(s2, c2);
  }
}

constexpr int kVecBlocks = 1024;

// the SpMV pass of one CG iteration: CSR-stream when the arrays are 16-byte aligned (lanes per
// row from csr_stream_lanes), else lanes per row (csr_lanes_per_row)
struct CgSpmvShape {
  int form;  // 0, 1, 2 as cg_spmv_kernel's FORM
  int lpr;
  int blocks;
};

CgSpmvShape cg_spmv_shape(int n, int64_t nnz, const void* indices, const void* vals) {
  const double avg = n > 0 ? double(nnz) / n : 1.0;
  CgSpmvShape sh;
  if (csr_stream_ok(indices, vals)) {
    sh.form = 2;
    sh.lpr = csr_stream_lanes(avg);
  } else {
    sh.lpr = csr_lanes_per_row(avg);
    sh.form = csr_use_vec(indices, vals, sh.lpr) ? 1 : 0;
  }
  sh.blocks = int((int64_t(n) * sh.lpr + kBlock - 1) / kBlock);
  return sh;
}

template <typename VT, int LPR>
void launch_cg_spmv(int form, const CgParams& P, const VT* vals, const double* z, const double* p_old, double* p_new,
                    double* w, const double* x, const double* st, double2* part, int nbs, hipStream_t s) {
  if (form == 2)
    hipLaunchKernelGGL((cg_spmv_kernel<VT, LPR, 2>), dim3(nbs), dim3(kBlock), 0, s, P, vals, z, p_old, p_new, w, x,
                       st, part);
  else if (form == 1)
    hipLaunchKernelGGL((cg_spmv_kernel<VT, LPR, 1>), dim3(nbs), dim3(kBlock), 0, s, P, vals, z, p_old, p_new, w, x,
                       st, part);
  else
    hipLaunchKernelGGL((cg_spmv_kernel<VT, LPR, 0>), dim3(nbs), dim3(kBlock), 0, s, P, vals, z, p_old, p_new, w, x,
                       st, part);
}

template <typename VT>
void cg_spmv(const CgSpmvShape& sh, const CgParams& P, const VT* vals, const double* z, const double* p_old,
             double* p_new, double* w, const double* x, const double* st, double2* part, hipStream_t s) {
  const int f = sh.form, nbs = sh.blocks;
  switch (sh.lpr) {
    case 1: launch_cg_spmv<VT, 1>(f, P, vals, z, p_old, p_new, w, x, st, part, nbs, s); break;
    case 2: launch_cg_spmv<VT, 2>(f, P, vals, z, p_old, p_new, w, x, st, part, nbs, s); break;
    case 4: launch_cg_spmv<VT, 4>(f, P, vals, z, p_old, p_new, w, x, st, part, nbs, s); break;
    case 8: launch_cg_spmv<VT, 8>(f, P, vals, z, p_old, p_new, w, x, st, part, nbs, s); break;
    case 16: launch_cg_spmv<VT, 16>(f, P, vals, z, p_old, p_new, w, x, st, part, nbs, s); break;
    case 32: launch_cg_spmv<VT, 32>(f, P, vals, z, p_old, p_new, w, x, st, part, nbs, s); break;
    default: launch_cg_spmv<VT, 64>(f, P, vals, z, p_old, p_new, w, x, st, part, nbs, s); break;
  }
}

// workgroups of the SpMV pass, the larger of the two forms (the workspace is sized before the
// arrays' alignment is known)
int64_t spmv_blocks_max(int n, int64_t nnz) {
  const double avg = n > 0 ? double(nnz) / n : 1.0;
  const int64_t l = std::max(csr_lanes_per_row(avg), csr_stream_lanes(avg));
  return (int64_t(n) * l + kBlock - 1) / kBlock;
}

}  // namespace

extern "C" {

int64_t fne_pcg_workspace_bytes(int32_t n, int64_t nnz) {
  if (n < 0 || nnz < 0) return FNE_EINVAL;
  int64_t nparts = 3 * spmv_blocks_max(n, nnz);  // p.w, and in trust mode p.p and x.p
  if (nparts < 2 * kVecBlocks) nparts = 2 * kVecBlocks;
  return int64_t(sizeof(double)) * (5 * int64_t(n) + kStateLen) + int64_t(sizeof(double2)) * nparts;
}

}  // extern "C"

namespace {

int pcg_run(const char* name, int trust, double radius, int32_t n, const int32_t* indptr, const int32_t* indices,
            const void* vals, int32_t precision, const double* inv_diag, const double* b, double* x, double rtol,
            double atol, double dtol, int32_t max_it, int32_t check_every, void* work, int64_t work_bytes,
            int32_t* its_out, int32_t* reason_out, double* rnorm_out, fne_stream_t stream) {
  if (n < 0 || !indptr || !indices || !vals || !b || !x || !work || max_it < 0 || (trust && !(radius >= 0.0))) {
    fne_set_error(std::string(name) + ": invalid arguments");
    return FNE_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  int64_t nnz64 = 0;
  {
    int32_t last = 0;
    if (n > 0 && hipMemcpyAsync(&last, indptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess) {
      fne_set_error(std::string(name) + ": reading indptr failed");
      return FNE_EHIP;
    }
    FNE_HIP_CHECK(hipStreamSynchronize(s));
    nnz64 = last;
  }
  if (work_bytes < fne_pcg_workspace_bytes(n, nnz64)) {
    fne_set_error(std::string(name) + ": workspace too small");
    return FNE_EINVAL;
  }
  if (check_every < 1) check_every = 1;
  const CgSpmvShape sh = cg_spmv_shape(n, nnz64, indices, vals);
  const int nbs = sh.blocks;
  double* r = reinterpret_cast<double*>(work);
  double* z = r + n;
  double* w = z + n;
  double* pa = w + n;
  double* pb = pa + n;
  double* st = pb + n;
  double2* part = reinterpret_cast<double2*>(st + kStateLen);
  CgParams P{n, indptr, indices, inv_diag, rtol, atol, dtol, trust, radius};
  const int nbv = n > 0 ? int(std::min<int64_t>(kVecBlocks, (int64_t(n) + kBlock - 1) / kBlock)) : 1;
  hipLaunchKernelGGL(cg_init_kernel, dim3(nbv), dim3(kBlock), 0, s, P, b, x, r, z, pa, part);
  hipLaunchKernelGGL(cg_init_finalize, dim3(1), dim3(1024), 0, s, P, part, nbv, st);
  FNE_LAUNCH_CHECK();
  double hst[kStateLen];
  int it = 0;
  bool done = false;
  while (!done) {
    FNE_HIP_CHECK(hipMemcpyAsync(hst, st, sizeof(hst), hipMemcpyDeviceToHost, s));
    FNE_HIP_CHECK(hipStreamSynchronize(s));
    if (hst[kDone] != 0.0 || it >= max_it) break;
    const int upto = std::min<int>(max_it, it + check_every);
    for (; it < upto; ++it) {
      double* p_old = (it & 1) ? pb : pa;
      double* p_new = (it & 1) ? pa : pb;
      if (precision == FNE_FP32)
        cg_spmv<float>(sh, P, (const float*)vals, z, p_old, p_new, w, x, st, part, s);
      else
        cg_spmv<double>(sh, P, (const double*)vals, z, p_old, p_new, w, x, st, part, s);
      hipLaunchKernelGGL(cg_alpha_finalize, dim3(1), dim3(1024), 0, s, P, part, nbs, st);
      hipLaunchKernelGGL(cg_update_kernel, dim3(nbv), dim3(kBlock), 0, s, P, p_new, w, x, r, z, st, part);
      hipLaunchKernelGGL(cg_beta_finalize, dim3(1), dim3(1024), 0, s, P, part, nbv, st);
    }
    FNE_LAUNCH_CHECK();
  }
  const int its = int(hst[kIts]);
  int reason = int(hst[kReason]);
  if (hst[kDone] == 0.0) reason = -3;  // KSP_DIVERGED_ITS
  if (its_out) *its_out = its;
  if (reason_out) *reason_out = reason;
  if (rnorm_out) *rnorm_out = hst[kRnorm];
  return FNE_OK;
}

int pcg_graph_run(const char* name, int trust, double radius, int32_t n, const int32_t* indptr,
                  const int32_t* indices, const void* vals, int32_t precision, void* pc_graph_exec, double* pc_in,
                  const double* pc_out, const double* b, double* x, double rtol, double atol, double dtol,
                  int32_t max_it, int32_t check_every, void* work, int64_t work_bytes, int32_t* its_out,
                  int32_t* reason_out, double* rnorm_out, fne_stream_t stream) {
  if (n < 0 || !indptr || !indices || !vals || !b || !x || !work || max_it < 0 || !pc_graph_exec || !pc_in ||
      !pc_out || pc_in == pc_out || (trust && !(radius >= 0.0))) {
    fne_set_error(std::string(name) + ": invalid arguments");
    return FNE_EINVAL;
  }
  if (precision != FNE_FP32 && precision != FNE_FP64) {
    fne_set_error(std::string(name) + ": precision must be FNE_FP64 or FNE_FP32");
    return FNE_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  hipGraphExec_t pc = (hipGraphExec_t)pc_graph_exec;
  int64_t nnz64 = 0;
  {
    int32_t last = 0;
    if (n > 0 && hipMemcpyAsync(&last, indptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess) {
      fne_set_error(std::string(name) + ": reading indptr failed");
      return FNE_EHIP;
    }
    FNE_HIP_CHECK(hipStreamSynchronize(s));
    nnz64 = last;
  }
  if (work_bytes < fne_pcg_workspace_bytes(n, nnz64)) {
    fne_set_error(std::string(name) + ": workspace too small");
    return FNE_EINVAL;
  }
  if (check_every < 1) check_every = 1;
  const CgSpmvShape sh = cg_spmv_shape(n, nnz64, indices, vals);
  const int nbs = sh.blocks;
  // r lives in the preconditioner's input buffer, z in its output buffer
  double* r = pc_in;
  const double* z = pc_out;
  double* scratch_z = reinterpret_cast<double*>(work);  // written by the shared init kernel, unused
  double* w = scratch_z + n;
  double* pa = w + n;
  double* pb = pa + n;
  double* st = pb + 2 * int64_t(n);
  double2* part = reinterpret_cast<double2*>(st + kStateLen);
  CgParams P{n, indptr, indices, nullptr, rtol, atol, dtol, trust, radius};
  const int nbv = n > 0 ? int(std::min<int64_t>(kVecBlocks, (int64_t(n) + kBlock - 1) / kBlock)) : 1;
  hipLaunchKernelGGL(cg_init_kernel, dim3(nbv), dim3(kBlock), 0, s, P, b, x, r, scratch_z, pa, part);
  FNE_HIP_CHECK(hipGraphLaunch(pc, s));
  hipLaunchKernelGGL(cg_dots_kernel, dim3(nbv), dim3(kBlock), 0, s, n, trust, z, r, nullptr, part);
  hipLaunchKernelGGL(cg_init_finalize, dim3(1), dim3(1024), 0, s, P, part, nbv, st);
  FNE_LAUNCH_CHECK();
  double hst[kStateLen];
  int it = 0;
  while (true) {
    FNE_HIP_CHECK(hipMemcpyAsync(hst, st, sizeof(hst), hipMemcpyDeviceToHost, s));
    FNE_HIP_CHECK(hipStreamSynchronize(s));
    if (hst[kDone] != 0.0 || it >= max_it) break;
    const int upto = std::min<int>(max_it, it + check_every);
    for (; it < upto; ++it) {
      double* p_old = (it & 1) ? pb : pa;
      double* p_new = (it & 1) ? pa : pb;
      if (precision == FNE_FP32)
        cg_spmv<float>(sh, P, (const float*)vals, z, p_old, p_new, w, x, st, part, s);
      else
        cg_spmv<double>(sh, P, (const double*)vals, z, p_old, p_new, w, x, st, part, s);
      hipLaunchKernelGGL(cg_alpha_finalize, dim3(1), dim3(1024), 0, s, P, part, nbs, st);
      hipLaunchKernelGGL(cg_update_xr_kernel, dim3(nbv), dim3(kBlock), 0, s, n, p_new, w, x, r, st);
      FNE_HIP_CHECK(hipGraphLaunch(pc, s));
      hipLaunchKernelGGL(cg_dots_kernel, dim3(nbv), dim3(kBlock), 0, s, n, trust, z, r, st, part);
      hipLaunchKernelGGL(cg_beta_finalize, dim3(1), dim3(1024), 0, s, P, part, nbv, st);
    }
    FNE_LAUNCH_CHECK();
  }
  const int its = int(hst[kIts]);
  int reason = int(hst[kReason]);
  if (hst[kDone] == 0.0) reason = -3;  // KSP_DIVERGED_ITS
  if (its_out) *its_out = its;
  if (reason_out) *reason_out = reason;
  if (rnorm_out) *rnorm_out = hst[kRnorm];
  return FNE_OK;
}

}  // namespace

extern "C" int fne_pcg(int32_t n, const int32_t* indptr, const int32_t* indices, const void* vals, int32_t precision,
                       const double* inv_diag, const double* b, double* x, double rtol, double atol, double dtol,
                       int32_t max_it, int32_t check_every, void* work, int64_t work_bytes, int32_t* its_out,
                       int32_t* reason_out, double* rnorm_out, fne_stream_t stream) {
  return pcg_run("fne_pcg", 0, INFINITY, n, indptr, indices, vals, precision, inv_diag, b, x, rtol, atol, dtol, max_it,
                 check_every, work, work_bytes, its_out, reason_out, rnorm_out, stream);
}

extern "C" int fne_pstcg(int32_t n, const int32_t* indptr, const int32_t* indices, const void* vals,
                         int32_t precision, const double* inv_diag, const double* b, double* x, double radius,
                         double rtol, double atol, double dtol, int32_t max_it, int32_t check_every, void* work,
                         int64_t work_bytes, int32_t* its_out, int32_t* reason_out, double* rnorm_out,
                         fne_stream_t stream) {
  return pcg_run("fne_pstcg", 1, radius, n, indptr, indices, vals, precision, inv_diag, b, x, rtol, atol, dtol,
                 max_it, check_every, work, work_bytes, its_out, reason_out, rnorm_out, stream);
}

extern "C" int fne_pcg_graph(int32_t n, const int32_t* indptr, const int32_t* indices, const void* vals,
                             int32_t precision, void* pc_graph_exec, double* pc_in, const double* pc_out,
                             const double* b, double* x, double rtol, double atol, double dtol, int32_t max_it,
                             int32_t check_every, void* work, int64_t work_bytes, int32_t* its_out,
                             int32_t* reason_out, double* rnorm_out, fne_stream_t stream) {
  return pcg_graph_run("fne_pcg_graph", 0, INFINITY, n, indptr, indices, vals, precision, pc_graph_exec, pc_in,
                       pc_out, b, x, rtol, atol, dtol, max_it, check_every, work, work_bytes, its_out, reason_out,
                       rnorm_out, stream);
}

extern "C" int fne_pstcg_graph(int32_t n, const int32_t* indptr, const int32_t* indices, const void* vals,
                               int32_t precision, void* pc_graph_exec, double* pc_in, const double* pc_out,
                               const double* b, double* x, double radius, double rtol, double atol, double dtol,
                               int32_t max_it, int32_t check_every, void* work, int64_t work_bytes, int32_t* its_out,
                               int32_t* reason_out, double* rnorm_out, fne_stream_t stream) {
  return pcg_graph_run("fne_pstcg_graph", 1, radius, n, indptr, indices, vals, precision, pc_graph_exec, pc_in,
                       pc_out, b, x, rtol, atol, dtol, max_it, check_every, work, work_bytes, its_out, reason_out,
                       rnorm_out, stream);
}

// ---------------------------------------------------------------------------------------------
// fne_pcg_noalloc: fne_pcg with its workspace in a file-scope device array instead of a caller
// buffer, for callers that cannot allocate device memory.  Each device holds its own copy of
// the array from module load to process exit; its address on the current device is looked up
// once per device with hipGetSymbolAddress and kept in g_scratch_ptr.  Calls on one device share
// the array, so calls that overlap in time on one device (two streams, two threads) race.  Every
// slot the solve reads is written by the solve first (r, z, x, p, the state and the partials are
// set by its init kernels), so nothing left by an earlier call is read.
// ---------------------------------------------------------------------------------------------
namespace {
constexpr int64_t kNoAllocScratchBytes = FNE_PCG_NOALLOC_MAX_BYTES;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    fne_set_error("fne_pcg_noalloc: no current HIP device");
    return FNE_EHIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fne_set_error(std::string("fne_pcg_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
      return FNE_EHIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return FNE_OK;
}
}  // namespace

extern "C" int fne_pcg_noalloc(int32_t n, const int32_t* indptr, const int32_t* indices, const void* vals,
                               int32_t precision, const double* inv_diag, const double* b, double* x, double rtol,
                               double atol, double dtol, int32_t max_it, int32_t check_every, int32_t* its_out,
                               int32_t* reason_out, double* rnorm_out, fne_stream_t stream) {
  if (n < 0 || (n > 0 && !indptr)) {
    fne_set_error("fne_pcg_noalloc: invalid arguments");
    return FNE_EINVAL;
  }
  int64_t nnz = 0;
  if (n > 0) {
    int32_t last = 0;
    hipStream_t s = (hipStream_t)stream;
    if (hipMemcpyAsync(&last, indptr + n, sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      fne_set_error("fne_pcg_noalloc: reading indptr failed");
      return FNE_EHIP;
    }
    nnz = last;
  }
  const int64_t need = fne_pcg_workspace_bytes(n, nnz);
  if (need < 0 || need > kNoAllocScratchBytes) {
    fne_set_error("fne_pcg_noalloc: the solve needs " + std::to_string(need) + " bytes of scratch (n = " +
                  std::to_string(n) + ", nnz = " + std::to_string(nnz) + "); the static scratch holds " +
                  std::to_string(kNoAllocScratchBytes) + " bytes: use fne_pcg with a caller workspace");
    return FNE_EINVAL;
  }
  void* work = nullptr;
  const int rc = noalloc_scratch(&work);
  if (rc != FNE_OK) return rc;
  return pcg_run("fne_pcg_noalloc", 0, INFINITY, n, indptr, indices, vals, precision, inv_diag, b, x, rtol, atol, dtol,
                 max_it, check_every, work, kNoAllocScratchBytes, its_out, reason_out, rnorm_out, stream);
}
