// Numeric sparse matrix product with a fixed result pattern (AMG Galerkin set-up).
//
// The smoothed-aggregation hierarchy is rebuilt on the device at every preconditioner
// set-up: the patterns of P = (I - w D^-1 A) P_tent, A P and P^T (A P) depend only on the
// Hessian pattern and the aggregates, so the host computes them once and every later set-up
// is a numeric product on the device (PETSc GAMG re-runs its Galerkin products at each
// PCSetUp the reference triggers per Newton iteration, core/petsc/scalar_problem_driver.py).
//
// C = A B restricted to C's pattern, which must hold every column a product of the row
// reaches (columns outside it are dropped; pattern entries no product reaches become 0).
// One 64-lane workgroup per row.  The row's result columns are staged in LDS in chunks of
// kChunk, and A's row with the bounds of the B rows it selects in chunks of kStage; for each
// nonzero a_ip of A's row, in order, the lanes take the entries of B's row p (distinct columns,
// so distinct accumulator slots), binary-search their slot in the staged columns and add
// a_ip * b_pj, while the first entries of the next B row are already on their way.  A barrier
// closes every p, so each result entry is the sum of its products in the order of A's row: the
// product is deterministic.
#include "fne_common.h"

namespace {

constexpr int kChunk = 512;   // result columns staged per pass
constexpr int kStage = 256;   // entries of A's row staged per pass
constexpr int kLanes = 64;

template <typename TA>
__global__ __launch_bounds__(kLanes) void spgemm_numeric_kernel(
    int n_rows, const int32_t* __restrict__ a_ptr, const int32_t* __restrict__ a_idx, const TA* __restrict__ a_val,
    const int32_t* __restrict__ b_ptr, const int32_t* __restrict__ b_idx, const double* __restrict__ b_val,
    const int32_t* __restrict__ c_ptr, const int32_t* __restrict__ c_idx, double* __restrict__ c_val) {
  __shared__ int32_t s_col[kChunk];
  __shared__ double s_acc[kChunk];
  __shared__ double s_a[kStage];    // A's row (values) ...
  __shared__ int32_t s_b0[kStage];  // ... and the bounds of the B rows its columns select
  __shared__ int32_t s_b1[kStage];
  const int row = xcd_remap(blockIdx.x, gridDim.x);
  if (row >= n_rows) return;  // the whole (single-wave) block leaves together
  const int lane = threadIdx.x;
  const int c0 = c_ptr[row], c1 = c_ptr[row + 1];
  const int a0 = a_ptr[row], a1 = a_ptr[row + 1];
  for (int cs = c0; cs < c1; cs += kChunk) {
    const int len = min(kChunk, c1 - cs);
    for (int t = lane; t < len; t += kLanes) {
      s_col[t] = c_idx[cs + t];
      s_acc[t] = 0.0;
    }
    __syncthreads();
    const int first = s_col[0], last = s_col[len - 1];
    auto add = [&](int j, double prod) {
      if (j < first || j > last) return;
      int lo = 0, hi = len - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s_col[mid] < j)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (s_col[lo] == j) s_acc[lo] += prod;
    };
    for (int ps = a0; ps < a1; ps += kStage) {
      const int na = min(kStage, a1 - ps);
      for (int t = lane; t < na; t += kLanes) {
        const int l = a_idx[ps + t];
        s_a[t] = static_cast<double>(a_val[ps + t]);
        s_b0[t] = b_ptr[l];
        s_b1[t] = b_ptr[l + 1];
      }
      __syncthreads();
      // the first 64 entries of the next B row are loaded while the current one is added
      int jn = -1;
      double bn = 0.0;
      {
        const int q = s_b0[0] + lane;
        if (q < s_b1[0]) {
          jn = b_idx[q];
          bn = b_val[q];
        }
      }
      for (int t = 0; t < na; ++t) {
        const int jc = jn;
        const double bc = bn;
        const double a = s_a[t];
        const int b0 = s_b0[t], b1 = s_b1[t];
        if (t + 1 < na) {
          const int q = s_b0[t + 1] + lane;
          jn = -1;
          if (q < s_b1[t + 1]) {
            jn = b_idx[q];
            bn = b_val[q];
          }
        }
        if (jc >= 0) add(jc, a * bc);
        for (int q = b0 + kLanes + lane; q < b1; q += kLanes) add(b_idx[q], a * b_val[q]);
        __syncthreads();  // entry t's additions land before entry t+1's (fixed order)
      }
    }
    for (int t = lane; t < len; t += kLanes) c_val[cs + t] = s_acc[t];
    __syncthreads();
  }
}

// Fused multigrid operators: one CSR pass per smoothing sweep, residual or prolongation.
//   kSweepFirst: y = w.r + w.(r - A (w.r))      (first damped-Jacobi sweep from x = 0)
//   kSweep:      y = x + w.(r - A x)
//   kResidual:   y = r - A x
//   kAddProduct: y = y_in + A x                 (prolongation added to the iterate; y may be y_in)
enum VcOp : int { kSweepFirst = 0, kSweep = 1, kResidual = 2, kAddProduct = 3 };

template <typename VT, int LPR, int OP, int FORM>
__global__ __launch_bounds__(256) void vcycle_op_kernel(int n, const int32_t* __restrict__ indptr,
                                                        const int32_t* __restrict__ indices,
                                                        const VT* __restrict__ vals, const double* __restrict__ x,
                                                        const double* __restrict__ r, const double* __restrict__ w,
                                                        const double* y_in, double* y) {
  __shared__ double prod[FORM == 2 ? kStreamCap : 1];
  int row;
  bool leader;
  const double acc = csr_form_dot<VT, LPR, FORM>(
      n, indptr, indices, vals, [&](int c) { return OP == kSweepFirst ? w[c] * r[c] : x[c]; }, prod, row, leader);
  if (leader) {
    double out;
    if (OP == kSweepFirst) {
      const double x0 = w[row] * r[row];
      out = x0 + w[row] * (r[row] - acc);
    } else if (OP == kSweep) {
      out = x[row] + w[row] * (r[row] - acc);
    } else if (OP == kResidual) {
      out = r[row] - acc;
    } else {
      out = y_in[row] + acc;
    }
    y[row] = out;
  }
}

// form of a row kernel's CSR pass: 2 CSR-stream when asked for (precision bits 16..19) and the
// arrays are aligned, else 1 / 0 lanes per row with / without grouped loads
int csr_kernel_form(int32_t precision, const void* indices, const void* vals, int lpr) {
  if (csr_form_of(precision) == kCsrFormStream && csr_stream_ok(indices, vals)) return 2;
  return csr_use_vec(indices, vals, lpr) ? 1 : 0;
}

template <typename VT, int OP>
void launch_vcycle_op(int n, const int32_t* indptr, const int32_t* indices, const VT* vals, const double* x,
                      const double* r, const double* w, const double* y_in, double* y, int lpr, int form,
                      hipStream_t s) {
  auto go = [&](auto kern, int L) {
    const long long threads = (long long)n * L;
    const int nb = int((threads + 255) / 256);
    hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, s, n, indptr, indices, vals, x, r, w, y_in, y);
  };
#define VC_L(L)                                                    \
  case L:                                                          \
    if (form == 2) go(vcycle_op_kernel<VT, L, OP, 2>, L);          \
    else if (form == 1) go(vcycle_op_kernel<VT, L, OP, 1>, L);     \
    else go(vcycle_op_kernel<VT, L, OP, 0>, L);                    \
    break;
  switch (lpr) {
    VC_L(1)
    VC_L(2)
    VC_L(4)
    VC_L(8)
    VC_L(16)
    VC_L(32)
    default:
      if (form == 2) go(vcycle_op_kernel<VT, 64, OP, 2>, 64);
      else if (form == 1) go(vcycle_op_kernel<VT, 64, OP, 1>, 64);
      else go(vcycle_op_kernel<VT, 64, OP, 0>, 64);
  }
#undef VC_L
}

template <typename VT>
int vcycle_op_dispatch(int op, int n, const int32_t* indptr, const int32_t* indices, const VT* vals,
                       const double* x, const double* r, const double* w, const double* y_in, double* y, int lpr,
                       int form, hipStream_t s) {
  switch (op) {
    case kSweepFirst:
      launch_vcycle_op<VT, kSweepFirst>(n, indptr, indices, vals, x, r, w, y_in, y, lpr, form, s);
      break;
    case kSweep: launch_vcycle_op<VT, kSweep>(n, indptr, indices, vals, x, r, w, y_in, y, lpr, form, s); break;
    case kResidual: launch_vcycle_op<VT, kResidual>(n, indptr, indices, vals, x, r, w, y_in, y, lpr, form, s); break;
    case kAddProduct:
      launch_vcycle_op<VT, kAddProduct>(n, indptr, indices, vals, x, r, w, y_in, y, lpr, form, s);
      break;
    default:
      fne_set_error("fne_vcycle_op: unknown op");
      return FNE_EINVAL;
  }
  return FNE_OK;
}

// Thin QR of every aggregate's rows of the near-nullspace (tentative prolongator).
// One wavefront per aggregate; a lane holds rows lane + 64 t (t < 4) of the m x K block in
// registers.  Modified Gram-Schmidt with one re-orthogonalisation pass, FP64, wave-wide sums by
// butterfly shuffles (a fixed order, so deterministic); R has a non-negative diagonal.  A
// column whose remaining norm falls below 1e-13 of its original norm becomes zero (R_jj = 0).
// When the aggregate has fewer rows than K (only the last one can), Q holds its first m
// columns and R's remaining columns are Q^T B.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int K>
__global__ __launch_bounds__(64) void aggregate_qr_kernel(int n, int agg, const double* __restrict__ B,
                                                          double* __restrict__ q_vals, double* __restrict__ bc) {
  const int a = blockIdx.x;
  const int lane = threadIdx.x;
  const int lo = a * agg;
  const int m = min(agg, n - lo);
  const int kk = min(m, K);
  double v[4][K];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int i = lane + 64 * t;
#pragma unroll
    for (int j = 0; j < K; ++j) v[t][j] = i < m ? B[int64_t(lo + i) * K + j] : 0.0;
  }
  double R[K][K];
#pragma unroll
  for (int i = 0; i < K; ++i)
#pragma unroll
    for (int j = 0; j < K; ++j) R[i][j] = 0.0;
#pragma unroll
  for (int j = 0; j < K; ++j) {
    if (j < kk) {
      double n0 = 0.0;
#pragma unroll
      for (int t = 0; t < 4; ++t) n0 += v[t][j] * v[t][j];
      n0 = sqrt(wave_sum(n0));
#pragma unroll
      for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
        for (int i = 0; i < K; ++i) {
          if (i < j) {
            double d = 0.0;
#pragma unroll
            for (int t = 0; t < 4; ++t) d += v[t][i] * v[t][j];
            d = wave_sum(d);
#pragma unroll
            for (int t = 0; t < 4; ++t) v[t][j] -= d * v[t][i];
            R[i][j] += d;
          }
        }
      }
      double nr = 0.0;
#pragma unroll
      for (int t = 0; t < 4; ++t) nr += v[t][j] * v[t][j];
      nr = sqrt(wave_sum(nr));
      const bool keep = nr > 1e-13 * n0 && nr > 0.0;
      R[j][j] = keep ? nr : 0.0;
      const double sc = keep ? 1.0 / nr : 0.0;
#pragma unroll
      for (int t = 0; t < 4; ++t) v[t][j] *= sc;
    } else {
#pragma unroll
      for (int i = 0; i < K; ++i) {
        if (i < kk) {
          double d = 0.0;
#pragma unroll
          for (int t = 0; t < 4; ++t) d += v[t][i] * v[t][j];
          R[i][j] = wave_sum(d);
        }
      }
    }
  }
  // Q in CSR order of the tentative prolongator (row-major m x kk, after a*agg*K values)
  double* q = q_vals + int64_t(lo) * K;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int i = lane + 64 * t;
    if (i < m) {
#pragma unroll
      for (int j = 0; j < K; ++j)
        if (j < kk) q[int64_t(i) * kk + j] = v[t][j];
    }
  }
  if (lane == 0) {
    double* r = bc + int64_t(a) * K * K;
#pragma unroll
    for (int i = 0; i < K; ++i)
      if (i < kk)
#pragma unroll
        for (int j = 0; j < K; ++j) r[i * K + j] = R[i][j];
  }
}

}  // namespace

extern "C" int fne_aggregate_qr(int32_t n, int32_t k, int32_t agg, const double* B, double* q_vals, double* bc,
                                fne_stream_t stream) {
  if (n < 0 || agg < k || agg > 256 || !(k == 1 || k == 2 || k == 3 || k == 6)) {
    fne_set_error("fne_aggregate_qr: need k in {1, 2, 3, 6} and k <= agg <= 256");
    return FNE_EINVAL;
  }
  if (n == 0) return FNE_OK;
  if (!B || !q_vals || !bc) {
    fne_set_error("fne_aggregate_qr: null pointer");
    return FNE_EINVAL;
  }
  const int na = (n + agg - 1) / agg;
  hipStream_t s = (hipStream_t)stream;
  switch (k) {
    case 1: hipLaunchKernelGGL(aggregate_qr_kernel<1>, dim3(na), dim3(64), 0, s, n, agg, B, q_vals, bc); break;
    case 2: hipLaunchKernelGGL(aggregate_qr_kernel<2>, dim3(na), dim3(64), 0, s, n, agg, B, q_vals, bc); break;
    case 3: hipLaunchKernelGGL(aggregate_qr_kernel<3>, dim3(na), dim3(64), 0, s, n, agg, B, q_vals, bc); break;
    default: hipLaunchKernelGGL(aggregate_qr_kernel<6>, dim3(na), dim3(64), 0, s, n, agg, B, q_vals, bc); break;
  }
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

extern "C" int fne_vcycle_op(int32_t op, int32_t n_rows, const int32_t* indptr, const int32_t* indices,
                             const void* vals, int32_t precision, const double* x, const double* r, const double* w,
                             const double* y_in, double* y, fne_stream_t stream) {
  if (n_rows < 0) {
    fne_set_error("fne_vcycle_op: n_rows < 0");
    return FNE_EINVAL;
  }
  if (n_rows == 0) return FNE_OK;
  const bool need_x = op != kSweepFirst, need_r = op != kAddProduct, need_w = op == kSweepFirst || op == kSweep;
  if (!indptr || !indices || !vals || !y || (need_x && !x) || (need_r && !r) || (need_w && !w) ||
      (op == kAddProduct && !y_in)) {
    fne_set_error("fne_vcycle_op: null pointer");
    return FNE_EINVAL;
  }
  // every row reads neighbouring entries of x (or of r and w for the first sweep): y must not
  // alias them; only the prolongation may update its iterate in place (y == y_in)
  if (y == x || y == r || y == w) {
    fne_set_error("fne_vcycle_op: the output aliases an input read across rows");
    return FNE_EINVAL;
  }
  int lpr = (precision >> 8) & 0xff;
  if (lpr == 0) lpr = 8;
  if (lpr != 1 && lpr != 2 && lpr != 4 && lpr != 8 && lpr != 16 && lpr != 32 && lpr != 64) {
    fne_set_error("fne_vcycle_op: lanes per row must be a power of two <= 64");
    return FNE_EINVAL;
  }
  const int prec = precision & 0xff;
  const int form = csr_kernel_form(precision, indices, vals, lpr);
  int rc;
  if (prec == FNE_FP64)
    rc = vcycle_op_dispatch<double>(op, n_rows, indptr, indices, (const double*)vals, x, r, w, y_in, y, lpr, form,
                                    (hipStream_t)stream);
  else if (prec == FNE_FP32)
    rc = vcycle_op_dispatch<float>(op, n_rows, indptr, indices, (const float*)vals, x, r, w, y_in, y, lpr, form,
                                   (hipStream_t)stream);
  else {
    fne_set_error("fne_vcycle_op: precision must be FNE_FP64 or FNE_FP32");
    return FNE_EINVAL;
  }
  if (rc != FNE_OK) return rc;
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

// Chebyshev smoothing step of the multigrid levels: y = x + c (x - x_prev) + w.(r - A x), one CSR
// pass.  With w = (2 rho_k / delta) D^-1 and c = rho_k rho_{k-1} it is step k of the Chebyshev
// iteration for D^-1 A on [a, b] (theta = (a + b) / 2, delta = (b - a) / 2); x_prev may be null
// when c = 0.  y may alias x_prev (each row reads its own x_prev entry before writing it).
namespace {
template <typename VT, int LPR, int FORM>
__global__ __launch_bounds__(256) void vcycle_cheb_kernel(int n, const int32_t* __restrict__ indptr,
                                                          const int32_t* __restrict__ indices,
                                                          const VT* __restrict__ vals, const double* __restrict__ x,
                                                          const double* x_prev, const double* __restrict__ r,
                                                          const double* __restrict__ w, double c, double* y) {
  __shared__ double prod[FORM == 2 ? kStreamCap : 1];
  int row;
  bool leader;
  const double acc =
      csr_form_dot<VT, LPR, FORM>(n, indptr, indices, vals, [&](int col) { return x[col]; }, prod, row, leader);
  if (leader) {
    const double xr = x[row];
    double out = xr + w[row] * (r[row] - acc);
    if (x_prev) out += c * (xr - x_prev[row]);
    y[row] = out;
  }
}

template <typename VT>
int launch_vcycle_cheb(int n, const int32_t* indptr, const int32_t* indices, const VT* vals, const double* x,
                       const double* x_prev, const double* r, const double* w, double c, double* y, int lpr,
                       int form, hipStream_t s) {
  auto go = [&](auto kern, int L) {
    const long long threads = (long long)n * L;
    const int nb = int((threads + 255) / 256);
    hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, s, n, indptr, indices, vals, x, x_prev, r, w, c, y);
  };
#define CHEB_L(L)                                                   \
  case L:                                                           \
    if (form == 2) go(vcycle_cheb_kernel<VT, L, 2>, L);              \
    else if (form == 1) go(vcycle_cheb_kernel<VT, L, 1>, L);         \
    else go(vcycle_cheb_kernel<VT, L, 0>, L);                        \
    break;
  switch (lpr) {
    CHEB_L(1)
    CHEB_L(2)
    CHEB_L(4)
    CHEB_L(8)
    CHEB_L(16)
    CHEB_L(32)
    default:
      if (form == 2) go(vcycle_cheb_kernel<VT, 64, 2>, 64);
      else if (form == 1) go(vcycle_cheb_kernel<VT, 64, 1>, 64);
      else go(vcycle_cheb_kernel<VT, 64, 0>, 64);
  }
#undef CHEB_L
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}
}  // namespace

extern "C" int fne_vcycle_cheb(int32_t n_rows, const int32_t* indptr, const int32_t* indices, const void* vals,
                               int32_t precision, const double* x, const double* x_prev, const double* r,
                               const double* w, double c, double* y, fne_stream_t stream) {
  if (n_rows < 0) {
    fne_set_error("fne_vcycle_cheb: n_rows < 0");
    return FNE_EINVAL;
  }
  if (n_rows == 0) return FNE_OK;
  if (!indptr || !indices || !vals || !x || !r || !w || !y || (c != 0.0 && !x_prev)) {
    fne_set_error("fne_vcycle_cheb: null pointer");
    return FNE_EINVAL;
  }
  if (y == x || y == r || y == w) {
    fne_set_error("fne_vcycle_cheb: the output aliases an input read across rows");
    return FNE_EINVAL;
  }
  int lpr = (precision >> 8) & 0xff;
  if (lpr == 0) lpr = 8;
  if (lpr != 1 && lpr != 2 && lpr != 4 && lpr != 8 && lpr != 16 && lpr != 32 && lpr != 64) {
    fne_set_error("fne_vcycle_cheb: lanes per row must be a power of two <= 64");
    return FNE_EINVAL;
  }
  const int prec = precision & 0xff;
  const int form = csr_kernel_form(precision, indices, vals, lpr);
  hipStream_t s = (hipStream_t)stream;
  if (prec == FNE_FP64)
    return launch_vcycle_cheb<double>(n_rows, indptr, indices, (const double*)vals, x, x_prev, r, w, c, y, lpr, form,
                                      s);
  if (prec == FNE_FP32)
    return launch_vcycle_cheb<float>(n_rows, indptr, indices, (const float*)vals, x, x_prev, r, w, c, y, lpr, form,
                                     s);
  fne_set_error("fne_vcycle_cheb: precision must be FNE_FP64 or FNE_FP32");
  return FNE_EINVAL;
}

extern "C" int fne_spgemm_numeric(int32_t n_rows, const int32_t* a_indptr, const int32_t* a_indices,
                                  const void* a_vals, int32_t a_precision, const int32_t* b_indptr,
                                  const int32_t* b_indices, const double* b_vals, const int32_t* c_indptr,
                                  const int32_t* c_indices, double* c_vals, fne_stream_t stream) {
  if (n_rows < 0) {
    fne_set_error("fne_spgemm_numeric: n_rows < 0");
    return FNE_EINVAL;
  }
  if (n_rows == 0) return FNE_OK;
  if (!a_indptr || !a_indices || !a_vals || !b_indptr || !b_indices || !b_vals || !c_indptr || !c_indices ||
      !c_vals) {
    fne_set_error("fne_spgemm_numeric: null pointer");
    return FNE_EINVAL;
  }
  const dim3 grid(n_rows), block(kLanes);
  hipStream_t s = (hipStream_t)stream;
  if (a_precision == FNE_FP64)
    hipLaunchKernelGGL(spgemm_numeric_kernel<double>, grid, block, 0, s, n_rows, a_indptr, a_indices,
                       (const double*)a_vals, b_indptr, b_indices, b_vals, c_indptr, c_indices, c_vals);
  else if (a_precision == FNE_FP32)
    hipLaunchKernelGGL(spgemm_numeric_kernel<float>, grid, block, 0, s, n_rows, a_indptr, a_indices,
                       (const float*)a_vals, b_indptr, b_indices, b_vals, c_indptr, c_indices, c_vals);
  else {
    fne_set_error("fne_spgemm_numeric: precision must be FNE_FP64 or FNE_FP32");
    return FNE_EINVAL;
  }
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}
