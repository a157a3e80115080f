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
// kChunk; for each nonzero a_ip of A's row, in order, the lanes take the entries of B's row p
// (distinct columns, so distinct accumulator slots), binary-search their slot in the staged
// columns and add a_ip * b_pj.  A barrier closes every p, so each result entry is the sum of
// its products in the order of A's row: the product is deterministic.
#include "fne_common.h"

namespace {

constexpr int kChunk = 1024;
constexpr int kLanes = 64;

template <typename TA>
__global__ __launch_bounds__(kLanes) void spgemm_numeric_kernel(
    int n_rows, const int32_t* __restrict__ a_ptr, const int32_t* __restrict__ a_idx, const TA* __restrict__ a_val,
    const int32_t* __restrict__ b_ptr, const int32_t* __restrict__ b_idx, const double* __restrict__ b_val,
    const int32_t* __restrict__ c_ptr, const int32_t* __restrict__ c_idx, double* __restrict__ c_val) {
  __shared__ int32_t s_col[kChunk];
  __shared__ double s_acc[kChunk];
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
    for (int p = a0; p < a1; ++p) {
      const int l = a_idx[p];
      const double a = static_cast<double>(a_val[p]);
      const int b0 = b_ptr[l], b1 = b_ptr[l + 1];
      for (int q = b0 + lane; q < b1; q += kLanes) {
        const int j = b_idx[q];
        if (j < first || j > last) continue;
        int lo = 0, hi = len - 1;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (s_col[mid] < j)
            lo = mid + 1;
          else
            hi = mid;
        }
        if (s_col[lo] == j) s_acc[lo] += a * b_val[q];
      }
      __syncthreads();
    }
    for (int t = lane; t < len; t += kLanes) c_val[cs + t] = s_acc[t];
    __syncthreads();
  }
}

}  // namespace

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
  if (a_precision == FNE_FP64)
    hipLaunchKernelGGL(spgemm_numeric_kernel<double>, grid, block, 0, (hipStream_t)stream, n_rows, a_indptr,
                       a_indices, (const double*)a_vals, b_indptr, b_indices, b_vals, c_indptr, c_indices, c_vals);
  else if (a_precision == FNE_FP32)
    hipLaunchKernelGGL(spgemm_numeric_kernel<float>, grid, block, 0, (hipStream_t)stream, n_rows, a_indptr,
                       a_indices, (const float*)a_vals, b_indptr, b_indices, b_vals, c_indptr, c_indices, c_vals);
  else {
    fne_set_error("fne_spgemm_numeric: precision must be FNE_FP64 or FNE_FP32");
    return FNE_EINVAL;
  }
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}
