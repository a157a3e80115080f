// Shared device/host helpers for libfne_hip (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/fne_hip.h"

// ---------------------------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------------------------
void fne_set_error(const std::string& msg);

#define FNE_HIP_CHECK(call)                                                                     \
  do {                                                                                          \
    hipError_t _e = (call);                                                                     \
    if (_e != hipSuccess) {                                                                     \
      fne_set_error(std::string("HIP error ") + hipGetErrorString(_e) + " at " + __FILE__ +     \
                    ":" + std::to_string(__LINE__));                                            \
      return FNE_EHIP;                                                                          \
    }                                                                                           \
  } while (0)

#define FNE_LAUNCH_CHECK() FNE_HIP_CHECK(hipGetLastError())

// ---------------------------------------------------------------------------------------------
// plan object
// ---------------------------------------------------------------------------------------------
struct fne_plan {
  fne_plan_desc d;
  double2* partials = nullptr;  // per-block (sum, compensation) pairs for energy reductions
  int n_partials = 0;
  int device = 0;
  int mode = 0;  // constitutive mode (Mohr-Coulomb: 0 full, 1 elastic)
  size_t max_lds = 65536;  // dynamic LDS bytes one workgroup may request on this device
  // Mohr-Coulomb: per-element gradient [n_elem][32] (FP64) and element matrices
  // [n_elem][n_loc][n_loc] (FP64 or FP32), written by the element kernel, read by the row gather
  double* mc_g = nullptr;
  void* mc_h = nullptr;
};

constexpr int kEnergyBlocks = 2048;  // fixed grid => fixed reduction order => deterministic
constexpr int kBlock = 256;

// ---------------------------------------------------------------------------------------------
// compensated (error-free) summation helpers
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bp = s - a;
  e = (a - (s - bp)) + (b - bp);
}

// (s, c) += v   (Neumaier / TwoSum accumulation)
__device__ __forceinline__ void acc_add(double& s, double& c, double v) {
  double t, e;
  two_sum(s, v, t, e);
  s = t;
  c += e;
}

// (s, c) += (s2, c2)
__device__ __forceinline__ void acc_merge(double& s, double& c, double s2, double c2) {
  double t, e;
  two_sum(s, s2, t, e);
  s = t;
  c += c2 + e;
}

// Deterministic block reduction of (s, c) pairs: fixed xor-shuffle tree inside each 64-wide
// wavefront, then wave 0 merges the per-wave results in wave order.
template <int NT = kBlock>
__device__ __forceinline__ void block_reduce_pair(double& s, double& c) {
  __shared__ double sh_s[NT / 64];
  __shared__ double sh_c[NT / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double s2 = __shfl_xor(s, off, 64);
    double c2 = __shfl_xor(c, off, 64);
    acc_merge(s, c, s2, c2);
  }
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (lane == 0) {
    sh_s[wave] = s;
    sh_c[wave] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double S = sh_s[0], C = sh_c[0];
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 1; w < nw; ++w) acc_merge(S, C, sh_s[w], sh_c[w]);
    s = S;
    c = C;
  }
}

// ---------------------------------------------------------------------------------------------
// XCD-aware block -> patch mapping.  MI355X dispatches consecutive workgroups round-robin over
// its 8 XCDs; neighbouring patches share halo elements, so we give each XCD a contiguous run of
// patches (its private L2 then sees the halo re-reads).  The map is a bijection on [0, n).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int xcd_remap(int b, int n) {
  constexpr int kXcd = 8;
  const int q = n / kXcd, rem = n % kXcd;
  const int x = b % kXcd, i = b / kXcd;
  const int start = x * q + (x < rem ? x : rem);
  return start + i;
}

// ---------------------------------------------------------------------------------------------
// CSR row dot product sum_k vals[k] * xf(indices[k]) over the row's entries [kb, ke), LPR lanes
// per row.  VEC (long rows: the vector problems' 40-80-entry rows): lane l walks the 4-aligned
// groups g = (kb & ~3) + 4 l + 4 LPR i with one 16-byte load of four column indices and one
// (FP32) or two (FP64) 16-byte loads of four values; every index of such a group is a column of
// some row (the group lies inside [0, nnz)), so the four gathers are issued unconditionally and
// an entry outside [kb, ke) is dropped by a select afterwards -- no branch between the loads.  A
// group running past the end of the arrays (g + 4 > nnz) is read entry by entry.  Needs indices
// and vals 16-byte aligned (the launchers check).  !VEC (short rows, e.g. the 7-entry P1 rows):
// one entry per lane per step.  Each lane sums its entries in increasing k; the caller reduces
// the LPR lanes with a fixed xor tree, so the result is bitwise reproducible.
// ---------------------------------------------------------------------------------------------
template <typename VT> struct CsrGroup4;
template <> struct CsrGroup4<float> {
  static __device__ __forceinline__ void load(const float* __restrict__ p, double v[4]) {
    const float4 t = *reinterpret_cast<const float4*>(p);
    v[0] = t.x;
    v[1] = t.y;
    v[2] = t.z;
    v[3] = t.w;
  }
};
template <> struct CsrGroup4<double> {
  static __device__ __forceinline__ void load(const double* __restrict__ p, double v[4]) {
    const double2 a = reinterpret_cast<const double2*>(p)[0];
    const double2 b = reinterpret_cast<const double2*>(p)[1];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
  }
};

template <typename VT, int LPR, bool VEC, typename XF>
__device__ __forceinline__ double csr_row_dot(int kb, int ke, int nnz, int lane, const int32_t* __restrict__ indices,
                                              const VT* __restrict__ vals, XF xf) {
  double acc = 0.0;
  if (!VEC) {
    for (int k = kb + lane; k < ke; k += LPR) acc += double(vals[k]) * xf(indices[k]);
    return acc;
  }
  for (int g = (kb & ~3) + 4 * lane; g < ke; g += 4 * LPR) {
    if (g + 4 <= nnz) {
      const int4 c4 = *reinterpret_cast<const int4*>(indices + g);
      double v[4];
      CsrGroup4<VT>::load(vals + g, v);
      const double x0 = xf(c4.x), x1 = xf(c4.y), x2 = xf(c4.z), x3 = xf(c4.w);
      const double t[4] = {v[0] * x0, v[1] * x1, v[2] * x2, v[3] * x3};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = g + j;
        acc += (k >= kb && k < ke) ? t[j] : 0.0;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = g + j;
        if (k >= kb && k < ke) acc += double(vals[k]) * xf(indices[k]);
      }
    }
  }
  return acc;
}

// grouped loads (VEC) for long rows, i.e. from 8 lanes per row up, when both arrays are aligned
__host__ __forceinline__ bool csr_use_vec(const void* indices, const void* vals, int lpr) {
  return lpr >= 8 && (reinterpret_cast<uintptr_t>(indices) & 15u) == 0 &&
         (reinterpret_cast<uintptr_t>(vals) & 15u) == 0;
}

// lanes per CSR row: about three entries per lane for short rows; from 8 lanes up (rows of 13+
// entries) enough groups of four for an average row and its unaligned start in one pass
__host__ __forceinline__ int csr_lanes_per_row(double avg_row_len) {
  int l = 1;
  while (l < 64 && 3.0 * l < avg_row_len) l *= 2;
  if (l >= 8)
    while (l < 64 && 4.0 * l < avg_row_len + 3.0) l *= 2;
  return l;
}

// ---------------------------------------------------------------------------------------------
// CSR-stream form (Greathouse and Daga, "Efficient sparse matrix-vector multiplication on GPUs
// using the CSR storage format", SC'14): a 256-thread workgroup takes R = 256 / LPR consecutive
// rows [r0, r1) and streams their entries [indptr[r0], indptr[r1]) through LDS in chunks of
// kStreamCap, each thread one 16-byte group of four column indices and four values per pass --
// independent, coalesced loads, many in flight -- as FP64 products vals[k] * xf(indices[k]);
// then LPR lanes per row sum their row's products from LDS (lane l: entries kb + l, kb + l + LPR,
// ... in increasing k, chunk by chunk).  The caller joins the LPR lanes with a fixed xor tree, so
// the result is bitwise reproducible.  Every thread of the workgroup must call it (barriers);
// indices and vals must be 16-byte aligned (csr_stream_ok); prod is kStreamCap doubles of LDS.
// ---------------------------------------------------------------------------------------------
constexpr int kStreamCap = 2048;

template <typename VT, int LPR, typename XF>
__device__ __forceinline__ double csr_stream_dot(int r0, int r1, int nnz, const int32_t* __restrict__ indptr,
                                                 const int32_t* __restrict__ indices, const VT* __restrict__ vals,
                                                 XF xf, double* prod) {
  const int tid = threadIdx.x;
  const int row = r0 + tid / LPR, lane = tid % LPR;
  const bool has = row < r1;
  const int base = indptr[r0], end = indptr[r1];
  const int kb = has ? indptr[row] : 0, ke = has ? indptr[row + 1] : 0;
  double acc = 0.0;
  for (int c0 = base & ~3; c0 < end; c0 += kStreamCap) {
#pragma unroll
    for (int j = 4 * tid; j < kStreamCap; j += 4 * 256) {
      const int g = c0 + j;
      double t[4] = {0.0, 0.0, 0.0, 0.0};
      if (g < end) {
        if (g + 4 <= nnz) {
          const int4 c4 = *reinterpret_cast<const int4*>(indices + g);
          double v[4];
          CsrGroup4<VT>::load(vals + g, v);
          const double x0 = xf(c4.x), x1 = xf(c4.y), x2 = xf(c4.z), x3 = xf(c4.w);
          const double p[4] = {v[0] * x0, v[1] * x1, v[2] * x2, v[3] * x3};
#pragma unroll
          for (int q = 0; q < 4; ++q) t[q] = (g + q >= base && g + q < end) ? p[q] : 0.0;
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (g + q >= base && g + q < end) t[q] = double(vals[g + q]) * xf(indices[g + q]);
        }
      }
      reinterpret_cast<double2*>(prod + j)[0] = make_double2(t[0], t[1]);
      reinterpret_cast<double2*>(prod + j)[1] = make_double2(t[2], t[3]);
    }
    __syncthreads();
    if (has) {
      const int lo = kb > c0 ? kb : c0;
      const int hi = ke < c0 + kStreamCap ? ke : c0 + kStreamCap;
      for (int k = lo + lane; k < hi; k += LPR) acc += prod[k - c0];
    }
    __syncthreads();
  }
  return acc;
}

// One CSR row pass of a 256-thread row kernel in either form, lanes joined by the fixed xor
// tree: FORM 0 lanes per row (row r = thread / LPR over the grid), 1 the same with grouped
// loads, 2 CSR-stream (rows blockIdx * 256 / LPR onward, prod = kStreamCap doubles of LDS).
// Sets row (n when the thread has none) and leader (the thread that writes the row's result).
// All threads of the workgroup must call it.  Workgroups: ceil(n * LPR / 256) in every form.
template <typename VT, int LPR, int FORM, typename XF>
__device__ __forceinline__ double csr_form_dot(int n, const int32_t* __restrict__ indptr,
                                               const int32_t* __restrict__ indices, const VT* __restrict__ vals,
                                               XF xf, double* prod, int& row, bool& leader) {
  double acc = 0.0;
  int lane;
  if constexpr (FORM == 2) {
    constexpr int R = 256 / LPR;
    const int r0 = xcd_remap(blockIdx.x, gridDim.x) * R;
    const int r1 = r0 + R < n ? r0 + R : n;
    acc = csr_stream_dot<VT, LPR>(r0, r1, indptr[n], indptr, indices, vals, xf, prod);
    row = r0 + int(threadIdx.x) / LPR;
    lane = int(threadIdx.x) % LPR;
    if (row >= r1) row = n;
  } else {
    const int gtid = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    row = gtid / LPR;
    lane = gtid % LPR;
    if (row < n) acc = csr_row_dot<VT, LPR, FORM == 1>(indptr[row], indptr[row + 1], indptr[n], lane, indices, vals, xf);
    else row = n;
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
  leader = row < n && lane == 0;
  return acc;
}

// lanes per row of the CSR-stream form: about 8 entries per lane (rows per workgroup 256 / lanes)
__host__ __forceinline__ int csr_stream_lanes(double avg_row_len) {
  int l = 1;
  while (l < 64 && 2.0 * l <= avg_row_len / 8.0) l *= 2;
  return l;
}

__host__ __forceinline__ bool csr_stream_ok(const void* indices, const void* vals) {
  return (reinterpret_cast<uintptr_t>(indices) & 15u) == 0 && (reinterpret_cast<uintptr_t>(vals) & 15u) == 0;
}

// the SpMV-form field of the precision argument of fne_spmv / fne_vcycle_op / fne_vcycle_cheb
// (bits 16..19): 0 = lanes per row, 1 = CSR-stream
constexpr int kCsrFormStream = 1;
__host__ __forceinline__ int csr_form_of(int32_t precision) { return (precision >> 16) & 0xF; }

// gather one element value: free DOF from x, constrained DOF from the Dirichlet vector
__device__ __forceinline__ double gather_dof(int c, int n_free, const double* __restrict__ x,
                                             const double* __restrict__ dvals) {
  return c < n_free ? x[c] : dvals[c - n_free];
}

// family launchers (one translation unit per family)
enum FneOp { kOpEnergy = 0, kOpGradient = 1, kOpHessian = 2, kOpElementEval = 3, kOpFused = 4 };

struct FneCall {
  const double* x;
  const double* dvals;
  const double* f;
  double* out;   // energy scalar
  double* g;     // gradient
  void* vals;    // CSR values
  int precision;
  double* e_out;
  double* ge_out;
  double* he_out;
  hipStream_t stream;
};

int fne_launch_p1(fne_plan* plan, FneOp op, const FneCall& c);
int fne_launch_tet(fne_plan* plan, FneOp op, const FneCall& c);
int fne_launch_mc(fne_plan* plan, FneOp op, const FneCall& c);
int fne_finalize_energy(fne_plan* plan, int nblocks, double* out, hipStream_t stream);
// bytes of the element-matrix scratch (and, via gbytes, of the gradient scratch) used by the
// row-gather assembly (Mohr-Coulomb always, tetrahedra with FNE_FLAG_ROW_GATHER)
size_t fne_mc_scratch_bytes(const fne_plan_desc& d, size_t* gbytes);
// CSR rows gather their incident element rows from the scratch (n_loc 12 or 30)
int fne_launch_row_gather(fne_plan* plan, bool want_g, bool want_h, const FneCall& c);
// Hessian-vector products of the P1 (staged plans) and tetrahedral (row-gather plans) families
int fne_hvp_p1(fne_plan* plan, const double* x, const double* dvals, const double* v, double* y, hipStream_t s);
int fne_hvp_tet(fne_plan* plan, const double* x, const double* dvals, const double* v, double* y, hipStream_t s);
