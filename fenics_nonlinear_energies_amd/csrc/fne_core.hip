// libfne_hip core: plan lifecycle, dispatch, energy finalisation, CSR SpMV, diagonal, dot.
#include <cstring>

#include "fne_common.h"

namespace {
thread_local std::string g_last_error;

// Final deterministic reduction of the per-block (sum, compensation) partials.
// One 1024-thread block; four independent loads in flight per thread, merged in index order.
constexpr int kFinalizeThreads = 1024;
__global__ __launch_bounds__(kFinalizeThreads) void finalize_pairs_kernel(const double2* __restrict__ partial,
                                                                          int n, double* __restrict__ out) {
  double s = 0.0, c = 0.0;
  constexpr int T = kFinalizeThreads;
  int i = threadIdx.x;
  for (; i + 3 * T < n; i += 4 * T) {
    const double2 a = partial[i], b = partial[i + T], e = partial[i + 2 * T], f = partial[i + 3 * T];
    acc_merge(s, c, a.x, a.y);
    acc_merge(s, c, b.x, b.y);
    acc_merge(s, c, e.x, e.y);
    acc_merge(s, c, f.x, f.y);
  }
  for (; i < n; i += T) acc_merge(s, c, partial[i].x, partial[i].y);
  block_reduce_pair<T>(s, c);
  if (threadIdx.x == 0) out[0] = s + c;
}

// y = A x with LPR lanes per row (LPR | 64), each lane taking groups of four entries
// (csr_row_dot).  Per-lane partial sums in a fixed order, then a fixed xor tree inside the lane
// group -> bitwise reproducible.
template <typename VT, int LPR, bool VEC>
__global__ __launch_bounds__(kBlock) void spmv_kernel(int n, const int32_t* __restrict__ indptr,
                                                      const int32_t* __restrict__ indices,
                                                      const VT* __restrict__ vals,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y) {
  // XCD-aware: each XCD's workgroups take a contiguous band of rows, so the gathered x
  // entries of a band stay in that XCD's L2
  const int gtid = xcd_remap(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int row = gtid / LPR;
  const int lane = gtid % LPR;
  double acc = 0.0;
  if (row < n) {
    const int kb = indptr[row], ke = indptr[row + 1];
    acc = csr_row_dot<VT, LPR, VEC>(kb, ke, indptr[n], lane, indices, vals, [&](int c) { return x[c]; });
  }
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
  if (row < n && lane == 0) y[row] = acc;
}

// y = A x in the CSR-stream form (csr_stream_dot): R = 256 / LPR rows per workgroup.
template <typename VT, int LPR>
__global__ __launch_bounds__(kBlock) void spmv_stream_kernel(int n, const int32_t* __restrict__ indptr,
                                                             const int32_t* __restrict__ indices,
                                                             const VT* __restrict__ vals,
                                                             const double* __restrict__ x,
                                                             double* __restrict__ y) {
  __shared__ double prod[kStreamCap];
  constexpr int R = kBlock / LPR;
  const int r0 = xcd_remap(blockIdx.x, gridDim.x) * R;
  const int r1 = min(r0 + R, n);
  double acc = csr_stream_dot<VT, LPR>(r0, r1, indptr[n], indptr, indices, vals, [&](int c) { return x[c]; }, prod);
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
  const int row = r0 + int(threadIdx.x) / LPR;
  if (row < r1 && int(threadIdx.x) % LPR == 0) y[row] = acc;
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void diag_kernel(int n, const int32_t* __restrict__ indptr,
                                                      const int32_t* __restrict__ indices,
                                                      const VT* __restrict__ vals,
                                                      double* __restrict__ diag) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  int lo = indptr[row], hi = indptr[row + 1];
  double dv = 0.0;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int cm = indices[mid];
    if (cm == row) {
      dv = double(vals[mid]);
      break;
    }
    if (cm < row)
      lo = mid + 1;
    else
      hi = mid;
  }
  diag[row] = dv;
}

__global__ __launch_bounds__(kBlock) void dot_kernel(int64_t n, const double* __restrict__ a,
                                                     const double* __restrict__ b,
                                                     double2* __restrict__ partial) {
  double s = 0.0, c = 0.0;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    acc_add(s, c, a[i] * b[i]);
  block_reduce_pair(s, c);
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(s, c);
}

template <typename VT>
int launch_spmv_stream(int n, const int32_t* indptr, const int32_t* indices, const VT* vals,
                       const double* x, double* y, hipStream_t stream, int lpr) {
  auto go = [&](auto kern, int L) {
    const int rows = kBlock / L;
    const int nb = int((static_cast<long long>(n) + rows - 1) / rows);
    hipLaunchKernelGGL(kern, dim3(nb), dim3(kBlock), 0, stream, n, indptr, indices, vals, x, y);
  };
  switch (lpr) {
    case 1: go(spmv_stream_kernel<VT, 1>, 1); break;
    case 2: go(spmv_stream_kernel<VT, 2>, 2); break;
    case 4: go(spmv_stream_kernel<VT, 4>, 4); break;
    case 8: go(spmv_stream_kernel<VT, 8>, 8); break;
    case 16: go(spmv_stream_kernel<VT, 16>, 16); break;
    case 32: go(spmv_stream_kernel<VT, 32>, 32); break;
    default: go(spmv_stream_kernel<VT, 64>, 64); break;
  }
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

template <typename VT>
int launch_spmv(int n, const int32_t* indptr, const int32_t* indices, const VT* vals,
                const double* x, double* y, hipStream_t stream, int lpr) {
  auto go = [&](auto kern, int L) {
    const long long threads = (long long)n * L;
    const int nb = int((threads + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(kern, dim3(nb), dim3(kBlock), 0, stream, n, indptr, indices, vals, x, y);
  };
  if (csr_use_vec(indices, vals, lpr)) {
    switch (lpr) {
      case 1: go(spmv_kernel<VT, 1, true>, 1); break;
      case 2: go(spmv_kernel<VT, 2, true>, 2); break;
      case 4: go(spmv_kernel<VT, 4, true>, 4); break;
      case 8: go(spmv_kernel<VT, 8, true>, 8); break;
      case 16: go(spmv_kernel<VT, 16, true>, 16); break;
      case 32: go(spmv_kernel<VT, 32, true>, 32); break;
      default: go(spmv_kernel<VT, 64, true>, 64); break;
    }
  } else {
    switch (lpr) {
      case 1: go(spmv_kernel<VT, 1, false>, 1); break;
      case 2: go(spmv_kernel<VT, 2, false>, 2); break;
      case 4: go(spmv_kernel<VT, 4, false>, 4); break;
      case 8: go(spmv_kernel<VT, 8, false>, 8); break;
      case 16: go(spmv_kernel<VT, 16, false>, 16); break;
      case 32: go(spmv_kernel<VT, 32, false>, 32); break;
      default: go(spmv_kernel<VT, 64, false>, 64); break;
    }
  }
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

int validate_desc(const fne_plan_desc* d) {
  if (!d) return FNE_EINVAL;
  if (d->n_elem < 0 || d->n_free < 0 || d->n_total < 0 || d->nnz < 0) {
    fne_set_error("negative size in plan descriptor");
    return FNE_EINVAL;
  }
  const bool node_geo = (d->flags & FNE_FLAG_NODE_GEO) != 0;
  if (d->n_elem > 0 && (!d->edof || (node_geo ? !d->dof_xy : !d->egeo))) {
    fne_set_error(node_geo ? "plan descriptor with FNE_FLAG_NODE_GEO is missing edof or dof_xy"
                           : "plan descriptor is missing element arrays");
    return FNE_EINVAL;
  }
  if (node_geo && d->kind != FNE_KIND_PLAPLACE && d->kind != FNE_KIND_GINZBURG_LANDAU) {
    fne_set_error("FNE_FLAG_NODE_GEO is defined for the P1 scalar families only");
    return FNE_EINVAL;
  }
  if (node_geo && (reinterpret_cast<uintptr_t>(d->dof_xy) & 15u) != 0) {
    fne_set_error("dof_xy must be 16-byte aligned");
    return FNE_EINVAL;
  }
  switch (d->kind) {
    case FNE_KIND_PLAPLACE:
    case FNE_KIND_GINZBURG_LANDAU:
      if (d->n_loc != 3 || d->n_geo != 7) {
        fne_set_error("P1 scalar plans need n_loc=3, n_geo=7");
        return FNE_EINVAL;
      }
      break;
    case FNE_KIND_NEOHOOKEAN:
      if (d->n_loc != 12 || d->n_geo != 13) {
        fne_set_error("Neo-Hookean plans need n_loc=12, n_geo=13");
        return FNE_EINVAL;
      }
      break;
    case FNE_KIND_MOHR_COULOMB:
      if (d->n_loc != 30 && d->n_loc != 12) {
        fne_set_error("Mohr-Coulomb plans need n_loc=30 (P4) or 12 (P2)");
        return FNE_EINVAL;
      }
      if (d->nq < 1 || d->nq > 21 || !d->qdata) {
        fne_set_error("Mohr-Coulomb plans need 1..21 quadrature points and a quadrature table");
        return FNE_EINVAL;
      }
      if (d->n_geo != 5 && d->n_geo != 5 + 3 * d->nq) {
        fne_set_error("Mohr-Coulomb plans need n_geo = 5 (K, |detJ|) or 5 + 3*nq (with eps_p_old)");
        return FNE_EINVAL;
      }
      break;
    default:
      fne_set_error("unknown energy kind " + std::to_string(d->kind));
      return FNE_EUNSUPPORTED;
  }
  if (d->n_patch > 0) {
    if (!d->prow || !d->pelem_ptr || !d->pelem || !d->inc_ptr || !d->inc || !d->indptr) {
      fne_set_error("patch schedule arrays missing");
      return FNE_EINVAL;
    }
    if (d->chunk <= 0 || d->chunk > 8192) {
      fne_set_error("chunk must be in 1..8192");
      return FNE_EINVAL;
    }
  }
  return FNE_OK;
}

int dispatch(fne_plan* plan, FneOp op, const FneCall& c) {
  switch (plan->d.kind) {
    case FNE_KIND_PLAPLACE:
    case FNE_KIND_GINZBURG_LANDAU:
      return fne_launch_p1(plan, op, c);
    case FNE_KIND_NEOHOOKEAN:
      return fne_launch_tet(plan, op, c);
    case FNE_KIND_MOHR_COULOMB:
      return fne_launch_mc(plan, op, c);
    default:
      fne_set_error("unknown energy kind");
      return FNE_EUNSUPPORTED;
  }
}

FneCall make_call(const double* x, const double* dvals, const double* f, hipStream_t s) {
  FneCall c;
  std::memset(&c, 0, sizeof(c));
  c.x = x;
  c.dvals = dvals;
  c.f = f;
  c.stream = s;
  return c;
}
}  // namespace

void fne_set_error(const std::string& msg) { g_last_error = msg; }

int fne_finalize_energy(fne_plan* plan, int nblocks, double* out, hipStream_t stream) {
  hipLaunchKernelGGL(finalize_pairs_kernel, dim3(1), dim3(kFinalizeThreads), 0, stream, plan->partials, nblocks,
                     out);
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

extern "C" {

const char* fne_last_error(void) { return g_last_error.c_str(); }
const char* fne_version(void) { return "fne_hip 0.1 gfx950"; }

int fne_plan_create(const fne_plan_desc* desc, fne_plan** out) {
  if (!out) return FNE_EINVAL;
  *out = nullptr;
  int rc = validate_desc(desc);
  if (rc) return rc;
  fne_plan* p = new (std::nothrow) fne_plan();
  if (!p) return FNE_ENOMEM;
  p->d = *desc;
  if (hipGetDevice(&p->device) != hipSuccess) p->device = 0;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device) == hipSuccess &&
      lds > 0)
    p->max_lds = size_t(lds);
  {
    // one partial per patch workgroup, per energy block (2*kBlock items each) or per fixed block
    const long items = desc->n_elem > desc->n_free ? desc->n_elem : desc->n_free;
    long need = (items + 2 * kBlock - 1) / (2 * kBlock);
    if (desc->n_patch > need) need = desc->n_patch;
    if (kEnergyBlocks > need) need = kEnergyBlocks;
    p->n_partials = int(need);
  }
  if (hipMalloc(&p->partials, sizeof(double2) * p->n_partials) != hipSuccess) {
    delete p;
    fne_set_error("hipMalloc of reduction scratch failed");
    return FNE_ENOMEM;
  }
  if ((desc->kind == FNE_KIND_MOHR_COULOMB ||
       (desc->kind == FNE_KIND_NEOHOOKEAN && (desc->flags & FNE_FLAG_ROW_GATHER))) &&
      desc->n_elem > 0) {
    size_t gbytes = 0;
    const size_t hbytes = fne_mc_scratch_bytes(*desc, &gbytes);
    if (hipMalloc(&p->mc_h, hbytes) != hipSuccess || hipMalloc(&p->mc_g, gbytes) != hipSuccess) {
      if (p->mc_h) (void)hipFree(p->mc_h);
      (void)hipFree(p->partials);
      delete p;
      fne_set_error("hipMalloc of the element scratch failed");
      return FNE_ENOMEM;
    }
  }
  *out = p;
  return FNE_OK;
}

int fne_plan_destroy(fne_plan* plan) {
  if (!plan) return FNE_OK;
  if (plan->partials) (void)hipFree(plan->partials);
  if (plan->mc_h) (void)hipFree(plan->mc_h);
  if (plan->mc_g) (void)hipFree(plan->mc_g);
  delete plan;
  return FNE_OK;
}

int fne_plan_set_params(fne_plan* plan, const double* param8) {
  if (!plan || !param8) return FNE_EINVAL;
  std::memcpy(plan->d.param, param8, sizeof(plan->d.param));
  return FNE_OK;
}

int fne_plan_set_mode(fne_plan* plan, int32_t mode) {
  if (!plan) return FNE_EINVAL;
  if (mode < 0 || mode > 1) {
    fne_set_error("constitutive mode must be 0 (full) or 1 (elastic)");
    return FNE_EINVAL;
  }
  plan->mode = mode;
  return FNE_OK;
}

int fne_energy(fne_plan* plan, const double* x, const double* dvals, const double* f, double* out,
               fne_stream_t stream) {
  if (!plan || !out) return FNE_EINVAL;
  FneCall c = make_call(x, dvals, f, (hipStream_t)stream);
  c.out = out;
  return dispatch(plan, kOpEnergy, c);
}

int fne_gradient(fne_plan* plan, const double* x, const double* dvals, const double* f, double* g,
                 fne_stream_t stream) {
  if (!plan || !g) return FNE_EINVAL;
  FneCall c = make_call(x, dvals, f, (hipStream_t)stream);
  c.g = g;
  return dispatch(plan, kOpGradient, c);
}

int fne_hessian(fne_plan* plan, const double* x, const double* dvals, const double* f, double* g,
                void* vals, int32_t precision, fne_stream_t stream) {
  if (!plan || !vals) return FNE_EINVAL;
  if (precision != FNE_FP64 && precision != FNE_FP32) {
    fne_set_error("precision must be FNE_FP64 or FNE_FP32");
    return FNE_EINVAL;
  }
  FneCall c = make_call(x, dvals, f, (hipStream_t)stream);
  c.g = g;
  c.vals = vals;
  c.precision = precision;
  return dispatch(plan, kOpHessian, c);
}

int fne_energy_grad_hessian(fne_plan* plan, const double* x, const double* dvals, const double* f,
                            double* out, double* g, void* vals, int32_t precision, fne_stream_t stream) {
  if (!plan || !vals) return FNE_EINVAL;
  if (precision != FNE_FP64 && precision != FNE_FP32) {
    fne_set_error("precision must be FNE_FP64 or FNE_FP32");
    return FNE_EINVAL;
  }
  FneCall c = make_call(x, dvals, f, (hipStream_t)stream);
  c.out = out;
  c.g = g;
  c.vals = vals;
  c.precision = precision;
  return dispatch(plan, kOpFused, c);
}

int fne_element_eval(fne_plan* plan, const double* x, const double* dvals, double* e_out,
                     double* ge_out, double* he_out, fne_stream_t stream) {
  if (!plan) return FNE_EINVAL;
  FneCall c = make_call(x, dvals, nullptr, (hipStream_t)stream);
  c.e_out = e_out;
  c.ge_out = ge_out;
  c.he_out = he_out;
  return dispatch(plan, kOpElementEval, c);
}

int fne_spmv(int32_t n_rows, const int32_t* indptr, const int32_t* indices, const void* vals,
             int32_t precision, const double* x, double* y, fne_stream_t stream) {
  if (n_rows < 0 || (n_rows > 0 && (!indptr || !indices || !vals || !x || !y))) return FNE_EINVAL;
  if (n_rows == 0) return FNE_OK;
  // lanes per row: bits 8..15 of precision (the caller knows the average row length; see
  // csr_lanes_per_row), 8 when not given
  int lpr = (precision >> 8) & 0xFF;
  if (lpr == 0) lpr = 8;
  if (lpr != 1 && lpr != 2 && lpr != 4 && lpr != 8 && lpr != 16 && lpr != 32 && lpr != 64) {
    fne_set_error("lanes per row must be a power of two <= 64");
    return FNE_EINVAL;
  }
  const int prec = precision & 0xFF;
  // bits 16..19: 1 = CSR-stream form (rows per workgroup 256 / lanes per row), when aligned
  if (csr_form_of(precision) == kCsrFormStream && csr_stream_ok(indices, vals)) {
    if (prec == FNE_FP32)
      return launch_spmv_stream<float>(n_rows, indptr, indices, (const float*)vals, x, y, (hipStream_t)stream, lpr);
    if (prec == FNE_FP64)
      return launch_spmv_stream<double>(n_rows, indptr, indices, (const double*)vals, x, y, (hipStream_t)stream,
                                        lpr);
  }
  if (prec == FNE_FP32)
    return launch_spmv<float>(n_rows, indptr, indices, (const float*)vals, x, y, (hipStream_t)stream,
                              lpr);
  if (prec == FNE_FP64)
    return launch_spmv<double>(n_rows, indptr, indices, (const double*)vals, x, y,
                               (hipStream_t)stream, lpr);
  fne_set_error("precision must be FNE_FP64 or FNE_FP32");
  return FNE_EINVAL;
}

int fne_csr_diag(int32_t n_rows, const int32_t* indptr, const int32_t* indices, const void* vals,
                 int32_t precision, double* diag, fne_stream_t stream) {
  if (n_rows <= 0) return n_rows == 0 ? FNE_OK : FNE_EINVAL;
  const int nb = (n_rows + kBlock - 1) / kBlock;
  if (precision == FNE_FP32)
    hipLaunchKernelGGL(diag_kernel<float>, dim3(nb), dim3(kBlock), 0, (hipStream_t)stream, n_rows,
                       indptr, indices, (const float*)vals, diag);
  else if (precision == FNE_FP64)
    hipLaunchKernelGGL(diag_kernel<double>, dim3(nb), dim3(kBlock), 0, (hipStream_t)stream, n_rows,
                       indptr, indices, (const double*)vals, diag);
  else {
    fne_set_error("precision must be FNE_FP64 or FNE_FP32");
    return FNE_EINVAL;
  }
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

int32_t fne_dot_work_len(void) { return kEnergyBlocks; }

int fne_dot(int64_t n, const double* a, const double* b, double* out, void* work_, fne_stream_t stream) {
  double2* work = static_cast<double2*>(work_);
  if (n < 0 || !out || !work_ || (n > 0 && (!a || !b))) {
    fne_set_error("fne_dot: negative length or null pointer");
    return FNE_EINVAL;
  }
  int nb = int((n + kBlock - 1) / kBlock);
  if (nb > kEnergyBlocks) nb = kEnergyBlocks;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(dot_kernel, dim3(nb), dim3(kBlock), 0, (hipStream_t)stream, n, a, b, work);
  FNE_LAUNCH_CHECK();
  hipLaunchKernelGGL(finalize_pairs_kernel, dim3(1), dim3(kFinalizeThreads), 0, (hipStream_t)stream, work, nb,
                     out);
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

}  // extern "C"
