// P1 triangle scalar energies: p-Laplace (A1) and Ginzburg-Landau (A2).
//
// Element data (SoA, element e of n_elem):
//   edof[a*n_elem+e], a=0..2          free index or n_free + Dirichlet index
//   egeo[k*n_elem+e], k=0..6          dvx0 dvx1 dvx2 dvy0 dvy1 dvy2 vol
// Math (hand-derived, replaces jax.grad / jax.hessian of
//   plaplace/jax_petsc/reordered_element_assembler.py:16-19 and
//   ginzburg_landau/jax_petsc/reordered_element_assembler.py:18-24):
//   gx = sum_a v_a dvx_a, gy = sum_a v_a dvy_a, s = gx^2 + gy^2, t_a = gx dvx_a + gy dvy_a
//   p-Laplace:  E = (1/p) s^(p/2) vol
//               dE/dv_a = vol s^(p/2-1) t_a
//               H_ab = vol [ s^(p/2-1) K_ab + (p-2) s^(p/2-2) t_a t_b ],  K_ab = dvx_a dvx_b + dvy_a dvy_b
//               (at s = 0 the second term is taken as its limit 0 for p > 2)
//   GL:         u_q = sum_a v_a ip[a][q]
//               E = vol [ eps/2 s + 1/4 sum_q w_q (u_q^2-1)^2 ]
//               dE/dv_a = vol [ eps t_a + sum_q w_q (u_q^2-1) u_q ip[a][q] ]
//               H_ab = vol [ eps K_ab + sum_q w_q (3 u_q^2 - 1) ip[a][q] ip[b][q] ]
#include "fne_common.h"

namespace {

constexpr int kMaxQ = 16;

struct P1Elem {
  double v[3];
  double dvx[3], dvy[3], vol;
};

__device__ __forceinline__ void p1_load(const fne_plan_desc& d, int e, const double* __restrict__ x,
                                        const double* __restrict__ dvals, P1Elem& el) {
  const int ne = d.n_elem;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int c = d.edof[a * ne + e];
    el.v[a] = gather_dof(c, d.n_free, x, dvals);
    el.dvx[a] = d.egeo[a * ne + e];
    el.dvy[a] = d.egeo[(3 + a) * ne + e];
  }
  el.vol = d.egeo[6 * ne + e];
}

// s^(p/2) and friends; p==2 and p==3 take exact fast paths
__device__ __forceinline__ double pl_pow_half(double s, double p) {
  if (p == 3.0) return s * sqrt(s);
  if (p == 2.0) return s;
  return pow(s, 0.5 * p);
}

__device__ __forceinline__ double p1_energy(int fam, const P1Elem& el, const double* prm,
                                            const double* __restrict__ qd, int nq) {
  double gx = 0.0, gy = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    gx += el.v[a] * el.dvx[a];
    gy += el.v[a] * el.dvy[a];
  }
  const double s = gx * gx + gy * gy;
  if (fam == FNE_KIND_PLAPLACE) {
    const double p = prm[0];
    return (1.0 / p) * pl_pow_half(s, p) * el.vol;
  }
  const double eps = prm[0];
  double pot = 0.0;
  for (int q = 0; q < nq; ++q) {
    const double uq = el.v[0] * qd[q] + el.v[1] * qd[nq + q] + el.v[2] * qd[2 * nq + q];
    const double t = uq * uq - 1.0;
    pot += t * t * qd[3 * nq + q];
  }
  return (0.5 * eps * s + 0.25 * pot) * el.vol;
}

// gradient (3) and symmetric Hessian (6: 00 11 22 01 02 12), FP64
__device__ __forceinline__ void p1_grad_hess(int fam, const P1Elem& el, const double* prm,
                                             const double* __restrict__ qd, int nq, bool want_g,
                                             bool want_h, double ge[3], double he[6]) {
  double gx = 0.0, gy = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    gx += el.v[a] * el.dvx[a];
    gy += el.v[a] * el.dvy[a];
  }
  const double s = gx * gx + gy * gy;
  double t[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) t[a] = gx * el.dvx[a] + gy * el.dvy[a];
  const int ia[6] = {0, 1, 2, 0, 0, 1};
  const int ib[6] = {0, 1, 2, 1, 2, 2};
  if (fam == FNE_KIND_PLAPLACE) {
    const double p = prm[0];
    // c1 = s^(p/2-1), c2 = (p-2) s^(p/2-2)
    double c1, c2;
    if (p == 2.0) {
      c1 = 1.0;
      c2 = 0.0;
    } else if (p == 3.0) {
      const double r = sqrt(s);
      c1 = r;
      c2 = s > 0.0 ? 1.0 / r : 0.0;
    } else {
      c1 = s > 0.0 ? pow(s, 0.5 * p - 1.0) : (p > 2.0 ? 0.0 : pow(s, 0.5 * p - 1.0));
      c2 = s > 0.0 ? (p - 2.0) * pow(s, 0.5 * p - 2.0) : 0.0;
    }
    if (want_g) {
#pragma unroll
      for (int a = 0; a < 3; ++a) ge[a] = el.vol * c1 * t[a];
    }
    if (want_h) {
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const int a = ia[k], b = ib[k];
        const double K = el.dvx[a] * el.dvx[b] + el.dvy[a] * el.dvy[b];
        he[k] = el.vol * (c1 * K + c2 * t[a] * t[b]);
      }
    }
    return;
  }
  // Ginzburg-Landau
  const double eps = prm[0];
  double gq[3] = {0.0, 0.0, 0.0};
  double hq[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int q = 0; q < nq; ++q) {
    const double i0 = qd[q], i1 = qd[nq + q], i2 = qd[2 * nq + q], w = qd[3 * nq + q];
    const double uq = el.v[0] * i0 + el.v[1] * i1 + el.v[2] * i2;
    const double u2 = uq * uq;
    const double gcoef = w * (u2 - 1.0) * uq;
    const double hcoef = w * (3.0 * u2 - 1.0);
    const double ip[3] = {i0, i1, i2};
#pragma unroll
    for (int a = 0; a < 3; ++a) gq[a] += gcoef * ip[a];
#pragma unroll
    for (int k = 0; k < 6; ++k) hq[k] += hcoef * ip[ia[k]] * ip[ib[k]];
  }
  if (want_g) {
#pragma unroll
    for (int a = 0; a < 3; ++a) ge[a] = el.vol * (eps * t[a] + gq[a]);
  }
  if (want_h) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int a = ia[k], b = ib[k];
      const double K = el.dvx[a] * el.dvx[b] + el.dvy[a] * el.dvy[b];
      he[k] = el.vol * (eps * K + hq[k]);
    }
  }
}

__device__ __forceinline__ int sym3(int a, int b) {
  // 00->0 11->1 22->2 01->3 02->4 12->5
  return a == b ? a : (a + b + 2);
}

// ---------------------------------------------------------------------------------------------
// energy: grid-stride over elements (+ load term over free DOFs), compensated block partials
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void p1_energy_kernel(fne_plan_desc d, const double* __restrict__ x,
                                                           const double* __restrict__ dvals,
                                                           const double* __restrict__ f,
                                                           double2* __restrict__ partial) {
  double s = 0.0, c = 0.0;
  const double* prm = d.param;
  const int stride = gridDim.x * blockDim.x;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < d.n_elem; e += stride) {
    P1Elem el;
    p1_load(d, e, x, dvals, el);
    double E = p1_energy(d.kind, el, prm, d.qdata, d.nq);
    if (d.eweight) E *= d.eweight[e];
    acc_add(s, c, E);
  }
  if (f) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n_free; i += stride)
      acc_add(s, c, -(f[i] * x[i]));
  }
  block_reduce_pair(s, c);
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(s, c);
}

// ---------------------------------------------------------------------------------------------
// owner-computes patch assembly.  One workgroup per patch of <= blockDim consecutive rows.
// Phase A: the patch's elements (owned + halo) are evaluated, chunk by chunk, into LDS.
// Phase B: each row thread walks its incidence list (sorted by element slot) and accumulates
//          its gradient entry and its CSR row slice in a fixed order -> deterministic, no atomics.
// Incidence item: slot[0:13) | a[13:15) | pos_b[15+5b : 20+5b) (31 = constrained column).
// ---------------------------------------------------------------------------------------------
template <typename HT, bool WANT_G, bool WANT_H>
__global__ __launch_bounds__(kBlock) void p1_patch_kernel(fne_plan_desc d, const double* __restrict__ x,
                                                          const double* __restrict__ dvals,
                                                          const double* __restrict__ f,
                                                          double* __restrict__ g, HT* __restrict__ vals) {
  extern __shared__ double smem[];
  const int p = xcd_remap(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int chunk = d.chunk;
  const int r0 = d.prow[p], r1 = d.prow[p + 1];
  const int k0 = d.indptr[r0];
  const int nslice = WANT_H ? d.indptr[r1] - k0 : 0;
  const int eb = d.pelem_ptr[p];
  const int ne = d.pelem_ptr[p + 1] - eb;

  double* sVal = smem;                       // [max_patch_nnz]
  double* sH = sVal + (WANT_H ? d.max_patch_nnz : 0);  // [6][chunk]
  double* sG = sH + (WANT_H ? 6 * chunk : 0);           // [3][chunk]

  if (WANT_H)
    for (int k = tid; k < nslice; k += blockDim.x) sVal[k] = 0.0;

  const int r = r0 + tid;
  const bool has_row = r < r1;
  int ip = has_row ? d.inc_ptr[r] : 0;
  const int ie = has_row ? d.inc_ptr[r + 1] : 0;
  const int roff = has_row ? d.indptr[r] - k0 : 0;
  double gacc = 0.0;
  const double* prm = d.param;
  __syncthreads();

  for (int c0 = 0; c0 < ne; c0 += chunk) {
    const int nc = min(chunk, ne - c0);
    for (int sl = tid; sl < nc; sl += blockDim.x) {
      const int e = d.pelem[eb + c0 + sl];
      P1Elem el;
      p1_load(d, e, x, dvals, el);
      double ge[3], he[6];
      p1_grad_hess(d.kind, el, prm, d.qdata, d.nq, WANT_G, WANT_H, ge, he);
      if (WANT_H) {
#pragma unroll
        for (int k = 0; k < 6; ++k) sH[k * chunk + sl] = he[k];
      }
      if (WANT_G) {
#pragma unroll
        for (int a = 0; a < 3; ++a) sG[a * chunk + sl] = ge[a];
      }
    }
    __syncthreads();
    if (has_row) {
      while (ip < ie) {
        const uint32_t it = d.inc[ip];
        const int slot = int(it & 0x1FFFu);
        if (slot >= c0 + nc) break;
        const int sl = slot - c0;
        const int a = int((it >> 13) & 3u);
        if (WANT_G) gacc += sG[a * chunk + sl];
        if (WANT_H) {
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            const int pos = int((it >> (15 + 5 * b)) & 31u);
            if (pos != 31) sVal[roff + pos] += sH[sym3(a, b) * chunk + sl];
          }
        }
        ++ip;
      }
    }
    __syncthreads();
  }
  if (WANT_H)
    for (int k = tid; k < nslice; k += blockDim.x) vals[k0 + k] = HT(sVal[k]);
  if (WANT_G && has_row) g[r] = f ? gacc - f[r] : gacc;
}

// dense per-element outputs (tests, autograd boundary)
__global__ __launch_bounds__(kBlock) void p1_element_kernel(fne_plan_desc d, const double* __restrict__ x,
                                                            const double* __restrict__ dvals,
                                                            double* __restrict__ e_out,
                                                            double* __restrict__ ge_out,
                                                            double* __restrict__ he_out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.n_elem) return;
  P1Elem el;
  p1_load(d, e, x, dvals, el);
  const int ne = d.n_elem;
  if (e_out) e_out[e] = p1_energy(d.kind, el, d.param, d.qdata, d.nq);
  if (ge_out || he_out) {
    double ge[3], he[6];
    p1_grad_hess(d.kind, el, d.param, d.qdata, d.nq, ge_out != nullptr, he_out != nullptr, ge, he);
    if (ge_out)
      for (int a = 0; a < 3; ++a) ge_out[a * ne + e] = ge[a];
    if (he_out)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) he_out[(a * 3 + b) * ne + e] = he[sym3(a, b)];
  }
}

}  // namespace

int fne_launch_p1(fne_plan* plan, FneOp op, const FneCall& c) {
  const fne_plan_desc& d = plan->d;
  if (d.kind == FNE_KIND_GINZBURG_LANDAU && (d.nq <= 0 || d.nq > kMaxQ || d.qdata == nullptr)) {
    fne_set_error("Ginzburg-Landau plan needs 1..16 quadrature points in qdata");
    return FNE_EINVAL;
  }
  if (op == kOpEnergy) {
    const int nb = plan->n_partials;
    hipLaunchKernelGGL(p1_energy_kernel, dim3(nb), dim3(kBlock), 0, c.stream, d, c.x, c.dvals, c.f,
                       plan->partials);
    FNE_LAUNCH_CHECK();
    return fne_finalize_energy(plan, nb, c.out, c.stream);
  }
  if (op == kOpElementEval) {
    const int nb = (d.n_elem + kBlock - 1) / kBlock;
    if (nb > 0) {
      hipLaunchKernelGGL(p1_element_kernel, dim3(nb), dim3(kBlock), 0, c.stream, d, c.x, c.dvals,
                         c.e_out, c.ge_out, c.he_out);
      FNE_LAUNCH_CHECK();
    }
    return FNE_OK;
  }
  if (d.n_patch <= 0) return FNE_OK;
  const bool want_h = (op == kOpHessian);
  const bool want_g = (op == kOpGradient) || (c.g != nullptr);
  size_t shm = 0;
  if (want_h) shm += sizeof(double) * (size_t(d.max_patch_nnz) + 6 * size_t(d.chunk));
  if (want_g) shm += sizeof(double) * 3 * size_t(d.chunk);
  if (shm > plan->max_lds) {
    fne_set_error("patch LDS request " + std::to_string(shm) + " B exceeds the device limit " +
                  std::to_string(plan->max_lds) + " B; rebuild the plan with a smaller chunk");
    return FNE_EINVAL;
  }
  if (d.max_row_len > 30) {
    fne_set_error("P1 patch kernel supports CSR rows of at most 30 entries");
    return FNE_EINVAL;
  }
  const dim3 grid(d.n_patch), block(kBlock);
  if (!want_h) {
    hipLaunchKernelGGL((p1_patch_kernel<double, true, false>), grid, block, shm, c.stream, d, c.x,
                       c.dvals, c.f, c.g, (double*)nullptr);
  } else if (c.precision == FNE_FP32) {
    if (want_g)
      hipLaunchKernelGGL((p1_patch_kernel<float, true, true>), grid, block, shm, c.stream, d, c.x,
                         c.dvals, c.f, c.g, (float*)c.vals);
    else
      hipLaunchKernelGGL((p1_patch_kernel<float, false, true>), grid, block, shm, c.stream, d, c.x,
                         c.dvals, c.f, c.g, (float*)c.vals);
  } else {
    if (want_g)
      hipLaunchKernelGGL((p1_patch_kernel<double, true, true>), grid, block, shm, c.stream, d, c.x,
                         c.dvals, c.f, c.g, (double*)c.vals);
    else
      hipLaunchKernelGGL((p1_patch_kernel<double, false, true>), grid, block, shm, c.stream, d, c.x,
                         c.dvals, c.f, c.g, (double*)c.vals);
  }
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}
