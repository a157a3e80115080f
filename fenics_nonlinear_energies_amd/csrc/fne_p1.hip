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
// compile-time families: the p-Laplace exponent gets exact specialisations for p = 3 and p = 2
constexpr int kFamPL = FNE_KIND_PLAPLACE;
constexpr int kFamGL = FNE_KIND_GINZBURG_LANDAU;
constexpr int kFamPL3 = 101;
constexpr int kFamPL2 = 102;
// Ginzburg-Landau with the 3-point rule of every GL mesh (nq == 3 at compile time): one pass for
// energy, gradient and Hessian with explicit FMAs and the products w_q ip[a][q] ip[b][q] read
// from the quadrature table
constexpr int kFamGL3 = 103;
template <int fam>
constexpr bool is_plaplace() {
  return fam == kFamPL || fam == kFamPL3 || fam == kFamPL2;
}
template <int fam>
constexpr bool is_gl() {
  return fam == kFamGL || fam == kFamGL3;
}

struct P1Elem {
  double v[3];
  double dvx[3], dvy[3], vol;
};

// Geometry of a triangle from its vertex coordinates (FNE_FLAG_NODE_GEO): 16 bytes per vertex,
// gathered like the nodal values and shared by the vertex's elements in cache, instead of 56
// bytes per element.  Each operation is rounded on its own (the library is built with
// -ffp-contract=off; FP64 division is correctly rounded) and in the order of the host's
// triangle operators, so dvx, dvy and vol equal the host-computed arrays bit for bit.
__device__ __forceinline__ void p1_node_geometry(const double2* __restrict__ xy, const int c[3], P1Elem& el) {
  const double2 p0 = xy[c[0]], p1 = xy[c[1]], p2 = xy[c[2]];
  const double det = (p1.x - p0.x) * (p2.y - p0.y) - (p2.x - p0.x) * (p1.y - p0.y);
  el.dvx[0] = (p1.y - p2.y) / det;
  el.dvx[1] = (p2.y - p0.y) / det;
  el.dvx[2] = (p0.y - p1.y) / det;
  el.dvy[0] = (p2.x - p1.x) / det;
  el.dvy[1] = (p0.x - p2.x) / det;
  el.dvy[2] = (p1.x - p0.x) / det;
  el.vol = 0.5 * fabs(det);
}

// element e of SoA arrays with stride ne: codes + geometry (NG: from the vertex coordinates xy)
template <bool NG>
__device__ __forceinline__ void p1_load_raw(const int32_t* __restrict__ edof, const double* __restrict__ egeo,
                                            const double2* __restrict__ xy, long ne, long e, int c[3],
                                            P1Elem& el) {
#pragma unroll
  for (int a = 0; a < 3; ++a) c[a] = edof[a * ne + e];
  if constexpr (NG) {
    p1_node_geometry(xy, c, el);
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      el.dvx[a] = egeo[a * ne + e];
      el.dvy[a] = egeo[(3 + a) * ne + e];
    }
    el.vol = egeo[6 * ne + e];
  }
}

__device__ __forceinline__ const double2* p1_xy(const fne_plan_desc& d) {
  return reinterpret_cast<const double2*>(d.dof_xy);
}

// the plan's geometry form chosen at run time (uniform over the launch)
__device__ __forceinline__ void p1_load_raw_any(const fne_plan_desc& d, const int32_t* __restrict__ edof,
                                                const double* __restrict__ egeo, long ne, long e, int c[3],
                                                P1Elem& el) {
  if (d.flags & FNE_FLAG_NODE_GEO)
    p1_load_raw<true>(edof, egeo, p1_xy(d), ne, e, c, el);
  else
    p1_load_raw<false>(edof, egeo, p1_xy(d), ne, e, c, el);
}

// element e of the plan's element arrays, then the gathered nodal values
__device__ __forceinline__ void p1_load(const fne_plan_desc& d, int e, const double* __restrict__ x,
                                        const double* __restrict__ dvals, P1Elem& el) {
  int c[3];
  p1_load_raw_any(d, d.edof, d.egeo, d.n_elem, e, c, el);
#pragma unroll
  for (int a = 0; a < 3; ++a) el.v[a] = gather_dof(c[a], d.n_free, x, dvals);
}

// s^(p/2) and friends; p==2 and p==3 take exact fast paths
__device__ __forceinline__ double pl_pow_half(double s, double p) {
  if (p == 3.0) return s * sqrt(s);
  if (p == 2.0) return s;
  return pow(s, 0.5 * p);
}

// Ginzburg-Landau, 3 quadrature points: energy E, gradient ge and Hessian he (sym3 order) in
// one pass.  Table layout (10 nq values): ip[a][q] at 3a+q, w[q] at 9+q, then the products
// W[k][q] = w_q ip[a_k][q] ip[b_k][q] at 12+3k+q for k = 00 11 22 01 02 12.  The same formulas
// as p1_energy / p1_grad_hess, evaluated with fused multiply-adds:
//   u_q = sum_a v_a ip[a][q],  t_q = u_q^2 - 1
//   E = vol (eps/2 s + 1/4 sum_q w_q t_q^2)
//   ge_a = vol (eps t_a + sum_q (w_q t_q u_q) ip[a][q])
//   he_k = vol (eps K_k + sum_q (3 u_q^2 - 1) W[k][q])
__device__ __forceinline__ void gl3_eval(const P1Elem& el, double eps, const double* __restrict__ qd, double& E,
                                         double ge[3], double he[6]) {
  const double gx = fma(el.v[2], el.dvx[2], fma(el.v[1], el.dvx[1], el.v[0] * el.dvx[0]));
  const double gy = fma(el.v[2], el.dvy[2], fma(el.v[1], el.dvy[1], el.v[0] * el.dvy[0]));
  const double s = fma(gy, gy, gx * gx);
  double pot = 0.0, gq[3] = {0.0, 0.0, 0.0}, hq[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double uq = fma(el.v[2], qd[6 + q], fma(el.v[1], qd[3 + q], el.v[0] * qd[q]));
    const double u2 = uq * uq;
    const double tq = u2 - 1.0;
    const double wt = qd[9 + q] * tq;
    pot = fma(wt, tq, pot);
    const double gc = wt * uq;
#pragma unroll
    for (int a = 0; a < 3; ++a) gq[a] = fma(gc, qd[3 * a + q], gq[a]);
    const double hc = fma(3.0, u2, -1.0);
#pragma unroll
    for (int k = 0; k < 6; ++k) hq[k] = fma(hc, qd[12 + 3 * k + q], hq[k]);
  }
  E = el.vol * fma(0.25, pot, (0.5 * eps) * s);
  const int ia[6] = {0, 1, 2, 0, 0, 1};
  const int ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int a = 0; a < 3; ++a) ge[a] = el.vol * fma(eps, fma(gy, el.dvy[a], gx * el.dvx[a]), gq[a]);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int a = ia[k], b = ib[k];
    he[k] = el.vol * fma(eps, fma(el.dvy[a], el.dvy[b], el.dvx[a] * el.dvx[b]), hq[k]);
  }
}

template <int fam>
__device__ __forceinline__ double p1_energy(const P1Elem& el, const double* prm, const double* __restrict__ qd,
                                            int nq) {
  double gx = 0.0, gy = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    gx += el.v[a] * el.dvx[a];
    gy += el.v[a] * el.dvy[a];
  }
  const double s = gx * gx + gy * gy;
  if constexpr (fam == kFamGL3) {
    double E, ge[3], he[6];
    gl3_eval(el, prm[0], qd, E, ge, he);
    return E;
  }
  if constexpr (fam == kFamPL3) {
    return (1.0 / 3.0) * (s * sqrt(s)) * el.vol;
  } else if constexpr (fam == kFamPL2) {
    return (1.0 / 2.0) * s * el.vol;
  } else if constexpr (fam == kFamPL) {
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
template <int fam>
__device__ __forceinline__ void p1_grad_hess(const P1Elem& el, const double* prm, const double* __restrict__ qd,
                                             int nq, bool want_g, bool want_h, double ge[3], double he[6]) {
  if constexpr (fam == kFamGL3) {
    double E;
    gl3_eval(el, prm[0], qd, E, ge, he);
    return;
  }
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
  if constexpr (is_plaplace<fam>()) {
    const double p = fam == kFamPL3 ? 3.0 : (fam == kFamPL2 ? 2.0 : prm[0]);
    // c1 = s^(p/2-1), c2 = (p-2) s^(p/2-2)
    double c1, c2;
    if constexpr (fam == kFamPL2) {
      c1 = 1.0;
      c2 = 0.0;
    } else if constexpr (fam == kFamPL3) {
      const double r = sqrt(s);
      c1 = r;
      c2 = s > 0.0 ? 1.0 / r : 0.0;
    } else if (p == 2.0) {
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

// energy, gradient and Hessian of one element for the given family (GL3 in one pass)
template <int fam>
__device__ __forceinline__ void p1_eval(const P1Elem& el, const double* prm, const double* __restrict__ qd, int nq,
                                        bool want_e, bool want_g, bool want_h, double& E, double ge[3],
                                        double he[6]) {
  if constexpr (fam == kFamGL3) {
    gl3_eval(el, prm[0], qd, E, ge, he);
  } else {
    if (want_e) E = p1_energy<fam>(el, prm, qd, nq);
    if (want_g || want_h) p1_grad_hess<fam>(el, prm, qd, nq, want_g, want_h, ge, he);
  }
}

__device__ __forceinline__ int sym3(int a, int b) {
  // 00->0 11->1 22->2 01->3 02->4 12->5
  return a == b ? a : (a + b + 2);
}

// ---------------------------------------------------------------------------------------------
// energy: grid-stride over elements (+ load term over free DOFs), compensated block partials
// ---------------------------------------------------------------------------------------------
template <int FAM>
__global__ __launch_bounds__(kBlock) void p1_energy_kernel(fne_plan_desc d, const double* __restrict__ x,
                                                           const double* __restrict__ dvals,
                                                           const double* __restrict__ f,
                                                           double2* __restrict__ partial) {
  // block b covers elements [b*2B, (b+1)*2B) and free DOFs [b*2B, (b+1)*2B) of the load term
  double s = 0.0, c = 0.0;
  const double* prm = d.param;
  const int base = blockIdx.x * 2 * blockDim.x + threadIdx.x;
  const int e1 = base, e2 = base + blockDim.x;
  P1Elem el, el2;
  if (e1 < d.n_elem) p1_load(d, e1, x, dvals, el);
  if (e2 < d.n_elem) p1_load(d, e2, x, dvals, el2);
  if (e1 < d.n_elem) {
    double E = p1_energy<FAM>(el, prm, d.qdata, d.nq);
    if (d.eweight) E *= d.eweight[e1];
    acc_add(s, c, E);
  }
  if (e2 < d.n_elem) {
    double E = p1_energy<FAM>(el2, prm, d.qdata, d.nq);
    if (d.eweight) E *= d.eweight[e2];
    acc_add(s, c, E);
  }
  if (f) {
    if (e1 < d.n_free) acc_add(s, c, -(f[e1] * x[e1]));
    if (e2 < d.n_free) acc_add(s, c, -(f[e2] * x[e2]));
  }
  block_reduce_pair(s, c);
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(s, c);
}

// ---------------------------------------------------------------------------------------------
// Colored patch assembly (FNE_FLAG_COLORED).  One workgroup per patch of <= blockDim consecutive
// rows; the patch's elements (owned + halo, stored patch-major) are evaluated once, EPT per
// thread, into registers.  Then, color by color, each element adds its contributions to the
// patch's OWNED rows directly into the LDS copy of the CSR slice (and gradient slice).  Elements
// of one color share no owned row, so the adds never race; the order (color, then slot) is
// fixed, so the result is bitwise reproducible.  No element staging, no incidence lists.
// item (uint64): pos_ab[5(3a+b) : +5) (31 = none) | energy owner [48] | color [52:60)
// ---------------------------------------------------------------------------------------------
template <int FAM, typename HT, bool WANT_E, bool WANT_G, bool WANT_H, int EPT>
__global__ __launch_bounds__(kBlock) void p1_color_kernel(fne_plan_desc d, const double* __restrict__ x,
                                                          const double* __restrict__ dvals,
                                                          const double* __restrict__ f,
                                                          double* __restrict__ g, HT* __restrict__ vals,
                                                          double2* __restrict__ partial) {
  extern __shared__ double smem[];
  const int p = xcd_remap(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int B = blockDim.x;
  const int r0 = d.prow[p], r1 = d.prow[p + 1];
  const int nrows = r1 - r0;
  const int k0 = d.indptr[r0];
  const int nslice = WANT_H ? d.indptr[r1] - k0 : 0;
  const long eb = d.pelem_ptr[p];
  const int ne = d.pelem_ptr[p + 1] - int(eb);
  const int ncol = d.pcolors[p];

  double* sVal = smem;                                 // [max_patch_nnz]
  double* sG = sVal + (WANT_H ? d.max_patch_nnz : 0);  // [B]
  int* sPtr = reinterpret_cast<int*>(sG + (WANT_G ? B : 0));  // [B + 1]
  if (WANT_H) {
    for (int k = tid; k < nslice; k += B) sVal[k] = 0.0;
    for (int i = tid; i <= nrows; i += B) sPtr[i] = d.indptr[r0 + i] - k0;
  }
  if (WANT_G)
    for (int i = tid; i < nrows; i += B) sG[i] = 0.0;

  double es = 0.0, ec = 0.0;
  if (WANT_E && f)
    for (int i = tid; i < nrows; i += B) acc_add(es, ec, -(f[r0 + i] * x[r0 + i]));
  const double* prm = d.param;
  P1Elem el[EPT];
  int code[EPT][3];
  uint64_t item[EPT];
  bool act[EPT];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int sl = tid + k * B;
    act[k] = sl < ne;
    if (act[k]) {
      p1_load_raw_any(d, d.pedof, d.pgeo, d.n_pe, eb + sl, code[k], el[k]);
      item[k] = d.pitem[eb + sl];
    }
  }
  double ge[EPT][3], he[EPT][6];
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    if (!act[k]) continue;
#pragma unroll
    for (int a = 0; a < 3; ++a) el[k].v[a] = gather_dof(code[k][a], d.n_free, x, dvals);
    if (WANT_E && ((item[k] >> 48) & 1ull)) acc_add(es, ec, p1_energy<FAM>(el[k], prm, d.qdata, d.nq));
    if (WANT_G || WANT_H) p1_grad_hess<FAM>(el[k], prm, d.qdata, d.nq, WANT_G, WANT_H, ge[k], he[k]);
  }
  __syncthreads();
  if (WANT_G || WANT_H) {
    for (int c = 0; c < ncol; ++c) {
#pragma unroll
      for (int k = 0; k < EPT; ++k) {
        if (!act[k] || int((item[k] >> 52) & 0xFFull) != c) continue;
        // the (up to) 9 targets of one element are distinct LDS words: issue every read, then
        // every write, instead of 9 dependent read-add-write round trips
        int rl[3], addr[9];
        double oldg[3], oldh[9];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          rl[a] = code[k][a] - r0;
          const bool own = rl[a] >= 0 && rl[a] < nrows;  // else: another patch's (or constrained) row
          if (!own) rl[a] = -1;
          if (WANT_H) {
            const int off = own ? sPtr[rl[a]] : 0;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
              const int pos = int((item[k] >> (5 * (3 * a + b))) & 31ull);
              addr[3 * a + b] = (own && pos != 31) ? off + pos : -1;
            }
          }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
          if (WANT_G && rl[a] >= 0) oldg[a] = sG[rl[a]];
        if (WANT_H) {
#pragma unroll
          for (int q = 0; q < 9; ++q)
            if (addr[q] >= 0) oldh[q] = sVal[addr[q]];
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
          if (WANT_G && rl[a] >= 0) sG[rl[a]] = oldg[a] + ge[k][a];
        if (WANT_H) {
#pragma unroll
          for (int q = 0; q < 9; ++q)
            if (addr[q] >= 0) sVal[addr[q]] = oldh[q] + he[k][sym3(q / 3, q % 3)];
        }
      }
      __syncthreads();
    }
  }
  if (WANT_H)
    for (int k = tid; k < nslice; k += B) vals[k0 + k] = HT(sVal[k]);
  if (WANT_G)
    for (int i = tid; i < nrows; i += B) g[r0 + i] = f ? sG[i] - f[r0 + i] : sG[i];
  if (WANT_E) {
    block_reduce_pair(es, ec);
    if (tid == 0) partial[p] = make_double2(es, ec);
  }
}

// ---------------------------------------------------------------------------------------------
// Staged patch assembly (FNE_FLAG_STAGED).  Phase A: the patch's elements (patch-major, owned
// first) are evaluated one per thread per iteration; their 6 Hessian values (type HT) and 3
// gradient values (FP64) are staged in LDS ([value][slot], conflict-free writes).  Phase B:
// thread t owns row r0+t and walks its incidence list (prefetched before phase A); the three
// columns of one incidence are distinct, so their read-modify-writes into the row's part of an
// LDS copy of the CSR slice are issued together.  Every slice word is touched by one thread in a
// fixed order (bitwise reproducible, no atomics); the slice then leaves as coalesced stores.
// The fused energy is the owned elements' energies plus the load term -f.x of the patch rows.
// HT = float accumulates in FP32 (FP32 Hessian mode), HT = double in FP64.
// ---------------------------------------------------------------------------------------------
constexpr int kStagedPrefetch = 8;

template <int FAM, typename HT, bool WANT_E, bool WANT_G, bool WANT_H, bool NG>
__global__ __launch_bounds__(kBlock) void p1_staged_kernel(fne_plan_desc d, const double* __restrict__ x,
                                                           const double* __restrict__ dvals,
                                                           const double* __restrict__ f,
                                                           double* __restrict__ g, HT* __restrict__ vals,
                                                           double2* __restrict__ partial) {
  extern __shared__ double smem[];
  const int p = xcd_remap(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int B = blockDim.x;
  const int cap = d.chunk;
  const int r0 = d.prow[p], r1 = d.prow[p + 1];
  const long eb = d.pelem_ptr[p];
  const int ne = d.pelem_ptr[p + 1] - int(eb);
  const int nown = WANT_E ? d.pown[p] : 0;
  const int k0 = WANT_H ? d.indptr[r0] : 0;
  const int nslice = WANT_H ? d.indptr[r1] - k0 : 0;
  // LDS: [gradient FP64: 3*cap] [slice HT: max_patch_nnz (8-aligned)] [Hessian HT: 6*cap]
  double* sG = smem;
  HT* sVal = reinterpret_cast<HT*>(smem + (WANT_G ? 3 * cap : 0));
  HT* sH = sVal + (WANT_H ? ((d.max_patch_nnz + 1) & ~1) : 0);

  const int r = r0 + tid;
  const bool has_row = r < r1;
  const int ib = has_row ? d.inc_ptr[r] : 0;
  const int cnt = has_row ? d.inc_ptr[r + 1] - ib : 0;
  const int roff = (WANT_H && has_row) ? d.indptr[r] - k0 : 0;
  uint32_t items[kStagedPrefetch];
#pragma unroll
  for (int k = 0; k < kStagedPrefetch; ++k) items[k] = (k < cnt) ? d.inc[ib + k] : 0u;
  if (WANT_H)
    for (int k = tid; k < nslice; k += B) sVal[k] = HT(0);

  double es = 0.0, ec = 0.0;
  if (WANT_E && f && has_row) acc_add(es, ec, -(f[r] * x[r]));
  const double* prm = d.param;
#pragma unroll 1
  for (int sl = tid; sl < ne; sl += B) {
    P1Elem el;
    int code[3];
    p1_load_raw<NG>(d.pedof, d.pgeo, p1_xy(d), d.n_pe, eb + sl, code, el);
#pragma unroll
    for (int a = 0; a < 3; ++a) el.v[a] = gather_dof(code[a], d.n_free, x, dvals);
    double E = 0.0, ge[3], he[6];
    p1_eval<FAM>(el, prm, d.qdata, d.nq, WANT_E && sl < nown, WANT_G, WANT_H, E, ge, he);
    if (WANT_E && sl < nown) acc_add(es, ec, E);
    if (WANT_G || WANT_H) {
      if (WANT_H) {
#pragma unroll
        for (int q = 0; q < 6; ++q) sH[q * cap + sl] = HT(he[q]);
      }
      if (WANT_G) {
#pragma unroll
        for (int a = 0; a < 3; ++a) sG[a * cap + sl] = ge[a];
      }
    }
  }
  if (WANT_G || WANT_H) {
    __syncthreads();
    if (has_row) {
      double gacc = 0.0;
      // the row's own column (b == a: its diagonal entry, hit by every incidence) accumulates in
      // a register in the same order; the two other columns go through the LDS row
      HT diag = HT(0);
      int pdiag = 31;
      auto consume = [&](uint32_t it) {
        const int s = int(it & 0x1FFFu);
        const int a = int((it >> 13) & 3u);
        if (WANT_G) gacc += sG[a * cap + s];
        if (WANT_H) {
          int pos[3];
          HT h[3], old[3];
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            pos[b] = int((it >> (15 + 5 * b)) & 31u);
            h[b] = sH[sym3(a, b) * cap + s];
          }
          diag += a == 0 ? h[0] : (a == 1 ? h[1] : h[2]);
          pdiag = a == 0 ? pos[0] : (a == 1 ? pos[1] : pos[2]);
#pragma unroll
          for (int b = 0; b < 3; ++b)
            if (b != a && pos[b] != 31) old[b] = sVal[roff + pos[b]];
#pragma unroll
          for (int b = 0; b < 3; ++b)
            if (b != a && pos[b] != 31) sVal[roff + pos[b]] = old[b] + h[b];
        }
      };
#pragma unroll
      for (int k = 0; k < kStagedPrefetch; ++k)
        if (k < cnt) consume(items[k]);
      for (int k = kStagedPrefetch; k < cnt; ++k) consume(d.inc[ib + k]);
      if (WANT_H && pdiag != 31) sVal[roff + pdiag] = diag;
      if (WANT_G) g[r] = f ? gacc - f[r] : gacc;
    }
    if (WANT_H) {
      __syncthreads();
      for (int k = tid; k < nslice; k += B) vals[k0 + k] = sVal[k];
    }
  }
  if (WANT_E) {
    block_reduce_pair(es, ec);
    if (tid == 0) partial[p] = make_double2(es, ec);
  }
}

// ---------------------------------------------------------------------------------------------
// Row-gather assembly (FNE_FLAG_ROW_GATHER): one thread per CSR row, no barriers, no staging.
// Each incident element is re-evaluated by each of its vertex rows (P1 element math is cheap;
// the re-reads hit L1/L2 because neighbouring rows share elements).  Row accumulators live in
// LDS, column-major per thread ([pos][thread]) so that dynamic positions stay conflict-free.
// item (uint64): e[0:32) | a[32:34) | pos_b[40+8b : 48+8b)  (255 = constrained column)
// ---------------------------------------------------------------------------------------------
template <int FAM, typename HT, bool WANT_G, bool WANT_H>
__global__ __launch_bounds__(kBlock) void p1_rowgather_kernel(fne_plan_desc d, const double* __restrict__ x,
                                                              const double* __restrict__ dvals,
                                                              const double* __restrict__ f,
                                                              double* __restrict__ g, HT* __restrict__ vals) {
  extern __shared__ double sAcc[];  // [max_row_len][blockDim]
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int tid = threadIdx.x;
  const int B = blockDim.x;
  const bool has_row = r < d.n_free;
  const int kb = has_row ? d.indptr[r] : 0;
  const int len = has_row ? d.indptr[r + 1] - kb : 0;
  if (WANT_H)
    for (int j = 0; j < len; ++j) sAcc[j * B + tid] = 0.0;
  if (!has_row) return;
  const uint64_t* inc = reinterpret_cast<const uint64_t*>(d.inc);
  const int ib = d.inc_ptr[r], ie = d.inc_ptr[r + 1];
  double gacc = 0.0;
  const double* prm = d.param;
  for (int it = ib; it < ie; ++it) {
    const uint64_t item = inc[it];
    const int e = int(item & 0xFFFFFFFFull);
    const int a = int((item >> 32) & 3ull);
    P1Elem el;
    p1_load(d, e, x, dvals, el);
    double ge[3], he[6];
    p1_grad_hess<FAM>(el, prm, d.qdata, d.nq, WANT_G, WANT_H, ge, he);
    if (WANT_G) gacc += ge[a];
    if (WANT_H) {
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const int pos = int((item >> (40 + 8 * b)) & 0xFFull);
        if (pos != 255) sAcc[pos * B + tid] += he[sym3(a, b)];
      }
    }
  }
  if (WANT_H)
    for (int j = 0; j < len; ++j) vals[kb + j] = HT(sAcc[j * B + tid]);
  if (WANT_G) g[r] = f ? gacc - f[r] : gacc;
}

// Hessian-vector product y = H(x) v over the staged patches: phase A stages every patch
// element's H_e v_e (element Hessian at x, v taken as zero on constrained DOFs) in LDS, phase B
// sums each row's incidence list in its fixed order -- the gradient path of p1_staged_kernel
// with the element gradient replaced by H_e v_e.  Deterministic, no atomics.
template <int FAM>
__global__ __launch_bounds__(kBlock) void p1_staged_hvp_kernel(fne_plan_desc d, const double* __restrict__ x,
                                                               const double* __restrict__ dvals,
                                                               const double* __restrict__ v,
                                                               double* __restrict__ y) {
  extern __shared__ double smem[];
  const int p = xcd_remap(blockIdx.x, gridDim.x);
  const int tid = threadIdx.x;
  const int B = blockDim.x;
  const int cap = d.chunk;
  const int r0 = d.prow[p], r1 = d.prow[p + 1];
  const long eb = d.pelem_ptr[p];
  const int ne = d.pelem_ptr[p + 1] - int(eb);
  double* sG = smem;  // [3][cap]
  for (int sl = tid; sl < ne; sl += B) {
    P1Elem el;
    int code[3];
    p1_load_raw_any(d, d.pedof, d.pgeo, d.n_pe, eb + sl, code, el);
    double ve[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      el.v[a] = gather_dof(code[a], d.n_free, x, dvals);
      ve[a] = code[a] < d.n_free ? v[code[a]] : 0.0;
    }
    double ge[3], he[6];
    p1_grad_hess<FAM>(el, d.param, d.qdata, d.nq, false, true, ge, he);
#pragma unroll
    for (int a = 0; a < 3; ++a)
      sG[a * cap + sl] = he[sym3(a, 0)] * ve[0] + he[sym3(a, 1)] * ve[1] + he[sym3(a, 2)] * ve[2];
  }
  __syncthreads();
  const int r = r0 + tid;
  if (r < r1) {
    const int ib = d.inc_ptr[r], ie = d.inc_ptr[r + 1];
    double acc = 0.0;
    for (int k = ib; k < ie; ++k) {
      const uint32_t it = d.inc[k];
      acc += sG[int((it >> 13) & 3u) * cap + int(it & 0x1FFFu)];
    }
    y[r] = acc;
  }
}

// dense per-element outputs (tests, autograd boundary)
template <int FAM>
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
  if (e_out) e_out[e] = p1_energy<FAM>(el, d.param, d.qdata, d.nq);
  if (ge_out || he_out) {
    double ge[3], he[6];
    p1_grad_hess<FAM>(el, d.param, d.qdata, d.nq, ge_out != nullptr, he_out != nullptr, ge, he);
    if (ge_out)
      for (int a = 0; a < 3; ++a) ge_out[a * ne + e] = ge[a];
    if (he_out)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) he_out[(a * 3 + b) * ne + e] = he[sym3(a, b)];
  }
}

}  // namespace

template <int FAM>
int launch_p1(fne_plan* plan, FneOp op, const FneCall& c) {
  const fne_plan_desc& d = plan->d;
  if (is_gl<FAM>() && (d.nq <= 0 || d.nq > kMaxQ || d.qdata == nullptr)) {
    fne_set_error("Ginzburg-Landau plan needs 1..16 quadrature points in qdata");
    return FNE_EINVAL;
  }
  if (op == kOpEnergy) {
    const int items = d.n_elem > d.n_free ? d.n_elem : d.n_free;
    int nb = (items + 2 * kBlock - 1) / (2 * kBlock);
    if (nb < 1) nb = 1;
    if (nb > plan->n_partials) {
      fne_set_error("energy reduction scratch too small");
      return FNE_EINVAL;
    }
    hipLaunchKernelGGL(p1_energy_kernel<FAM>, dim3(nb), dim3(kBlock), 0, c.stream, d, c.x, c.dvals, c.f,
                       plan->partials);
    FNE_LAUNCH_CHECK();
    return fne_finalize_energy(plan, nb, c.out, c.stream);
  }
  if (op == kOpElementEval) {
    const int nb = (d.n_elem + kBlock - 1) / kBlock;
    if (nb > 0) {
      hipLaunchKernelGGL(p1_element_kernel<FAM>, dim3(nb), dim3(kBlock), 0, c.stream, d, c.x, c.dvals,
                         c.e_out, c.ge_out, c.he_out);
      FNE_LAUNCH_CHECK();
    }
    return FNE_OK;
  }
  const bool want_h = (op == kOpHessian);
  const bool want_g = (op == kOpGradient) || (c.g != nullptr);
  if (d.flags & FNE_FLAG_ROW_GATHER) {
    if (op == kOpFused) {
      if (c.out) {
        const int rc = launch_p1<FAM>(plan, kOpEnergy, c);
        if (rc) return rc;
      }
      return launch_p1<FAM>(plan, kOpHessian, c);
    }
    if (d.n_free <= 0) return FNE_OK;
    const size_t shm_rg = want_h ? sizeof(double) * size_t(d.max_row_len) * kBlock : 0;
    if (shm_rg > plan->max_lds) {
      fne_set_error("row-gather LDS request exceeds the device limit");
      return FNE_EINVAL;
    }
    const dim3 grid((d.n_free + kBlock - 1) / kBlock), block(kBlock);
    if (!want_h)
      hipLaunchKernelGGL((p1_rowgather_kernel<FAM, double, true, false>), grid, block, 0, c.stream, d, c.x, c.dvals,
                         c.f, c.g, (double*)nullptr);
    else if (c.precision == FNE_FP32) {
      if (want_g)
        hipLaunchKernelGGL((p1_rowgather_kernel<FAM, float, true, true>), grid, block, shm_rg, c.stream, d, c.x,
                           c.dvals, c.f, c.g, (float*)c.vals);
      else
        hipLaunchKernelGGL((p1_rowgather_kernel<FAM, float, false, true>), grid, block, shm_rg, c.stream, d, c.x,
                           c.dvals, c.f, c.g, (float*)c.vals);
    } else {
      if (want_g)
        hipLaunchKernelGGL((p1_rowgather_kernel<FAM, double, true, true>), grid, block, shm_rg, c.stream, d, c.x,
                           c.dvals, c.f, c.g, (double*)c.vals);
      else
        hipLaunchKernelGGL((p1_rowgather_kernel<FAM, double, false, true>), grid, block, shm_rg, c.stream, d, c.x,
                           c.dvals, c.f, c.g, (double*)c.vals);
    }
    FNE_LAUNCH_CHECK();
    return FNE_OK;
  }
  if (d.n_patch <= 0) {
    if (op == kOpFused && c.out) return launch_p1<FAM>(plan, kOpEnergy, c);
    return FNE_OK;
  }
  const bool want_e = (op == kOpFused) && (c.out != nullptr);
  if (d.flags & FNE_FLAG_STAGED) {
    if (op == kOpFused && c.vals == nullptr) {
      fne_set_error("fused assembly needs the CSR value buffer");
      return FNE_EINVAL;
    }
    const bool sh = want_h || op == kOpFused;
    const bool sg = want_g || (op == kOpFused && c.g != nullptr);
    const bool ng = (d.flags & FNE_FLAG_NODE_GEO) != 0;
    if (!d.pedof || (!ng && !d.pgeo) || !d.pown || !d.inc || !d.inc_ptr) {
      fne_set_error("staged patch arrays missing from the plan");
      return FNE_EINVAL;
    }
    const size_t hsz = c.precision == FNE_FP32 ? 4 : 8;
    size_t shm = 0;
    if (sg) shm += sizeof(double) * 3 * size_t(d.chunk);
    if (sh) shm += hsz * (size_t((d.max_patch_nnz + 1) & ~1) + 6 * size_t(d.chunk));
    if (shm > plan->max_lds) {
      fne_set_error("staged patch LDS request exceeds the device limit");
      return FNE_EINVAL;
    }
    if (d.max_row_len > 30) {
      fne_set_error("P1 staged kernel supports CSR rows of at most 30 entries");
      return FNE_EINVAL;
    }
    const dim3 grid(d.n_patch), block(kBlock);
    double2* part = plan->partials;
#define ST_GO(HT, E, G, H, NG)                                                                          \
  hipLaunchKernelGGL((p1_staged_kernel<FAM, HT, E, G, H, NG>), grid, block, shm, c.stream, d, c.x, c.dvals, c.f, \
                     c.g, (HT*)c.vals, part)
#define ST_MR(HT, E, G, H) \
  do {                     \
    if (ng)                \
      ST_GO(HT, E, G, H, true); \
    else                   \
      ST_GO(HT, E, G, H, false); \
  } while (0)
    if (!sh) {
      ST_MR(double, false, true, false);
    } else if (c.precision == FNE_FP32) {
      if (want_e && sg) ST_MR(float, true, true, true);
      else if (want_e) ST_MR(float, true, false, true);
      else if (sg) ST_MR(float, false, true, true);
      else ST_MR(float, false, false, true);
    } else {
      if (want_e && sg) ST_MR(double, true, true, true);
      else if (want_e) ST_MR(double, true, false, true);
      else if (sg) ST_MR(double, false, true, true);
      else ST_MR(double, false, false, true);
    }
#undef ST_MR
#undef ST_GO
    FNE_LAUNCH_CHECK();
    if (want_e) return fne_finalize_energy(plan, d.n_patch, c.out, c.stream);
    return FNE_OK;
  }
  if (!(d.flags & FNE_FLAG_COLORED)) {
    fne_set_error("P1 plan has none of FNE_FLAG_COLORED, FNE_FLAG_STAGED, FNE_FLAG_ROW_GATHER");
    return FNE_EINVAL;
  }
  if (op == kOpFused && c.vals == nullptr) {
    fne_set_error("fused assembly needs the CSR value buffer");
    return FNE_EINVAL;
  }
  const bool want_hh = want_h || op == kOpFused;
  const bool want_gg = want_g || (op == kOpFused && c.g != nullptr);
  if (!d.pedof || (!(d.flags & FNE_FLAG_NODE_GEO) && !d.pgeo) || !d.pitem || !d.pcolors) {
    fne_set_error("colored patch arrays missing from the plan");
    return FNE_EINVAL;
  }
  const int ept = d.chunk;  // elements per thread (patch elements <= ept * blockDim)
  if (ept < 1 || ept > 3) {
    fne_set_error("colored P1 plans support 1..3 elements per thread");
    return FNE_EINVAL;
  }
  size_t shm = 0;
  if (want_hh) shm += sizeof(double) * size_t(d.max_patch_nnz) + sizeof(int) * (kBlock + 1);
  if (want_gg) shm += sizeof(double) * kBlock;
  if (shm > plan->max_lds) {
    fne_set_error("patch LDS request " + std::to_string(shm) + " B exceeds the device limit");
    return FNE_EINVAL;
  }
  const dim3 grid(d.n_patch), block(kBlock);
  double2* part = plan->partials;
#define P1_GO(HT, E, G, H, EP)                                                                          \
  hipLaunchKernelGGL((p1_color_kernel<FAM, HT, E, G, H, EP>), grid, block, shm, c.stream, d, c.x, c.dvals, c.f, \
                     c.g, (HT*)c.vals, part)
#define P1_EPT(HT, E, G, H)          \
  do {                               \
    if (ept == 1) P1_GO(HT, E, G, H, 1); \
    else if (ept == 2) P1_GO(HT, E, G, H, 2); \
    else P1_GO(HT, E, G, H, 3);      \
  } while (0)
  if (!want_hh) {
    P1_EPT(double, false, true, false);
  } else if (c.precision == FNE_FP32) {
    if (want_e && want_gg) P1_EPT(float, true, true, true);
    else if (want_e) P1_EPT(float, true, false, true);
    else if (want_gg) P1_EPT(float, false, true, true);
    else P1_EPT(float, false, false, true);
  } else {
    if (want_e && want_gg) P1_EPT(double, true, true, true);
    else if (want_e) P1_EPT(double, true, false, true);
    else if (want_gg) P1_EPT(double, false, true, true);
    else P1_EPT(double, false, false, true);
  }
#undef P1_EPT
#undef P1_GO
  FNE_LAUNCH_CHECK();
  if (want_e) return fne_finalize_energy(plan, d.n_patch, c.out, c.stream);
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

// Ginzburg-Landau plans with 3 quadrature points whose table carries the w ip ip products
// (FNE_FLAG_GL_PRODUCTS) take the one-pass GL3 kernels
bool gl3_table(const fne_plan_desc& d) {
  return d.nq == 3 && d.qdata != nullptr && (d.flags & FNE_FLAG_GL_PRODUCTS);
}

template <int FAM>
int launch_p1_hvp(fne_plan* plan, const double* x, const double* dvals, const double* v, double* y,
                  hipStream_t stream) {
  const fne_plan_desc& d = plan->d;
  const bool ng = (d.flags & FNE_FLAG_NODE_GEO) != 0;
  if (!(d.flags & FNE_FLAG_STAGED) || !d.pedof || (!ng && !d.pgeo) || !d.inc || !d.inc_ptr) {
    fne_set_error("P1 Hessian-vector products need a staged patch plan (FNE_FLAG_STAGED)");
    return FNE_EUNSUPPORTED;
  }
  if (is_gl<FAM>() && (d.nq <= 0 || d.nq > kMaxQ || d.qdata == nullptr)) {
    fne_set_error("Ginzburg-Landau plan needs 1..16 quadrature points in qdata");
    return FNE_EINVAL;
  }
  if (d.n_patch <= 0) return FNE_OK;
  const size_t shm = sizeof(double) * 3 * size_t(d.chunk);
  if (shm > plan->max_lds) {
    fne_set_error("staged patch LDS request exceeds the device limit");
    return FNE_EINVAL;
  }
  hipLaunchKernelGGL(p1_staged_hvp_kernel<FAM>, dim3(d.n_patch), dim3(kBlock), shm, stream, d, x, dvals, v, y);
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

int fne_hvp_p1(fne_plan* plan, const double* x, const double* dvals, const double* v, double* y, hipStream_t s) {
  if (plan->d.kind == FNE_KIND_PLAPLACE) {
    const double p = plan->d.param[0];
    if (p == 3.0) return launch_p1_hvp<kFamPL3>(plan, x, dvals, v, y, s);
    if (p == 2.0) return launch_p1_hvp<kFamPL2>(plan, x, dvals, v, y, s);
    return launch_p1_hvp<kFamPL>(plan, x, dvals, v, y, s);
  }
  if (gl3_table(plan->d)) return launch_p1_hvp<kFamGL3>(plan, x, dvals, v, y, s);
  return launch_p1_hvp<kFamGL>(plan, x, dvals, v, y, s);
}

int fne_launch_p1(fne_plan* plan, FneOp op, const FneCall& c) {
  if (plan->d.kind == FNE_KIND_PLAPLACE) {
    const double p = plan->d.param[0];
    if (p == 3.0) return launch_p1<kFamPL3>(plan, op, c);
    if (p == 2.0) return launch_p1<kFamPL2>(plan, op, c);
    return launch_p1<kFamPL>(plan, op, c);
  }
  if (gl3_table(plan->d)) return launch_p1<kFamGL3>(plan, op, c);
  return launch_p1<kFamGL>(plan, op, c);
}
