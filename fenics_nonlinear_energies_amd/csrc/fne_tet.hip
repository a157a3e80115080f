// Compressible Neo-Hookean P1 tetrahedra (A3): energy, gradient, CSR Hessian.
//
// Element data (SoA): edof[(3a+i)*n_elem+e] (deformed position y_{a,i}, a=0..3, i=x,y,z),
// egeo[k*n_elem+e]: dphix[0..3], dphiy[0..3], dphiz[0..3], vol.
// Math (hand-derived; replaces jax.hessian of _he_energy_density,
// hyperelasticity/jax_petsc/parallel_hessian_dof.py:17-52):
//   F_ij = sum_a y_{a,i} D_{a,j}   (D_a = grad phi_a),  J = det F  (|det F| with FNE_FLAG_ABS_DET)
//   W = C1 (|F|^2 - 3 - 2 ln J) + D1 (J-1)^2,   E = W vol
//   G = F^{-T},  q_a = G D_a
//   dE/dy_{a,i} = vol [ 2 C1 (F D_a)_i + c q_{a,i} ],            c = -2 C1 + 2 D1 J (J-1)
//   H_{(a,i),(b,k)} = vol [ 2 C1 d_ik (D_a.D_b) + alpha q_{b,i} q_{a,k} + beta q_{a,i} q_{b,k} ]
//                     alpha = 2 C1 - 2 D1 J (J-1),  beta = 2 D1 J (2J-1)
#include "fne_common.h"

namespace {

struct TetElem {
  double y[4][3];
  double D[4][3];
  double vol;
};

__device__ __forceinline__ void tet_load(const fne_plan_desc& d, int e, const double* __restrict__ x,
                                         const double* __restrict__ dvals, TetElem& t) {
  const int ne = d.n_elem;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int i = 0; i < 3; ++i) t.y[a][i] = gather_dof(d.edof[(3 * a + i) * ne + e], d.n_free, x, dvals);
    t.D[a][0] = d.egeo[a * ne + e];
    t.D[a][1] = d.egeo[(4 + a) * ne + e];
    t.D[a][2] = d.egeo[(8 + a) * ne + e];
  }
  t.vol = d.egeo[12 * ne + e];
}

struct TetKin {
  double F[3][3];
  double det, J;
};

__device__ __forceinline__ void tet_kinematics(const TetElem& t, bool absdet, TetKin& k) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      k.F[i][j] = t.y[0][i] * t.D[0][j] + t.y[1][i] * t.D[1][j] + t.y[2][i] * t.D[2][j] + t.y[3][i] * t.D[3][j];
  const double(&F)[3][3] = k.F;
  k.det = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) - F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
          F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
  k.J = absdet ? fabs(k.det) : k.det;
}

__device__ __forceinline__ double tet_energy(const TetElem& t, const double* prm, bool absdet) {
  TetKin k;
  tet_kinematics(t, absdet, k);
  double I1 = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) I1 += k.F[i][j] * k.F[i][j];
  const double C1 = prm[0], D1 = prm[1];
  const double W = C1 * (I1 - 3.0 - 2.0 * log(k.J)) + D1 * ((k.J - 1.0) * (k.J - 1.0));
  return W * t.vol;
}

// compact per-element state: the Hessian is expanded from (coefs, q, D) where it is needed
struct TetState {
  double s0, s1, s2;  // vol*2C1, vol*alpha, vol*beta
  double q[4][3];
  double g[4][3];
};

__device__ __forceinline__ void tet_state(const TetElem& t, const double* prm, bool absdet, TetState& s) {
  TetKin k;
  tet_kinematics(t, absdet, k);
  const double(&F)[3][3] = k.F;
  const double C1 = prm[0], D1 = prm[1];
  const double J = k.J;
  // G = F^{-T} = cof(F) / det
  const double inv = 1.0 / k.det;
  double G[3][3];
  G[0][0] = (F[1][1] * F[2][2] - F[1][2] * F[2][1]) * inv;
  G[0][1] = -(F[1][0] * F[2][2] - F[1][2] * F[2][0]) * inv;
  G[0][2] = (F[1][0] * F[2][1] - F[1][1] * F[2][0]) * inv;
  G[1][0] = -(F[0][1] * F[2][2] - F[0][2] * F[2][1]) * inv;
  G[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) * inv;
  G[1][2] = -(F[0][0] * F[2][1] - F[0][1] * F[2][0]) * inv;
  G[2][0] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) * inv;
  G[2][1] = -(F[0][0] * F[1][2] - F[0][2] * F[1][0]) * inv;
  G[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) * inv;
  const double c = -2.0 * C1 + 2.0 * D1 * J * (J - 1.0);
  s.s0 = t.vol * 2.0 * C1;
  s.s1 = t.vol * (2.0 * C1 - 2.0 * D1 * J * (J - 1.0));
  s.s2 = t.vol * (2.0 * D1 * J * (2.0 * J - 1.0));
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const double qa = G[i][0] * t.D[a][0] + G[i][1] * t.D[a][1] + G[i][2] * t.D[a][2];
      const double fa = F[i][0] * t.D[a][0] + F[i][1] * t.D[a][1] + F[i][2] * t.D[a][2];
      s.q[a][i] = qa;
      s.g[a][i] = t.vol * (2.0 * C1 * fa + c * qa);
    }
  }
}

// staged layout in LDS (FP64, [field][chunk]): 0..2 coefs, 3..14 q, 15..26 D, 27..38 g
constexpr int kTetStage = 39;

__global__ __launch_bounds__(kBlock) void tet_energy_kernel(fne_plan_desc d, const double* __restrict__ x,
                                                            const double* __restrict__ dvals,
                                                            const double* __restrict__ f,
                                                            double2* __restrict__ partial) {
  double s = 0.0, c = 0.0;
  const bool absdet = (d.flags & FNE_FLAG_ABS_DET) != 0;
  const int stride = gridDim.x * blockDim.x;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < d.n_elem; e += stride) {
    TetElem t;
    tet_load(d, e, x, dvals, t);
    double E = tet_energy(t, d.param, absdet);
    if (d.eweight) E *= d.eweight[e];
    acc_add(s, c, E);
  }
  if (f) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n_free; i += stride) acc_add(s, c, -(f[i] * x[i]));
  }
  block_reduce_pair(s, c);
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(s, c);
}

// Patch kernel.  Rows come in node triples; thread t of the workgroup owns scalar row r0+t.
// Incidence lists are per node row: item (uint64) = slot[0:16) | a[16:18) | bpos_b[24+8b : 32+8b)
// (bpos = block column of element node b inside the node's block row, 255 = constrained node).
template <typename HT, bool WANT_G, bool WANT_H>
__global__ __launch_bounds__(kBlock) void tet_patch_kernel(fne_plan_desc d, const double* __restrict__ x,
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
  const bool absdet = (d.flags & FNE_FLAG_ABS_DET) != 0;

  HT* sVal = reinterpret_cast<HT*>(smem);
  double* sS = smem + (WANT_H ? (size_t(d.max_patch_nnz) * sizeof(HT) + 7) / 8 : 0);  // [39][chunk]
  if (WANT_H)
    for (int k = tid; k < nslice; k += blockDim.x) sVal[k] = HT(0);

  const int r = r0 + tid;
  const bool has_row = r < r1;
  const int node_row = r / 3;
  const int comp = r - 3 * node_row;
  const uint64_t* inc = reinterpret_cast<const uint64_t*>(d.inc);
  int ip = has_row ? d.inc_ptr[node_row] : 0;
  const int ie = has_row ? d.inc_ptr[node_row + 1] : 0;
  const int roff = has_row ? d.indptr[r] - k0 : 0;
  double gacc = 0.0;
  __syncthreads();

  for (int c0 = 0; c0 < ne; c0 += chunk) {
    const int nc = min(chunk, ne - c0);
    for (int sl = tid; sl < nc; sl += blockDim.x) {
      const int e = d.pelem[eb + c0 + sl];
      TetElem t;
      tet_load(d, e, x, dvals, t);
      TetState st;
      tet_state(t, d.param, absdet, st);
      if (WANT_H) {
        sS[0 * chunk + sl] = st.s0;
        sS[1 * chunk + sl] = st.s1;
        sS[2 * chunk + sl] = st.s2;
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          if (WANT_H) {
            sS[(3 + 3 * a + i) * chunk + sl] = st.q[a][i];
            sS[(15 + 3 * a + i) * chunk + sl] = t.D[a][i];
          }
          if (WANT_G) sS[(27 + 3 * a + i) * chunk + sl] = st.g[a][i];
        }
    }
    __syncthreads();
    if (has_row) {
      while (ip < ie) {
        const uint64_t it = inc[ip];
        const int slot = int(it & 0xFFFFull);
        if (slot >= c0 + nc) break;
        const int sl = slot - c0;
        const int a = int((it >> 16) & 3ull);
        if (WANT_G) gacc += sS[(27 + 3 * a + comp) * chunk + sl];
        if (WANT_H) {
          const double s0 = sS[0 * chunk + sl], s1 = sS[1 * chunk + sl], s2 = sS[2 * chunk + sl];
          double qa[3], Da[3];
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            qa[i] = sS[(3 + 3 * a + i) * chunk + sl];
            Da[i] = sS[(15 + 3 * a + i) * chunk + sl];
          }
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const int bp = int((it >> (24 + 8 * b)) & 0xFFull);
            if (bp == 255) continue;
            double qb[3], Db[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              qb[i] = sS[(3 + 3 * b + i) * chunk + sl];
              Db[i] = sS[(15 + 3 * b + i) * chunk + sl];
            }
            const double dd = Da[0] * Db[0] + Da[1] * Db[1] + Da[2] * Db[2];
            const double qbi = qb[comp], qai = qa[comp];
            HT* dst = sVal + roff + 3 * bp;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              double h = s1 * qbi * qa[k] + s2 * qai * qb[k];
              if (k == comp) h += s0 * dd;
              dst[k] += HT(h);
            }
          }
        }
        ++ip;
      }
    }
    __syncthreads();
  }
  if (WANT_H)
    for (int k = tid; k < nslice; k += blockDim.x) vals[k0 + k] = sVal[k];
  if (WANT_G && has_row) g[r] = f ? gacc - f[r] : gacc;
}

__global__ __launch_bounds__(kBlock) void tet_element_kernel(fne_plan_desc d, const double* __restrict__ x,
                                                             const double* __restrict__ dvals,
                                                             double* __restrict__ e_out,
                                                             double* __restrict__ ge_out,
                                                             double* __restrict__ he_out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.n_elem) return;
  const bool absdet = (d.flags & FNE_FLAG_ABS_DET) != 0;
  TetElem t;
  tet_load(d, e, x, dvals, t);
  const int ne = d.n_elem;
  if (e_out) e_out[e] = tet_energy(t, d.param, absdet);
  if (!ge_out && !he_out) return;
  TetState s;
  tet_state(t, d.param, absdet, s);
  if (ge_out)
    for (int a = 0; a < 4; ++a)
      for (int i = 0; i < 3; ++i) ge_out[(3 * a + i) * ne + e] = s.g[a][i];
  if (he_out)
    for (int a = 0; a < 4; ++a)
      for (int i = 0; i < 3; ++i)
        for (int b = 0; b < 4; ++b) {
          const double dd = t.D[a][0] * t.D[b][0] + t.D[a][1] * t.D[b][1] + t.D[a][2] * t.D[b][2];
          for (int k = 0; k < 3; ++k) {
            double h = s.s1 * s.q[b][i] * s.q[a][k] + s.s2 * s.q[a][i] * s.q[b][k];
            if (i == k) h += s.s0 * dd;
            he_out[((3 * a + i) * 12 + 3 * b + k) * ne + e] = h;
          }
        }
}

}  // namespace

// Row-gather form (FNE_FLAG_ROW_GATHER): one thread per element writes its gradient
// (gbuf[e][32], FP64) and 12x12 Hessian (hbuf[e][144], HT) element-major; the CSR rows then
// gather them (fne_launch_row_gather).
template <typename HT, bool WANT_G, bool WANT_H>
__global__ __launch_bounds__(kBlock) void tet_store_kernel(fne_plan_desc d, const double* __restrict__ x,
                                                           const double* __restrict__ dvals,
                                                           double* __restrict__ gbuf, HT* __restrict__ hbuf) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.n_elem) return;
  const bool absdet = (d.flags & FNE_FLAG_ABS_DET) != 0;
  TetElem t;
  tet_load(d, e, x, dvals, t);
  TetState s;
  tet_state(t, d.param, absdet, s);
  if (WANT_G)
    for (int a = 0; a < 4; ++a)
      for (int i = 0; i < 3; ++i) gbuf[size_t(e) * 32 + 3 * a + i] = s.g[a][i];
  if (WANT_H) {
    HT* hb = hbuf + size_t(e) * 144;
    for (int a = 0; a < 4; ++a)
      for (int i = 0; i < 3; ++i)
        for (int b = 0; b < 4; ++b) {
          const double dd = t.D[a][0] * t.D[b][0] + t.D[a][1] * t.D[b][1] + t.D[a][2] * t.D[b][2];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            double h = s.s1 * s.q[b][i] * s.q[a][k] + s.s2 * s.q[a][i] * s.q[b][k];
            if (i == k) h += s.s0 * dd;
            hb[(3 * a + i) * 12 + 3 * b + k] = HT(h);
          }
        }
  }
}

// Hessian-vector product, row-gather form: one thread per element writes y_e = H_e v_e (the
// 12x12 element Hessian at x applied to the element's v, zero on constrained DOFs) into the
// element gradient scratch; the gradient row gather then sums each row in its fixed order.
__global__ __launch_bounds__(kBlock) void tet_hvp_store_kernel(fne_plan_desc d, const double* __restrict__ x,
                                                               const double* __restrict__ dvals,
                                                               const double* __restrict__ v,
                                                               double* __restrict__ gbuf) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.n_elem) return;
  const bool absdet = (d.flags & FNE_FLAG_ABS_DET) != 0;
  TetElem t;
  tet_load(d, e, x, dvals, t);
  TetState s;
  tet_state(t, d.param, absdet, s);
  double ve[4][3];
  for (int b = 0; b < 4; ++b)
    for (int k = 0; k < 3; ++k) {
      const int code = d.edof[size_t(3 * b + k) * d.n_elem + e];
      ve[b][k] = code < d.n_free ? v[code] : 0.0;
    }
  for (int a = 0; a < 4; ++a)
    for (int i = 0; i < 3; ++i) {
      double acc = 0.0;
      for (int b = 0; b < 4; ++b) {
        const double dd = t.D[a][0] * t.D[b][0] + t.D[a][1] * t.D[b][1] + t.D[a][2] * t.D[b][2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          double h = s.s1 * s.q[b][i] * s.q[a][k] + s.s2 * s.q[a][i] * s.q[b][k];
          if (i == k) h += s.s0 * dd;
          acc += h * ve[b][k];
        }
      }
      gbuf[size_t(e) * 32 + 3 * a + i] = acc;
    }
}

int fne_hvp_tet(fne_plan* plan, const double* x, const double* dvals, const double* v, double* y, hipStream_t stream) {
  const fne_plan_desc& d = plan->d;
  if (!(d.flags & FNE_FLAG_ROW_GATHER) || !plan->mc_g || !d.inc || !d.inc_ptr) {
    fne_set_error("tetrahedral Hessian-vector products need a row-gather plan (FNE_FLAG_ROW_GATHER)");
    return FNE_EUNSUPPORTED;
  }
  const int nb = (d.n_elem + kBlock - 1) / kBlock;
  if (nb > 0) {
    hipLaunchKernelGGL(tet_hvp_store_kernel, dim3(nb), dim3(kBlock), 0, stream, d, x, dvals, v, plan->mc_g);
    FNE_LAUNCH_CHECK();
  }
  FneCall c{};
  c.f = nullptr;
  c.g = y;
  c.precision = FNE_FP64;
  c.stream = stream;
  return fne_launch_row_gather(plan, true, false, c);
}

int fne_launch_tet(fne_plan* plan, FneOp op, const FneCall& c) {
  const fne_plan_desc& d = plan->d;
  if (op == kOpEnergy) {
    const int nb = plan->n_partials;
    hipLaunchKernelGGL(tet_energy_kernel, dim3(nb), dim3(kBlock), 0, c.stream, d, c.x, c.dvals, c.f,
                       plan->partials);
    FNE_LAUNCH_CHECK();
    return fne_finalize_energy(plan, nb, c.out, c.stream);
  }
  if (op == kOpFused) {
    if (c.out) {
      const int rc = fne_launch_tet(plan, kOpEnergy, c);
      if (rc) return rc;
    }
    return fne_launch_tet(plan, kOpHessian, c);
  }
  if (op == kOpElementEval) {
    const int nb = (d.n_elem + kBlock - 1) / kBlock;
    if (nb > 0) {
      hipLaunchKernelGGL(tet_element_kernel, dim3(nb), dim3(kBlock), 0, c.stream, d, c.x, c.dvals, c.e_out,
                         c.ge_out, c.he_out);
      FNE_LAUNCH_CHECK();
    }
    return FNE_OK;
  }
  const bool want_h = (op == kOpHessian);
  const bool want_g = (op == kOpGradient) || (c.g != nullptr);
  if (d.flags & FNE_FLAG_ROW_GATHER) {
    if (!plan->mc_g || !plan->mc_h || !d.inc || !d.inc_ptr || (want_h && !d.pitem)) {
      fne_set_error("row-gather arrays missing from the tetrahedron plan");
      return FNE_EINVAL;
    }
    const int nb = (d.n_elem + kBlock - 1) / kBlock;
    if (nb > 0) {
      if (!want_h)
        hipLaunchKernelGGL((tet_store_kernel<double, true, false>), dim3(nb), dim3(kBlock), 0, c.stream, d, c.x,
                           c.dvals, plan->mc_g, (double*)plan->mc_h);
      else if (c.precision == FNE_FP32)
        hipLaunchKernelGGL((tet_store_kernel<float, true, true>), dim3(nb), dim3(kBlock), 0, c.stream, d, c.x,
                           c.dvals, plan->mc_g, (float*)plan->mc_h);
      else
        hipLaunchKernelGGL((tet_store_kernel<double, true, true>), dim3(nb), dim3(kBlock), 0, c.stream, d, c.x,
                           c.dvals, plan->mc_g, (double*)plan->mc_h);
      FNE_LAUNCH_CHECK();
    }
    return fne_launch_row_gather(plan, want_g, want_h, c);
  }
  if (d.n_patch <= 0) return FNE_OK;
  const size_t hsz = c.precision == FNE_FP32 ? 4 : 8;
  size_t shm = sizeof(double) * size_t(kTetStage) * size_t(d.chunk);
  if (want_h) shm += ((size_t(d.max_patch_nnz) * hsz + 7) / 8) * 8;
  if (shm > plan->max_lds) {
    fne_set_error("tet patch LDS request " + std::to_string(shm) + " B exceeds the device limit " +
                  std::to_string(plan->max_lds) + " B");
    return FNE_EINVAL;
  }
  const dim3 grid(d.n_patch), block(kBlock);
  if (!want_h) {
    hipLaunchKernelGGL((tet_patch_kernel<double, true, false>), grid, block, shm, c.stream, d, c.x, c.dvals,
                       c.f, c.g, (double*)nullptr);
  } else if (c.precision == FNE_FP32) {
    if (want_g)
      hipLaunchKernelGGL((tet_patch_kernel<float, true, true>), grid, block, shm, c.stream, d, c.x, c.dvals,
                         c.f, c.g, (float*)c.vals);
    else
      hipLaunchKernelGGL((tet_patch_kernel<float, false, true>), grid, block, shm, c.stream, d, c.x, c.dvals,
                         c.f, c.g, (float*)c.vals);
  } else {
    if (want_g)
      hipLaunchKernelGGL((tet_patch_kernel<double, true, true>), grid, block, shm, c.stream, d, c.x, c.dvals,
                         c.f, c.g, (double*)c.vals);
    else
      hipLaunchKernelGGL((tet_patch_kernel<double, false, true>), grid, block, shm, c.stream, d, c.x, c.dvals,
                         c.f, c.g, (double*)c.vals);
  }
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}
