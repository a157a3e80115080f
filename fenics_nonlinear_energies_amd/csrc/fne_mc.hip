// Plane-strain Mohr-Coulomb on P2/P4 triangles (rows A4/A5 of the hot path).
//
// Element energy (slope_stability/jax/jax_energy.py:45-142): per quadrature point q,
// eps = B_q u_e (engineering shear), sigma_tr = C eps, principal values s1,2 = m +- r with
// r = sqrt(d^2 + tau^2 + reg^2); elastic if alpha s1 - beta s2 - kappa <= 0, else a return
// to the yield line along z = A^-1 n (apex if s1 < s2 afterwards); sigma is rebuilt with the
// trial principal direction; psi = sigma.eps - 1/2 sigma.S sigma; E_e = sum_q w_q psi_q.
// The gradient is sum_q w_q B_q^T sigma_q and the Hessian sum_q w_q B_q^T D_q B_q with
// D = d sigma / d eps derived by hand (line return: chain rule through (m, r, theta); apex:
// zero), symmetrised.
//
// Two kernels per Newton-step assembly:
//   mc_elem_kernel   one wavefront evaluates 3 elements at a time.  FP64 per quadrature point
//                    (strain, return mapping, energy, stress) on 3 x n_q lanes; the element
//                    gradient in FP64; the 2n_p x 2n_p element Hessian as three 16x16 tiles
//                    (xx, xy, yy blocks; yx = xy^T) of v_mfma 16x16x4 over K = 2 n_q,
//                    A[n][(q,a)] = dphi_n/dx_a(q), B^{ij}[(q,a)][m] = sum_b wD^{ij}_{ab} dphi_m/dx_b.
//                    Results go to a per-element scratch (element order of the plan).
//   mc_row_gather    32 lanes per CSR row: for each incident element (fixed ascending order)
//                    lane c adds the element row's entry c into an LDS copy of the CSR row at
//                    the precomputed byte position; the row is then written once.  No atomics:
//                    results are bitwise reproducible.
#include "fne_common.h"

#include <cmath>

namespace {

constexpr int kMcMaxQ = 21;  // 3 elements x 21 points fill one wavefront
constexpr int kMcEpw = 3;    // elements per wavefront iteration
constexpr int kMcWaves = kBlock / 64;
// lanes per row group in the row gather: lane c < NL/2 carries element columns 2c and 2c+1, so
// 8 lanes cover a 12-column element row and 16 the 30 of P4
__host__ __device__ constexpr int mc_row_lanes(int nl) {
  return nl <= 16 ? 8 : 16;
}
constexpr unsigned char kPosNone = 255;

typedef float mc_f4 __attribute__((ext_vector_type(4)));
typedef double mc_d4 __attribute__((ext_vector_type(4)));

struct McConst {
  double C00, C01, C22;  // C = [[C00, C01, 0], [C01, C00, 0], [0, 0, C22]]
  double S00, S01, S22;  // S = C^-1 in plane strain
  double alpha, beta, kappa, reg;
  double z1, z2, nz;     // return direction z = A^-1 n, n = (alpha, -beta)
  double s_apex;
  int mode;              // 0 full Mohr-Coulomb, 1 elastic
};

__device__ __forceinline__ McConst mc_const(const double* prm, int mode) {
  McConst k;
  const double E = prm[0], nu = prm[1], phi_deg = prm[2], coh = prm[3];
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = E / (2.0 * (1.0 + nu));
  k.C00 = lam + 2.0 * mu;
  k.C01 = lam;
  k.C22 = mu;
  k.S00 = (1.0 - nu * nu) / E;
  k.S01 = -(nu * (1.0 + nu)) / E;
  k.S22 = 1.0 / mu;
  const double phi = phi_deg * (M_PI / 180.0);
  const double sphi = sin(phi), cphi = cos(phi);
  k.alpha = 1.0 + sphi;
  k.beta = 1.0 - sphi;
  k.kappa = 2.0 * coh * cphi;
  k.reg = prm[4];
  const double A11 = k.S00, A12 = k.S01, A22 = k.S00;
  const double n1 = k.alpha, n2 = -k.beta;
  const double det = A11 * A22 - A12 * A12;
  k.z1 = (A22 * n1 - A12 * n2) / det;
  k.z2 = (-A12 * n1 + A11 * n2) / det;
  k.nz = n1 * k.z1 + n2 * k.z2;
  const double den = fabs(k.alpha - k.beta) > 1.0e-14 ? k.alpha - k.beta : 1.0;
  k.s_apex = k.kappa / den;
  k.mode = mode;
  return k;
}

// One quadrature point: energy density, stress and (optionally) the symmetric tangent
// D = [d00 d01 d02 d11 d12 d22] (Voigt, engineering shear).
template <bool WANT_D>
__device__ __forceinline__ double mc_point(const McConst& k, double e0, double e1, double e2, double sig[3],
                                           double D[6]) {
  const double st0 = k.C00 * e0 + k.C01 * e1;
  const double st1 = k.C01 * e0 + k.C00 * e1;
  const double st2 = k.C22 * e2;
  bool elastic = k.mode == 1;
  double m = 0.0, d = 0.0, r = 0.0, s1 = 0.0, s2 = 0.0;
  if (!elastic) {
    m = 0.5 * (st0 + st1);
    d = 0.5 * (st0 - st1);
    r = sqrt(d * d + st2 * st2 + k.reg * k.reg);
    s1 = m + r;
    s2 = m - r;
    elastic = k.alpha * s1 - k.beta * s2 - k.kappa <= 0.0;
  }
  if (elastic) {
    sig[0] = st0;
    sig[1] = st1;
    sig[2] = st2;
    if (WANT_D) {
      D[0] = k.C00;
      D[1] = k.C01;
      D[2] = 0.0;
      D[3] = k.C00;
      D[4] = 0.0;
      D[5] = k.C22;
    }
  } else {
    const double n1 = k.alpha, n2 = -k.beta;
    const double lam_mc = (n1 * s1 + n2 * s2 - k.kappa) / k.nz;
    const double s1l = s1 - lam_mc * k.z1;
    const double s2l = s2 - lam_mc * k.z2;
    const bool line = s1l >= s2l;
    const double s1p = line ? s1l : k.s_apex;
    const double s2p = line ? s2l : k.s_apex;
    const double mp = 0.5 * (s1p + s2p);
    const double rp = 0.5 * (s1p - s2p);
    // 2 theta = atan2(2 tau, sxx - syy): cos/sin from the trial deviator (atan2(0,0) = 0)
    const double r0 = sqrt(d * d + st2 * st2);
    const double c2 = r0 > 0.0 ? d / r0 : 1.0;
    const double sn2 = r0 > 0.0 ? st2 / r0 : 0.0;
    sig[0] = mp + rp * c2;
    sig[1] = mp - rp * c2;
    sig[2] = rp * sn2;
    if (WANT_D) {
      if (!line) {
#pragma unroll
        for (int i = 0; i < 6; ++i) D[i] = 0.0;
      } else {
        const double P11 = 1.0 - k.z1 * n1 / k.nz, P12 = -k.z1 * n2 / k.nz;
        const double P21 = -k.z2 * n1 / k.nz, P22 = 1.0 - k.z2 * n2 / k.nz;
        const double am = 0.5 * (P11 + P12 + P21 + P22), ar = 0.5 * (P11 - P12 + P21 - P22);
        const double bm = 0.5 * (P11 + P12 - P21 - P22), br = 0.5 * (P11 - P12 - P21 + P22);
        const double ir = r > 0.0 ? 1.0 / r : 0.0;
        const double ir0 = r0 > 0.0 ? 1.0 / r0 : 0.0;
        // derivatives with respect to the trial stress (sxx, syy, txy)
        const double dr[3] = {0.5 * d * ir, -0.5 * d * ir, st2 * ir};
        const double dph[3] = {0.5 * sn2 * ir0, -0.5 * sn2 * ir0, -c2 * ir0};
        const double dm[3] = {0.5, 0.5, 0.0};
        double Ds[3][3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const double dmp = am * dm[j] + ar * dr[j];
          const double drp = bm * dm[j] + br * dr[j];
          Ds[0][j] = dmp + c2 * drp + rp * sn2 * dph[j];
          Ds[1][j] = dmp - c2 * drp - rp * sn2 * dph[j];
          Ds[2][j] = sn2 * drp - rp * c2 * dph[j];
        }
        // D = Ds * C, then symmetrise
        double Dm[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          Dm[i][0] = Ds[i][0] * k.C00 + Ds[i][1] * k.C01;
          Dm[i][1] = Ds[i][0] * k.C01 + Ds[i][1] * k.C00;
          Dm[i][2] = Ds[i][2] * k.C22;
        }
        D[0] = Dm[0][0];
        D[1] = 0.5 * (Dm[0][1] + Dm[1][0]);
        D[2] = 0.5 * (Dm[0][2] + Dm[2][0]);
        D[3] = Dm[1][1];
        D[4] = 0.5 * (Dm[1][2] + Dm[2][1]);
        D[5] = Dm[2][2];
      }
    }
  }
  const double Ss0 = k.S00 * sig[0] + k.S01 * sig[1];
  const double Ss1 = k.S01 * sig[0] + k.S00 * sig[1];
  const double Ss2 = k.S22 * sig[2];
  return sig[0] * e0 + sig[1] * e1 + sig[2] * e2 - 0.5 * (sig[0] * Ss0 + sig[1] * Ss1 + sig[2] * Ss2);
}

__device__ __forceinline__ mc_f4 mfma16(float a, float b, mc_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ mc_d4 mfma16(double a, double b, mc_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// Result layout of the 16x16x4 products: lane l holds D[row][l % 16] in slot r, with
// row = 4 (l / 16) + r for the FP32 form and row = l / 16 + 4 r for the FP64 form.
template <typename T> struct Vec4;
template <> struct Vec4<float> {
  typedef mc_f4 type;
  static __device__ __forceinline__ int row(int kk, int r) { return 4 * kk + r; }
};
template <> struct Vec4<double> {
  typedef mc_d4 type;
  static __device__ __forceinline__ int row(int kk, int r) { return kk + 4 * r; }
};

// LDS of one wavefront
template <typename HT>
struct McWaveLds {
  double u[kMcEpw][32];            // element DOF values
  double geo[kMcEpw][8];           // K00 K01 K10 K11 |detJ|
  double P[kMcEpw][kMcMaxQ][4];    // w * sigma_{i a} K[a][g], i-major
  HT wD[kMcEpw][kMcMaxQ][6];       // w * D (Voigt, symmetric)
  double psi[kMcEpw][kMcMaxQ];     // w * psi (per-element energies)
};

// NN nodes per element (6: P2, 15: P4); element matrices have NL = 2 NN rows.
template <int NN, typename HT, bool WE, bool WG, bool WH>
__global__ __launch_bounds__(kBlock) void mc_elem_kernel(fne_plan_desc d, int mode, const double* __restrict__ x,
                                                         const double* __restrict__ dvals,
                                                         const double* __restrict__ f,
                                                         double2* __restrict__ partial,
                                                         double* __restrict__ eelem,
                                                         double* __restrict__ gbuf, HT* __restrict__ hbuf) {
  constexpr int NL = 2 * NN;
  __shared__ double sTab[kMcMaxQ * 16 * 2];  // reference gradients [q][n][g]
  __shared__ double sW[kMcMaxQ];
  __shared__ McWaveLds<HT> sWave[kMcWaves];
  const int nq = d.nq;
  for (int i = threadIdx.x; i < nq * NN * 2; i += blockDim.x) sTab[i] = d.qdata[nq + i];
  for (int i = threadIdx.x; i < nq; i += blockDim.x) sW[i] = d.qdata[i];
  __syncthreads();
  const McConst k = mc_const(d.param, mode);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  McWaveLds<HT>& L = sWave[wave];
  const int ne = d.n_elem;
  const bool has_eps = d.n_geo >= 5 + 3 * nq;
  double es = 0.0, ec = 0.0;

  const int nwaves = gridDim.x * kMcWaves;
  for (int base = (blockIdx.x * kMcWaves + wave) * kMcEpw; base < ne; base += nwaves * kMcEpw) {
    // ---- gather element values and geometry
    for (int i = lane; i < kMcEpw * NL; i += 64) {
      const int t = i / NL, c = i - t * NL;
      const int e = base + t;
      L.u[t][c] = e < ne ? gather_dof(d.edof[size_t(c) * ne + e], d.n_free, x, dvals) : 0.0;
    }
    if (lane < kMcEpw * 5) {
      const int t = lane / 5, c = lane - 5 * t;
      const int e = base + t;
      L.geo[t][c] = e < ne ? d.egeo[size_t(c) * ne + e] : (c == 4 ? 0.0 : 1.0);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // ---- quadrature points: lanes t*nq + q
    if (lane < kMcEpw * nq) {
      const int t = lane / nq, q = lane - t * nq;
      const int e = base + t;
      if (e < ne) {
        double U00 = 0.0, U01 = 0.0, U10 = 0.0, U11 = 0.0;  // U[i][g] = sum_n u_{n,i} dphi_n/dxi_g
#pragma unroll
        for (int n = 0; n < NN; ++n) {
          const double g0 = sTab[(q * NN + n) * 2 + 0], g1 = sTab[(q * NN + n) * 2 + 1];
          const double ux = L.u[t][2 * n], uy = L.u[t][2 * n + 1];
          U00 += ux * g0;
          U01 += ux * g1;
          U10 += uy * g0;
          U11 += uy * g1;
        }
        const double K00 = L.geo[t][0], K01 = L.geo[t][1], K10 = L.geo[t][2], K11 = L.geo[t][3];
        double e0 = K00 * U00 + K01 * U01;
        double e1 = K10 * U10 + K11 * U11;
        double e2 = (K10 * U00 + K11 * U01) + (K00 * U10 + K01 * U11);
        if (has_eps) {
          const size_t o = size_t(5 + 3 * q) * ne + e;
          e0 -= d.egeo[o];
          e1 -= d.egeo[o + ne];
          e2 -= d.egeo[o + 2 * size_t(ne)];
        }
        double sig[3], D[6];
        const double w = sW[q] * L.geo[t][4];
        double psi;
        if (WH)
          psi = mc_point<true>(k, e0, e1, e2, sig, D);
        else
          psi = mc_point<false>(k, e0, e1, e2, sig, D);
        double wpsi = w * psi;
        if (d.eweight) wpsi *= d.eweight[e];
        if (WE) acc_add(es, ec, wpsi);
        if (eelem) L.psi[t][q] = wpsi;
        if (WG) {
          const double sx0 = w * sig[0], sy1 = w * sig[1], sxy = w * sig[2];
          L.P[t][q][0] = sx0 * K00 + sxy * K10;  // i = x, g = 0
          L.P[t][q][1] = sx0 * K01 + sxy * K11;  // i = x, g = 1
          L.P[t][q][2] = sxy * K00 + sy1 * K10;  // i = y, g = 0
          L.P[t][q][3] = sxy * K01 + sy1 * K11;  // i = y, g = 1
        }
        if (WH) {
#pragma unroll
          for (int i = 0; i < 6; ++i) L.wD[t][q][i] = HT(w * D[i]);
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (eelem && lane < kMcEpw) {
      const int e = base + lane;
      if (e < ne) {
        double s = 0.0;
        for (int q = 0; q < nq; ++q) s += L.psi[lane][q];
        eelem[e] = s;
      }
    }
    // ---- element gradient (FP64): lanes (t, row)
    if (WG) {
      for (int i = lane; i < kMcEpw * NL; i += 64) {
        const int t = i / NL, rr = i - t * NL;
        const int e = base + t;
        if (e >= ne) continue;
        const int n = rr >> 1, comp = rr & 1;
        double s = 0.0;
        for (int q = 0; q < nq; ++q) {
          const double g0 = sTab[(q * NN + n) * 2 + 0], g1 = sTab[(q * NN + n) * 2 + 1];
          s += g0 * L.P[t][q][2 * comp] + g1 * L.P[t][q][2 * comp + 1];
        }
        gbuf[size_t(e) * 32 + rr] = s;
      }
    }
    // ---- element Hessian: three 16x16 MFMA tiles per element
    if (WH) {
      typedef typename Vec4<HT>::type V4;
      const int n = lane & 15;
      const int kk = lane >> 4;
      const int steps = (2 * nq + 3) >> 2;
      for (int t = 0; t < kMcEpw; ++t) {
        const int e = base + t;
        if (e >= ne) break;  // uniform over the wavefront
        const HT K00 = HT(L.geo[t][0]), K01 = HT(L.geo[t][1]), K10 = HT(L.geo[t][2]), K11 = HT(L.geo[t][3]);
        V4 axx = {0, 0, 0, 0}, axy = {0, 0, 0, 0}, ayy = {0, 0, 0, 0};
        for (int s = 0; s < steps; ++s) {
          const int kidx = 4 * s + kk;
          const int q = kidx >> 1, a = kidx & 1;
          HT av = HT(0), bxx = HT(0), bxy = HT(0), byy = HT(0);
          if (q < nq && n < NN) {
            const HT g0 = HT(sTab[(q * NN + n) * 2 + 0]), g1 = HT(sTab[(q * NN + n) * 2 + 1]);
            const HT gx = K00 * g0 + K01 * g1;
            const HT gy = K10 * g0 + K11 * g1;
            const HT* D = L.wD[t][q];  // d00 d01 d02 d11 d12 d22
            // row of D for (x, a): v(x,x)=0, v(x,y)=2; for (y, a): v(y,x)=2, v(y,y)=1
            const HT dx0 = a == 0 ? D[0] : D[2];  // D[v(x,a)][0]
            const HT dx1 = a == 0 ? D[1] : D[4];  // D[v(x,a)][1]
            const HT dx2 = a == 0 ? D[2] : D[5];  // D[v(x,a)][2]
            const HT dy1 = a == 0 ? D[4] : D[3];  // D[v(y,a)][1]
            const HT dy2 = a == 0 ? D[5] : D[4];  // D[v(y,a)][2]
            av = a == 0 ? gx : gy;
            bxx = dx0 * gx + dx2 * gy;
            bxy = dx2 * gx + dx1 * gy;
            byy = dy2 * gx + dy1 * gy;
          }
          axx = mfma16(av, bxx, axx);
          axy = mfma16(av, bxy, axy);
          ayy = mfma16(av, byy, ayy);
        }
        HT* hb = hbuf + size_t(e) * (NL * NL);
        const int m = lane & 15;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int nr = Vec4<HT>::row(kk, r);
          if (nr < NN && m < NN) {
            hb[(2 * nr) * NL + 2 * m] = axx[r];
            hb[(2 * nr) * NL + 2 * m + 1] = axy[r];
            hb[(2 * m + 1) * NL + 2 * nr] = axy[r];
            hb[(2 * nr + 1) * NL + 2 * m + 1] = ayy[r];
          }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  if (WE) {
    if (f) {
      const int stride = gridDim.x * blockDim.x;
      for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n_free; i += stride) acc_add(es, ec, -(f[i] * x[i]));
    }
    block_reduce_pair(es, ec);
    if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(es, ec);
  }
}

template <typename T> struct Pair;
template <> struct Pair<float> {
  typedef float2 type;
  static __device__ __forceinline__ float2 zero() { return make_float2(0.f, 0.f); }
};
template <> struct Pair<double> {
  typedef double2 type;
  static __device__ __forceinline__ double2 zero() { return make_double2(0.0, 0.0); }
};

// Row gather.  A group of mc_row_lanes(NL) lanes owns mc_row_batch(NL) consecutive CSR rows; lane c < NL/2 handles
// the element columns 2c and 2c+1 (paired loads).  inc[it] = element * 32 + local row lists the incidences of
// every row (rows consecutive, elements ascending); pos[it*32 + c] is the position of the
// element's local column c inside the row (255 = constrained).  The incidences of the whole
// batch are read in one load (lane j holds item j), up to 8 items' element rows and positions
// are fetched together, and then added in item order into LDS copies of the rows, which are
// written once.  Every CSR entry therefore sums its contributions in a fixed order.
// rows per lane group: one for 12-column element rows (P2, tets), two for the 30 of P4
// (measured on MI355X: P4 0.50 ms with 2 rows against 0.52 with 4 and 0.55 with 1; the
// 320x16x16 tet beam 0.335 ms with 1 row against 0.37 with 2 and 0.40 with 4)
__host__ __device__ constexpr int mc_row_batch(int nl) {
  return nl <= 12 ? 1 : 2;
}
template <int NL, typename HT, bool WG, bool WH>
__global__ __launch_bounds__(kBlock) void mc_row_gather_kernel(fne_plan_desc d, const HT* __restrict__ hbuf,
                                                               const double* __restrict__ gbuf,
                                                               const double* __restrict__ f,
                                                               double* __restrict__ g, HT* __restrict__ vals) {
  extern __shared__ unsigned char mc_smem[];
  constexpr int kL = mc_row_lanes(NL);
  constexpr int kRB = mc_row_batch(NL);
  constexpr int kRowsBlk = (kBlock / kL) * kRB;
  constexpr int kPf = 8;
  const int cap = (d.max_row_len + 1) & ~1;
  const int grp = threadIdx.x / kL;
  const int c = threadIdx.x & (kL - 1);
  HT* sRows = reinterpret_cast<HT*>(mc_smem) + size_t(grp) * kRB * cap;
  const int r0 = xcd_remap(blockIdx.x, gridDim.x) * kRowsBlk + grp * kRB;
  if (r0 >= d.n_free) return;  // uniform over the lane group; no block-level barrier below
  const int nr = min(kRB, d.n_free - r0);
  const unsigned char* pos = reinterpret_cast<const unsigned char*>(d.pitem);
  const unsigned int* inc = reinterpret_cast<const unsigned int*>(d.inc);
  const int ip_l = c <= nr ? d.indptr[r0 + c] : 0;
  const int ic_l = c <= nr ? d.inc_ptr[r0 + c] : 0;
  int kb[kRB + 1], ib[kRB + 1];
#pragma unroll
  for (int t = 0; t <= kRB; ++t) {
    kb[t] = __shfl(ip_l, t < nr ? t : nr, kL);
    ib[t] = __shfl(ic_l, t < nr ? t : nr, kL);
  }
  if (WH) {
#pragma unroll
    for (int t = 0; t < kRB; ++t)
      if (t < nr)
        for (int j = c; j < kb[t + 1] - kb[t]; j += kL) sRows[t * cap + j] = HT(0);
  }
  __builtin_amdgcn_wave_barrier();
  double gacc = 0.0;  // lane t < nr: gradient of row r0 + t
  for (int base = ib[0]; base < ib[nr]; base += kL) {
    const int nb = min(kL, ib[nr] - base);
    const unsigned int mine = c < nb ? inc[base + c] : 0u;
    int myslot = 0;
#pragma unroll
    for (int t = 1; t < kRB; ++t)
      if (t < nr && base + c >= ib[t]) myslot = t;
    if (WG) {
      const double gv = c < nb ? gbuf[size_t(mine >> 5) * 32 + (mine & 31u)] : 0.0;
      for (int j = 0; j < nb; ++j) {
        const double v = __shfl(gv, j, kL);
        const int t = __shfl(myslot, j, kL);
        if (c == t) gacc += v;
      }
    }
    if (WH) {
      // lane c < NL/2 carries the element columns 2c and 2c+1: one paired value load and one
      // 16-bit position load per item
      typedef typename Pair<HT>::type HT2;
      constexpr int kHalf = NL / 2;
      for (int j0 = 0; j0 < nb; j0 += kPf) {
        HT2 v[kPf];
        unsigned int pp[kPf];
        int sl[kPf];
#pragma unroll
        for (int jj = 0; jj < kPf; ++jj) {
          const int j = j0 + jj;
          const unsigned int item = __shfl(mine, j < nb ? j : 0, kL);
          sl[jj] = __shfl(myslot, j < nb ? j : 0, kL);
          v[jj] = Pair<HT>::zero();
          pp[jj] = 0xFFFFu;
          if (j < nb && c < kHalf) {
            pp[jj] = *reinterpret_cast<const unsigned short*>(pos + size_t(base + j) * 32 + 2 * c);
            v[jj] = *reinterpret_cast<const HT2*>(hbuf + size_t(item >> 5) * (NL * NL) + (item & 31u) * NL + 2 * c);
          }
        }
#pragma unroll
        for (int jj = 0; jj < kPf; ++jj) {
          const unsigned int p0 = pp[jj] & 0xFFu, p1 = pp[jj] >> 8;
          if (p0 != kPosNone) sRows[sl[jj] * cap + p0] += v[jj].x;
          if (p1 != kPosNone) sRows[sl[jj] * cap + p1] += v[jj].y;
          __builtin_amdgcn_wave_barrier();
        }
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (WH) {
#pragma unroll
    for (int t = 0; t < kRB; ++t)
      if (t < nr)
        for (int j = c; j < kb[t + 1] - kb[t]; j += kL) vals[kb[t] + j] = sRows[t * cap + j];
  }
  if (WG && c < nr) g[r0 + c] = f ? gacc - f[r0 + c] : gacc;
}

// Element outputs for fne_element_eval: scratch (element-major) -> SoA outputs.
template <int NL>
__global__ __launch_bounds__(kBlock) void mc_element_out_kernel(int ne, const double* __restrict__ gbuf,
                                                                const double* __restrict__ hbuf,
                                                                double* __restrict__ ge_out,
                                                                double* __restrict__ he_out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  if (ge_out)
    for (int i = 0; i < NL; ++i) ge_out[size_t(i) * ne + e] = gbuf[size_t(e) * 32 + i];
  if (he_out)
    for (int i = 0; i < NL * NL; ++i) he_out[size_t(i) * ne + e] = hbuf[size_t(e) * NL * NL + i];
}

template <int NN>
int launch_mc(fne_plan* plan, FneOp op, const FneCall& c) {
  constexpr int NL = 2 * NN;
  const fne_plan_desc& d = plan->d;
  const int ne = d.n_elem;
  int nb = (ne + kMcWaves * kMcEpw - 1) / (kMcWaves * kMcEpw);
  if (nb > kEnergyBlocks) nb = kEnergyBlocks;
  if (nb < 1) nb = 1;
  if (nb > plan->n_partials) {
    fne_set_error("energy reduction scratch too small");
    return FNE_EINVAL;
  }
  const int mode = plan->mode;
  const bool want_e = op == kOpEnergy || (op == kOpFused && c.out);
  const bool want_g = op == kOpGradient || ((op == kOpHessian || op == kOpFused) && c.g);
  const bool want_h = op == kOpHessian || op == kOpFused;
  double* gb = plan->mc_g;
  void* hb = plan->mc_h;

  if (op == kOpElementEval) {
    hipLaunchKernelGGL((mc_elem_kernel<NN, double, false, true, true>), dim3(nb), dim3(kBlock), 0, c.stream, d,
                       mode, c.x, c.dvals, (const double*)nullptr, plan->partials, c.e_out, gb, (double*)hb);
    FNE_LAUNCH_CHECK();
    if (c.ge_out || c.he_out) {
      hipLaunchKernelGGL(mc_element_out_kernel<NL>, dim3((ne + kBlock - 1) / kBlock), dim3(kBlock), 0, c.stream,
                         ne, gb, (const double*)hb, c.ge_out, c.he_out);
      FNE_LAUNCH_CHECK();
    }
    return FNE_OK;
  }
  if ((want_g || want_h) && (!d.inc_ptr || !d.inc || (want_h && !d.pitem))) {
    fne_set_error("row incidence arrays missing from the plan");
    return FNE_EINVAL;
  }
  if (want_h && d.max_row_len > 254) {
    fne_set_error("CSR rows longer than 254 entries are not supported");
    return FNE_EINVAL;
  }
  const bool fp32 = c.precision == FNE_FP32;
#define MC_ELEM(HT, E, G, H)                                                                          \
  hipLaunchKernelGGL((mc_elem_kernel<NN, HT, E, G, H>), dim3(nb), dim3(kBlock), 0, c.stream, d, mode, c.x, \
                     c.dvals, c.f, plan->partials, (double*)nullptr, gb, (HT*)hb)
  if (want_h) {
    if (fp32) {
      if (want_e && want_g) MC_ELEM(float, true, true, true);
      else if (want_e) MC_ELEM(float, true, false, true);
      else if (want_g) MC_ELEM(float, false, true, true);
      else MC_ELEM(float, false, false, true);
    } else {
      if (want_e && want_g) MC_ELEM(double, true, true, true);
      else if (want_e) MC_ELEM(double, true, false, true);
      else if (want_g) MC_ELEM(double, false, true, true);
      else MC_ELEM(double, false, false, true);
    }
  } else if (want_g) {
    if (want_e) MC_ELEM(double, true, true, false);
    else MC_ELEM(double, false, true, false);
  } else {
    MC_ELEM(double, true, false, false);
  }
#undef MC_ELEM
  FNE_LAUNCH_CHECK();
  if (want_e) {
    const int rc = fne_finalize_energy(plan, nb, c.out, c.stream);
    if (rc) return rc;
  }
  if (!want_g && !want_h) return FNE_OK;
  const int rows_per_block = (kBlock / mc_row_lanes(NL)) * mc_row_batch(NL);
  const dim3 grid((d.n_free + rows_per_block - 1) / rows_per_block);
  const int cap = (d.max_row_len + 1) & ~1;
  if (want_h && size_t(rows_per_block) * cap * (fp32 ? 4 : 8) > plan->max_lds) {
    fne_set_error("row gather LDS request exceeds the device limit");
    return FNE_EINVAL;
  }
  if (want_h) {
    const size_t shm = size_t(rows_per_block) * cap * (fp32 ? 4 : 8);
    if (fp32) {
      if (want_g)
        hipLaunchKernelGGL((mc_row_gather_kernel<NL, float, true, true>), grid, dim3(kBlock), shm, c.stream, d,
                           (const float*)hb, gb, c.f, c.g, (float*)c.vals);
      else
        hipLaunchKernelGGL((mc_row_gather_kernel<NL, float, false, true>), grid, dim3(kBlock), shm, c.stream, d,
                           (const float*)hb, gb, c.f, c.g, (float*)c.vals);
    } else {
      if (want_g)
        hipLaunchKernelGGL((mc_row_gather_kernel<NL, double, true, true>), grid, dim3(kBlock), shm, c.stream, d,
                           (const double*)hb, gb, c.f, c.g, (double*)c.vals);
      else
        hipLaunchKernelGGL((mc_row_gather_kernel<NL, double, false, true>), grid, dim3(kBlock), shm, c.stream, d,
                           (const double*)hb, gb, c.f, c.g, (double*)c.vals);
    }
  } else {
    hipLaunchKernelGGL((mc_row_gather_kernel<NL, double, true, false>), grid, dim3(kBlock), 0, c.stream, d,
                       (const double*)hb, gb, c.f, c.g, (double*)nullptr);
  }
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

// ---------------------------------------------------------------------------------------------
// Matrix-free Hessian products (slope_stability/jax_petsc/reordered_element_assembler.py:313-562).
// AUX = 1: element vectors H_e v_e = sum_q w_q B_q^T D_q (B_q v_e) with D_q the tangent of the
// return mapping at the state x (the matrix the element kernel assembles); AUX = 2: the element
// diagonals diag(H_e)_a = sum_q w_q b_a^T D_q b_a.  Both are written to the element-vector
// scratch and summed per row by the row gather, like the gradient.  FP64 throughout.
// ---------------------------------------------------------------------------------------------
struct McAuxLds {
  double u[kMcEpw][32];
  double v[kMcEpw][32];
  double geo[kMcEpw][8];
  double P[kMcEpw][kMcMaxQ][4];  // AUX 1: w * (D B v)_{i a} K[a][g]
  double D[kMcEpw][kMcMaxQ][6];  // AUX 2: w * D
};

template <int NN, int AUX>
__global__ __launch_bounds__(kBlock) void mc_aux_kernel(fne_plan_desc d, int mode, const double* __restrict__ x,
                                                        const double* __restrict__ dvals,
                                                        const double* __restrict__ vin, double* __restrict__ gbuf) {
  constexpr int NL = 2 * NN;
  __shared__ double sTab[kMcMaxQ * 16 * 2];
  __shared__ double sW[kMcMaxQ];
  __shared__ McAuxLds sWave[kMcWaves];
  const int nq = d.nq;
  for (int i = threadIdx.x; i < nq * NN * 2; i += blockDim.x) sTab[i] = d.qdata[nq + i];
  for (int i = threadIdx.x; i < nq; i += blockDim.x) sW[i] = d.qdata[i];
  __syncthreads();
  const McConst k = mc_const(d.param, mode);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  McAuxLds& L = sWave[wave];
  const int ne = d.n_elem;
  const bool has_eps = d.n_geo >= 5 + 3 * nq;
  const int nwaves = gridDim.x * kMcWaves;
  for (int base = (blockIdx.x * kMcWaves + wave) * kMcEpw; base < ne; base += nwaves * kMcEpw) {
    for (int i = lane; i < kMcEpw * NL; i += 64) {
      const int t = i / NL, c = i - t * NL;
      const int e = base + t;
      const int code = e < ne ? d.edof[size_t(c) * ne + e] : d.n_free;
      L.u[t][c] = e < ne ? gather_dof(code, d.n_free, x, dvals) : 0.0;
      if (AUX == 1) L.v[t][c] = (e < ne && code < d.n_free) ? vin[code] : 0.0;  // v vanishes on constrained DOFs
    }
    if (lane < kMcEpw * 5) {
      const int t = lane / 5, c = lane - 5 * t;
      const int e = base + t;
      L.geo[t][c] = e < ne ? d.egeo[size_t(c) * ne + e] : (c == 4 ? 0.0 : 1.0);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (lane < kMcEpw * nq) {
      const int t = lane / nq, q = lane - t * nq;
      const int e = base + t;
      if (e < ne) {
        double U00 = 0.0, U01 = 0.0, U10 = 0.0, U11 = 0.0, V00 = 0.0, V01 = 0.0, V10 = 0.0, V11 = 0.0;
#pragma unroll
        for (int n = 0; n < NN; ++n) {
          const double g0 = sTab[(q * NN + n) * 2 + 0], g1 = sTab[(q * NN + n) * 2 + 1];
          U00 += L.u[t][2 * n] * g0;
          U01 += L.u[t][2 * n] * g1;
          U10 += L.u[t][2 * n + 1] * g0;
          U11 += L.u[t][2 * n + 1] * g1;
          if (AUX == 1) {
            V00 += L.v[t][2 * n] * g0;
            V01 += L.v[t][2 * n] * g1;
            V10 += L.v[t][2 * n + 1] * g0;
            V11 += L.v[t][2 * n + 1] * g1;
          }
        }
        const double K00 = L.geo[t][0], K01 = L.geo[t][1], K10 = L.geo[t][2], K11 = L.geo[t][3];
        double e0 = K00 * U00 + K01 * U01;
        double e1 = K10 * U10 + K11 * U11;
        double e2 = (K10 * U00 + K11 * U01) + (K00 * U10 + K01 * U11);
        if (has_eps) {
          const size_t o = size_t(5 + 3 * q) * ne + e;
          e0 -= d.egeo[o];
          e1 -= d.egeo[o + ne];
          e2 -= d.egeo[o + 2 * size_t(ne)];
        }
        double sig[3], D[6];
        mc_point<true>(k, e0, e1, e2, sig, D);
        const double w = sW[q] * L.geo[t][4];
        if (AUX == 1) {
          const double v0 = K00 * V00 + K01 * V01;
          const double v1 = K10 * V10 + K11 * V11;
          const double v2 = (K10 * V00 + K11 * V01) + (K00 * V10 + K01 * V11);
          const double s0 = w * (D[0] * v0 + D[1] * v1 + D[2] * v2);
          const double s1 = w * (D[1] * v0 + D[3] * v1 + D[4] * v2);
          const double s2 = w * (D[2] * v0 + D[4] * v1 + D[5] * v2);
          L.P[t][q][0] = s0 * K00 + s2 * K10;
          L.P[t][q][1] = s0 * K01 + s2 * K11;
          L.P[t][q][2] = s2 * K00 + s1 * K10;
          L.P[t][q][3] = s2 * K01 + s1 * K11;
        } else {
#pragma unroll
          for (int i = 0; i < 6; ++i) L.D[t][q][i] = w * D[i];
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int i = lane; i < kMcEpw * NL; i += 64) {
      const int t = i / NL, rr = i - t * NL;
      const int e = base + t;
      if (e >= ne) continue;
      const int n = rr >> 1, comp = rr & 1;
      double s = 0.0;
      for (int q = 0; q < nq; ++q) {
        const double g0 = sTab[(q * NN + n) * 2 + 0], g1 = sTab[(q * NN + n) * 2 + 1];
        if (AUX == 1) {
          s += g0 * L.P[t][q][2 * comp] + g1 * L.P[t][q][2 * comp + 1];
        } else {
          const double gx = L.geo[t][0] * g0 + L.geo[t][1] * g1;
          const double gy = L.geo[t][2] * g0 + L.geo[t][3] * g1;
          const double* Dq = L.D[t][q];
          // b = (gx, 0, gy) for an x component, (0, gy, gx) for a y component
          s += comp == 0 ? Dq[0] * gx * gx + 2.0 * Dq[2] * gx * gy + Dq[5] * gy * gy
                         : Dq[3] * gy * gy + 2.0 * Dq[4] * gx * gy + Dq[5] * gx * gx;
        }
      }
      gbuf[size_t(e) * 32 + rr] = s;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
}

template <int NN, int AUX>
int launch_mc_aux(fne_plan* plan, const double* x, const double* dvals, const double* v, double* out,
                  hipStream_t stream) {
  const fne_plan_desc& d = plan->d;
  int nb = (d.n_elem + kMcWaves * kMcEpw - 1) / (kMcWaves * kMcEpw);
  if (nb > kEnergyBlocks) nb = kEnergyBlocks;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL((mc_aux_kernel<NN, AUX>), dim3(nb), dim3(kBlock), 0, stream, d, plan->mode, x, dvals, v,
                     plan->mc_g);
  FNE_LAUNCH_CHECK();
  FneCall c{};
  c.f = nullptr;
  c.g = out;
  c.precision = FNE_FP64;
  c.stream = stream;
  return fne_launch_row_gather(plan, true, false, c);
}

int mc_aux(fne_plan* plan, int aux, const double* x, const double* dvals, const double* v, double* out,
           fne_stream_t stream) {
  if (!plan || !x || !out || (aux == 1 && (!v || v == out))) {
    fne_set_error("invalid arguments");
    return FNE_EINVAL;
  }
  const fne_plan_desc& d = plan->d;
  if (d.kind != FNE_KIND_MOHR_COULOMB) {
    fne_set_error("the matrix-free Hessian diagonal is implemented for the Mohr-Coulomb family");
    return FNE_EUNSUPPORTED;
  }
  if (!plan->mc_g || !d.inc_ptr || !d.inc) {
    fne_set_error("the plan needs FNE_FLAG_ROW_GATHER (row incidences and element scratch)");
    return FNE_EINVAL;
  }
  if (d.n_total > 0 && !dvals) {
    fne_set_error("Dirichlet values missing");
    return FNE_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  if (d.n_loc == 12) return aux == 1 ? launch_mc_aux<6, 1>(plan, x, dvals, v, out, s)
                                     : launch_mc_aux<6, 2>(plan, x, dvals, v, out, s);
  return aux == 1 ? launch_mc_aux<15, 1>(plan, x, dvals, v, out, s) : launch_mc_aux<15, 2>(plan, x, dvals, v, out, s);
}

}  // namespace

extern "C" int fne_hvp(fne_plan* plan, const double* x, const double* dvals, const double* v, double* y,
                       fne_stream_t stream) {
  if (!plan || !x || !v || !y || v == y) {
    fne_set_error("fne_hvp: invalid arguments");
    return FNE_EINVAL;
  }
  const int kind = plan->d.kind;
  if (kind == FNE_KIND_MOHR_COULOMB) return mc_aux(plan, 1, x, dvals, v, y, stream);
  if (plan->d.n_total > 0 && !dvals) {
    fne_set_error("Dirichlet values missing");
    return FNE_EINVAL;
  }
  if (kind == FNE_KIND_PLAPLACE || kind == FNE_KIND_GINZBURG_LANDAU)
    return fne_hvp_p1(plan, x, dvals, v, y, (hipStream_t)stream);
  if (kind == FNE_KIND_NEOHOOKEAN) return fne_hvp_tet(plan, x, dvals, v, y, (hipStream_t)stream);
  fne_set_error("fne_hvp: unknown energy family");
  return FNE_EUNSUPPORTED;
}

extern "C" int fne_hessian_diag(fne_plan* plan, const double* x, const double* dvals, double* diag,
                                fne_stream_t stream) {
  return mc_aux(plan, 2, x, dvals, nullptr, diag, stream);
}

int fne_launch_row_gather(fne_plan* plan, bool want_g, bool want_h, const FneCall& c) {
  const fne_plan_desc& d = plan->d;
  const bool fp32 = c.precision == FNE_FP32;
  const int rows_per_block = (kBlock / mc_row_lanes(d.n_loc)) * mc_row_batch(d.n_loc);
  const dim3 grid((d.n_free + rows_per_block - 1) / rows_per_block);
  const int cap = (d.max_row_len + 1) & ~1;
  const size_t shm = want_h ? size_t(rows_per_block) * cap * (fp32 ? 4 : 8) : 0;
  if (shm > plan->max_lds || d.max_row_len > 254) {
    fne_set_error("row gather: CSR rows too long for the LDS row copies");
    return FNE_EINVAL;
  }
  if (d.n_free == 0 || (!want_g && !want_h)) return FNE_OK;
  const void* hb = plan->mc_h;
  const double* gb = plan->mc_g;
#define RG(NL, HT, G, H)                                                                                   \
  hipLaunchKernelGGL((mc_row_gather_kernel<NL, HT, G, H>), grid, dim3(kBlock), shm, c.stream, d, (const HT*)hb, \
                     gb, c.f, c.g, (HT*)c.vals)
  if (d.n_loc == 12) {
    if (!want_h) RG(12, double, true, false);
    else if (fp32) { if (want_g) RG(12, float, true, true); else RG(12, float, false, true); }
    else { if (want_g) RG(12, double, true, true); else RG(12, double, false, true); }
  } else if (d.n_loc == 30) {
    if (!want_h) RG(30, double, true, false);
    else if (fp32) { if (want_g) RG(30, float, true, true); else RG(30, float, false, true); }
    else { if (want_g) RG(30, double, true, true); else RG(30, double, false, true); }
  } else {
    fne_set_error("row gather supports elements with 12 or 30 DOFs");
    return FNE_EINVAL;
  }
#undef RG
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

size_t fne_mc_scratch_bytes(const fne_plan_desc& d, size_t* gbytes) {
  const size_t nl = size_t(d.n_loc);
  if (gbytes) *gbytes = size_t(d.n_elem) * 32 * sizeof(double);
  return size_t(d.n_elem) * nl * nl * sizeof(double);
}

int fne_launch_mc(fne_plan* plan, FneOp op, const FneCall& c) {
  if (!plan->mc_g || !plan->mc_h) {
    fne_set_error("Mohr-Coulomb element scratch missing");
    return FNE_EINVAL;
  }
  switch (plan->d.n_loc) {
    case 12:
      return launch_mc<6>(plan, op, c);
    case 30:
      return launch_mc<15>(plan, op, c);
    default:
      fne_set_error("Mohr-Coulomb plans need n_loc = 12 (P2) or 30 (P4)");
      return FNE_EINVAL;
  }
}
