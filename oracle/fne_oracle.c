/*
 * CPU oracle, C restatement -- TEST INFRASTRUCTURE ONLY (the CPU baseline of bench.py and
 * tests/test_oracle_c.py load it; the product package never does).
 *
 * P1 triangle assembly of the p-Laplace and Ginzburg-Landau energies: per element the energy,
 * gradient and 3x3 Hessian (the formulas of oracle/elements.py, which restate
 * plaplace/jax_petsc/reordered_element_assembler.py:16-63 and
 * ginzburg_landau/jax_petsc/reordered_element_assembler.py:18-97), then the sum into the
 * gradient and the owned CSR values (the np.add.at scatter of
 * core/petsc/reordered_element_base.py:1749-1757).  OpenMP over elements, then over rows:
 * phase 1 writes each element's quantities to scratch, phase 2 adds, for every free row, its
 * incident element rows in element order (a fixed order, so gradient and Hessian do not depend
 * on the thread count).  The energy is summed per thread in element order and the thread sums
 * in thread order, so it depends on the thread count at the last bits.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* family 1: p-Laplace (param = p), family 2: Ginzburg-Landau (param = eps, ip[3][nq], w[nq]) */
static void p1_element(int family, double param, const double* v, const double dvx[3], const double dvy[3],
                       double vol, const double* ip, const double* w, int nq, double* e, double g[3], double h[9]) {
  const double fx = v[0] * dvx[0] + v[1] * dvx[1] + v[2] * dvx[2];
  const double fy = v[0] * dvy[0] + v[1] * dvy[1] + v[2] * dvy[2];
  const double s = fx * fx + fy * fy;
  if (family == 1) {
    const double p = param;
    /* E = vol/p s^(p/2); g_a = vol s^(p/2-1) t_a; H_ab = vol (s^(p/2-1) K_ab + (p-2) s^(p/2-2) t_a t_b) */
    const double sp = pow(s, 0.5 * p);
    *e = vol * sp / p;
    const double c1 = s > 0.0 ? pow(s, 0.5 * p - 1.0) : (p == 2.0 ? 1.0 : 0.0);
    const double c2 = s > 0.0 ? (p - 2.0) * pow(s, 0.5 * p - 2.0) : 0.0;
    double t[3];
    for (int a = 0; a < 3; ++a) {
      t[a] = fx * dvx[a] + fy * dvy[a];
      g[a] = vol * c1 * t[a];
    }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) h[3 * a + b] = vol * (c1 * (dvx[a] * dvx[b] + dvy[a] * dvy[b]) + c2 * t[a] * t[b]);
    return;
  }
  const double eps = param;
  double pot_e = 0.0, pot_g[3] = {0.0, 0.0, 0.0}, m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = 0; q < nq; ++q) {
    const double i0 = ip[q], i1 = ip[nq + q], i2 = ip[2 * nq + q];
    const double uq = v[0] * i0 + v[1] * i1 + v[2] * i2;
    const double r = uq * uq - 1.0;
    pot_e += r * r * w[q];
    const double gq = r * uq * w[q];
    pot_g[0] += gq * i0;
    pot_g[1] += gq * i1;
    pot_g[2] += gq * i2;
    const double cq = (3.0 * uq * uq - 1.0) * w[q];
    const double iv[3] = {i0, i1, i2};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) m[3 * a + b] += cq * iv[a] * iv[b];
  }
  *e = (0.5 * eps * s + 0.25 * pot_e) * vol;
  for (int a = 0; a < 3; ++a) g[a] = vol * (eps * (fx * dvx[a] + fy * dvy[a]) + pot_g[a]);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) h[3 * a + b] = vol * (eps * (dvx[a] * dvx[b] + dvy[a] * dvy[b]) + m[3 * a + b]);
}

int fne_oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/*
 * elems[3 n_elem] node ids; dvx, dvy [3 n_elem]; vol [n_elem]; u_full [n_nodes] nodal values;
 * inc_ptr [n_free + 1], inc [n_inc]: for free row r the items 3 e + a (element e, local row a),
 * ascending; pos [9 n_elem]: CSR position of entry (a, b) of element e, or -1 (constrained row
 * or column); scratch [13 n_elem] doubles.  Outputs: energy (sum over all elements minus
 * f . u_free when f != NULL), grad [n_free] (minus f), vals [nnz] (zeroed here).
 */
int fne_oracle_p1_assemble(int family, double param, int n_elem, const int32_t* elems, const double* dvx,
                           const double* dvy, const double* vol, const double* u_full, const double* ip,
                           const double* w, int nq, int n_free, const int32_t* free_nodes, const double* f,
                           const int64_t* inc_ptr, const int64_t* inc, const int64_t* pos, int64_t nnz,
                           double* scratch, double* energy, double* grad, double* vals) {
  const int nth = fne_oracle_threads();
  double* esums = (double*)calloc((size_t)nth, sizeof(double));
  if (!esums) return -3;
#pragma omp parallel num_threads(nth)
  {
    double esum = 0.0;
#pragma omp for schedule(static)
    for (int e = 0; e < n_elem; ++e) {
      double v[3], ge[3], he[9], ee;
      for (int a = 0; a < 3; ++a) v[a] = u_full[elems[3 * e + a]];
      p1_element(family, param, v, dvx + 3 * e, dvy + 3 * e, vol[e], ip, w, nq, &ee, ge, he);
      double* sc = scratch + 13 * (int64_t)e;
      sc[0] = ge[0];
      sc[1] = ge[1];
      sc[2] = ge[2];
      for (int k = 0; k < 9; ++k) sc[3 + k] = he[k];
      sc[12] = ee;
      esum += ee;
    }
#pragma omp for schedule(static)
    for (int64_t k = 0; k < nnz; ++k) vals[k] = 0.0;
#pragma omp for schedule(static)
    for (int r = 0; r < n_free; ++r) {
      double gr = 0.0;
      for (int64_t it = inc_ptr[r]; it < inc_ptr[r + 1]; ++it) {
        const int64_t e = inc[it] / 3;
        const int a = (int)(inc[it] % 3);
        const double* sc = scratch + 13 * e;
        gr += sc[a];
        for (int b = 0; b < 3; ++b) {
          const int64_t pk = pos[9 * e + 3 * a + b];
          if (pk >= 0) vals[pk] += sc[3 + 3 * a + b];
        }
      }
      grad[r] = f ? gr - f[r] : gr;
    }
#ifdef _OPENMP
    esums[omp_get_thread_num()] = esum;
#else
    esums[0] = esum;
#endif
  }
  double etot = 0.0;
  for (int t = 0; t < nth; ++t) etot += esums[t];
  free(esums);
  if (f) {
    for (int r = 0; r < n_free; ++r) etot -= f[r] * u_full[free_nodes[r]];
  }
  *energy = etot;
  return 0;
}
