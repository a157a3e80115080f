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
