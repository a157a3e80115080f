// Golden-section line search as a device state machine (core/petsc/minimizers.py:14-64).
//
// The Newton driver's line search minimises J(x + alpha p) over [lo, hi] by golden-section
// search; each step needs one energy, and the host's only work between energies is comparing
// two numbers.  Here that comparison runs in a one-thread kernel after each energy: the state
// holds the bracket, the two interior points and their values, the best point so far and the
// next alpha to evaluate, so the host enqueues trial-point / energy / step triples back to back
// and synchronises once at the end.  The arithmetic and the order of every comparison are the
// reference's (same points, same tie-breaking), so the chosen alpha is the one the host
// search would choose from the same energies.
#include "fne_common.h"

#include <algorithm>
#include <cmath>

namespace {

// state slots (doubles)
enum : int {
  kLo = 0,
  kHi,
  kC,
  kD,
  kFLo,
  kFHi,
  kFC,
  kFD,
  kBestA,
  kBestF,
  kAlpha,  // point whose energy the next step consumes
  kPhase,
  kEvals,
  kTol,
  kMid,
  kGoldenLen = 16
};

// phases: 0..3 initial points lo, hi, c, d; 4 new c; 5 new d; 6 midpoint; 7 done
__device__ void golden_next(double* st) {
  const double lo = st[kLo], hi = st[kHi];
  if (hi - lo > st[kTol]) {
    if (st[kFC] < st[kFD]) {
      const double nhi = st[kD];
      const double c = st[kC];
      st[kHi] = nhi;
      st[kD] = c;
      st[kFD] = st[kFC];
      st[kC] = lo + nhi - c;
      st[kAlpha] = st[kC];
      st[kPhase] = 4.0;
    } else {
      const double nlo = st[kC];
      const double d = st[kD];
      st[kLo] = nlo;
      st[kC] = d;
      st[kFC] = st[kFD];
      st[kD] = nlo + hi - d;
      st[kAlpha] = st[kD];
      st[kPhase] = 5.0;
    }
  } else {
    st[kMid] = 0.5 * (lo + hi);
    st[kAlpha] = st[kMid];
    st[kPhase] = 6.0;
  }
}

__device__ __forceinline__ void golden_probe(double* st, double a, double f) {
  if (f < st[kBestF]) {
    st[kBestA] = a;
    st[kBestF] = f;
  }
}

__global__ void golden_step_kernel(double* __restrict__ st, const double* __restrict__ f_new) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int phase = int(st[kPhase]);
  if (phase >= 7) return;
  double v = f_new[0];
  if (!isfinite(v)) v = INFINITY;  // the host search maps non-finite energies to +inf
  switch (phase) {
    case 0:
      st[kFLo] = v;
      st[kBestA] = st[kLo];
      st[kBestF] = v;
      st[kAlpha] = st[kHi];
      st[kPhase] = 1.0;
      break;
    case 1: {
      st[kFHi] = v;
      const double gamma = 0.5 + sqrt(5.0) / 2.0;
      const double d = (st[kHi] - st[kLo]) / gamma + st[kLo];
      st[kD] = d;
      st[kC] = st[kLo] + st[kHi] - d;
      st[kAlpha] = st[kC];
      st[kPhase] = 2.0;
      break;
    }
    case 2:
      st[kFC] = v;
      st[kAlpha] = st[kD];
      st[kPhase] = 3.0;
      break;
    case 3:
      st[kFD] = v;
      st[kEvals] = 4.0;
      // best over (hi, c, d) after lo, in the host's order
      golden_probe(st, st[kHi], st[kFHi]);
      golden_probe(st, st[kC], st[kFC]);
      golden_probe(st, st[kD], st[kFD]);
      golden_next(st);
      break;
    case 4:  // new c after hi = d
      st[kFC] = v;
      golden_probe(st, st[kD], st[kFD]);
      golden_probe(st, st[kC], st[kFC]);
      st[kEvals] += 1.0;
      golden_next(st);
      break;
    case 5:  // new d after lo = c
      st[kFD] = v;
      golden_probe(st, st[kC], st[kFC]);
      golden_probe(st, st[kD], st[kFD]);
      st[kEvals] += 1.0;
      golden_next(st);
      break;
    case 6:
      st[kEvals] += 1.0;
      golden_probe(st, st[kMid], v);
      st[kAlpha] = st[kBestA];
      st[kPhase] = 7.0;
      break;
    default:
      break;
  }
}

// trial = x + alpha p with alpha from the state (skipped once the search is done)
__global__ __launch_bounds__(256) void golden_trial_kernel(int n, const double* __restrict__ x,
                                                           const double* __restrict__ p,
                                                           const double* __restrict__ st,
                                                           double* __restrict__ trial) {
  if (st[kPhase] >= 7.0) return;
  const double a = st[kAlpha];
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) trial[i] = x[i] + a * p[i];
}

}  // namespace

extern "C" int fne_golden_state_len(void) { return kGoldenLen; }

extern "C" int fne_golden_step(double* state, const double* f_new, fne_stream_t stream) {
  if (!state || !f_new) {
    fne_set_error("fne_golden_step: null pointer");
    return FNE_EINVAL;
  }
  hipLaunchKernelGGL(golden_step_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, state, f_new);
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

extern "C" int fne_golden_trial(int32_t n, const double* x, const double* p, const double* state, double* trial,
                                fne_stream_t stream) {
  if (n < 0 || !x || !p || !state || !trial) {
    fne_set_error("fne_golden_trial: invalid arguments");
    return FNE_EINVAL;
  }
  if (n == 0) return FNE_OK;
  const int nb = int(std::min<int64_t>(1024, (int64_t(n) + 255) / 256));
  hipLaunchKernelGGL(golden_trial_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, n, x, p, state, trial);
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}
