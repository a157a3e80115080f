// Per-node MLP for the learned warm start: BF16-MFMA forward, backward and Adam.
//
// Network: n_layers linear layers y = W x + b (W: [out][in] BF16, b: FP32), tanh after every
// layer but the last.  Activations, gradients and the optimiser state are FP32; every matrix
// product runs on v_mfma_f32_16x16x16bf16_1k with BF16 operands and FP32 accumulation.
//
// One GEMM kernel serves every product: C[m][n] = sum_k A(m,k) B(k,n) with arbitrary element
// strides for A and B (so X W^T, dZ^T H and dZ W are the same kernel), a 64x64 block tile,
// 16-deep K steps staged through LDS as BF16, four wavefronts each owning a 32x32 sub-tile
// (2x2 MFMA tiles).  Long reductions (the weight gradient sums over all nodes) are split over
// grid.z into FP32 partials that a second kernel adds in split order, so every result is
// bitwise reproducible.
#include "fne_common.h"

namespace {

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(unsigned(h) << 16); }

__device__ __forceinline__ unsigned short load_bf(const float* p, size_t i) { return f2bf(p[i]); }
__device__ __forceinline__ unsigned short load_bf(const unsigned short* p, size_t i) { return p[i]; }

enum MlpEpi : int { kEpiNone = 0, kEpiTanh = 1, kEpiDTanh = 2 };

struct GemmArgs {
  int M, N, K;
  const void* A;
  long sam, sak;
  const void* B;
  long sbk, sbn;
  float* C;       // [M][ldc] (final) or partials [splits][M][N]
  long ldc;
  const float* bias;  // [N] or null
  const float* H;     // kEpiDTanh: C *= (1 - H^2), H [M][ldh]
  long ldh;
  int epi;
  int kchunk;         // K range per grid.z slice
  int split;          // 1: write raw partials
};

constexpr int kT = 64;   // block tile
constexpr int kKs = 16;  // K step

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(GemmArgs g) {
  __shared__ unsigned short As[kT][kKs + 4];
  __shared__ unsigned short Bs[kKs][kT + 4];
  const TA* A = reinterpret_cast<const TA*>(g.A);
  const TB* B = reinterpret_cast<const TB*>(g.B);
  const int m0 = blockIdx.y * kT, n0 = blockIdx.x * kT;
  const int kbeg = blockIdx.z * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wy = wave >> 1, wx = wave & 1;
  mlp_f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = mlp_f4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kbeg; k0 < kend; k0 += kKs) {
    // 64x16 A elements and 16x64 B elements, 4 of each per thread
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + 256 * r;
      const int am = idx / kKs, ak = idx % kKs;  // A tile [m][k]
      const int gm = m0 + am, gk = k0 + ak;
      As[am][ak] = (gm < g.M && gk < kend) ? load_bf(A, size_t(gm) * g.sam + size_t(gk) * g.sak) : 0;
      const int bk = idx / kT, bn = idx % kT;  // B tile [k][n]
      const int gk2 = k0 + bk, gn = n0 + bn;
      Bs[bk][bn] = (gn < g.N && gk2 < kend) ? load_bf(B, size_t(gk2) * g.sbk + size_t(gn) * g.sbn) : 0;
    }
    __syncthreads();
    const int kq = 4 * (lane >> 4);
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
      const int row = wy * 32 + ti * 16 + (lane & 15);
      mlp_s4 a;
      a[0] = (short)As[row][kq + 0];
      a[1] = (short)As[row][kq + 1];
      a[2] = (short)As[row][kq + 2];
      a[3] = (short)As[row][kq + 3];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) {
        const int col = wx * 32 + tj * 16 + (lane & 15);
        mlp_s4 b;
        b[0] = (short)Bs[kq + 0][col];
        b[1] = (short)Bs[kq + 1][col];
        b[2] = (short)Bs[kq + 2][col];
        b[3] = (short)Bs[kq + 3][col];
        acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[ti][tj], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wy * 32 + ti * 16 + 4 * (lane >> 4) + r;
        const int gn = n0 + wx * 32 + tj * 16 + (lane & 15);
        if (gm >= g.M || gn >= g.N) continue;
        float v = acc[ti][tj][r];
        if (g.split) {
          g.C[(size_t(blockIdx.z) * g.M + gm) * g.N + gn] = v;
          continue;
        }
        if (g.bias) v += g.bias[gn];
        if (g.epi == kEpiTanh) v = tanhf(v);
        if (g.epi == kEpiDTanh) {
          const float h = g.H[size_t(gm) * g.ldh + gn];
          v *= 1.0f - h * h;
        }
        g.C[size_t(gm) * g.ldc + gn] = v;
      }
}

// out[i] = sum_s part[s][i] in split order
__global__ __launch_bounds__(256) void mlp_split_reduce_kernel(const float* __restrict__ part, int splits, long n,
                                                               float* __restrict__ out) {
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int k = 0; k < splits; ++k) s += part[size_t(k) * n + i];
  out[i] = s;
}

// column sums of X [rows][cols] over row chunks of 1024 -> part[chunk][cols]
__global__ __launch_bounds__(256) void mlp_colsum_kernel(const float* __restrict__ X, long rows, int cols,
                                                         float* __restrict__ part) {
  const int c = blockIdx.y * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  const long r0 = long(blockIdx.x) * 1024;
  const long r1 = r0 + 1024 < rows ? r0 + 1024 : rows;
  float s = 0.f;
  for (long r = r0; r < r1; ++r) s += X[size_t(r) * cols + c];
  part[size_t(blockIdx.x) * cols + c] = s;
}

__global__ __launch_bounds__(256) void mlp_adam_kernel(long n, float* __restrict__ w, unsigned short* __restrict__ wb,
                                                       const float* __restrict__ grad, float* __restrict__ m,
                                                       float* __restrict__ v, float lr, float b1, float b2,
                                                       float eps, float c1, float c2, float wd) {
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = grad[i] + wd * w[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

__global__ __launch_bounds__(256) void mlp_to_bf16_kernel(long n, const float* __restrict__ w,
                                                          unsigned short* __restrict__ wb) {
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = f2bf(w[i]);
}

template <typename TA, typename TB>
int run_gemm(GemmArgs g, int splits, hipStream_t s, int* used_splits = nullptr) {
  if (g.M <= 0 || g.N <= 0) return FNE_OK;
  if (splits < 1) splits = 1;
  g.kchunk = ((g.K + splits - 1) / splits + kKs - 1) / kKs * kKs;
  if (g.kchunk < kKs) g.kchunk = kKs;
  splits = (g.K + g.kchunk - 1) / g.kchunk;
  if (splits < 1) splits = 1;
  g.split = splits > 1 ? 1 : g.split;
  if (used_splits) *used_splits = splits;
  dim3 grid((g.N + kT - 1) / kT, (g.M + kT - 1) / kT, splits);
  hipLaunchKernelGGL((mlp_gemm_kernel<TA, TB>), grid, dim3(256), 0, s, g);
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

bool net_ok(const fne_mlp_desc* net) {
  if (!net || net->n_layers < 1 || net->n_layers > FNE_MLP_MAX_LAYERS) return false;
  for (int l = 0; l < net->n_layers; ++l)
    if (net->dims[l] < 1 || net->dims[l + 1] < 1 || !net->w[l] || !net->b[l]) return false;
  return true;
}

long max_width(const fne_mlp_desc* net) {
  long w = 0;
  for (int l = 0; l <= net->n_layers; ++l) w = net->dims[l] > w ? net->dims[l] : w;
  return w;
}

constexpr int kWgradSplits = 64;

}  // namespace

extern "C" {

int64_t fne_mlp_workspace_bytes(const fne_mlp_desc* net, int64_t n) {
  if (!net_ok(net) || n < 0) return FNE_EINVAL;
  const long w = max_width(net);
  const long chunks = (n + 1023) / 1024;
  // two dZ buffers, weight-gradient split partials, bias-gradient chunk partials
  return int64_t(sizeof(float)) * (2 * n * w + long(kWgradSplits) * w * w + chunks * w);
}

int fne_mlp_forward(const fne_mlp_desc* net, int64_t n, const float* x, float* y, float* acts,
                    fne_stream_t stream) {
  if (!net_ok(net) || n < 0 || !x || !y) {
    fne_set_error("fne_mlp_forward: invalid arguments");
    return FNE_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const float* in = x;
  size_t off = 0;
  for (int l = 0; l < net->n_layers; ++l) {
    const int din = net->dims[l], dout = net->dims[l + 1];
    const bool last = l == net->n_layers - 1;
    float* out = last ? y : acts + off;
    if (!last && !acts) {
      fne_set_error("fne_mlp_forward: activation buffer required");
      return FNE_EINVAL;
    }
    GemmArgs g{};
    g.M = int(n);
    g.N = dout;
    g.K = din;
    g.A = in;
    g.sam = din;
    g.sak = 1;
    g.B = net->w[l];  // B(k, n) = W[n][k]
    g.sbk = 1;
    g.sbn = din;
    g.C = out;
    g.ldc = dout;
    g.bias = net->b[l];
    g.epi = last ? kEpiNone : kEpiTanh;
    const int rc = run_gemm<float, unsigned short>(g, 1, s);
    if (rc) return rc;
    if (!last) off += size_t(n) * dout;
    in = out;
  }
  return FNE_OK;
}

int fne_mlp_backward(const fne_mlp_desc* net, int64_t n, const float* x, const float* acts, const float* dy,
                     float* const* dw, float* const* db, void* workspace, int64_t ws_bytes, fne_stream_t stream) {
  if (!net_ok(net) || n < 1 || !x || !dy || !dw || !db || !workspace) {
    fne_set_error("fne_mlp_backward: invalid arguments");
    return FNE_EINVAL;
  }
  if (ws_bytes < fne_mlp_workspace_bytes(net, n)) {
    fne_set_error("fne_mlp_backward: workspace too small");
    return FNE_EINVAL;
  }
  if (net->n_layers > 1 && !acts) {
    fne_set_error("fne_mlp_backward: activation buffer required");
    return FNE_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const long w = max_width(net);
  float* dz0 = reinterpret_cast<float*>(workspace);
  float* dz1 = dz0 + size_t(n) * w;
  float* part = dz1 + size_t(n) * w;
  float* cpart = part + size_t(kWgradSplits) * w * w;
  const long chunks = (n + 1023) / 1024;
  // offsets of each hidden activation in acts
  size_t offs[FNE_MLP_MAX_LAYERS + 1];
  offs[0] = 0;
  for (int l = 0; l + 1 < net->n_layers; ++l) offs[l + 1] = offs[l] + size_t(n) * net->dims[l + 1];
  const float* dz = dy;  // gradient w.r.t. the pre-activation of layer l
  float* bufs[2] = {dz0, dz1};
  for (int l = net->n_layers - 1; l >= 0; --l) {
    const int din = net->dims[l], dout = net->dims[l + 1];
    const float* hin = l == 0 ? x : acts + offs[l - 1];
    // dW[j][k] = sum_i dz[i][j] hin[i][k]  (M = dout, N = din, K = n): split over nodes
    GemmArgs g{};
    g.M = dout;
    g.N = din;
    g.K = int(n);
    g.A = dz;
    g.sam = 1;
    g.sak = dout;
    g.B = hin;
    g.sbk = din;
    g.sbn = 1;
    g.C = part;
    g.ldc = din;
    g.split = 1;
    int splits = 1;
    int rc = run_gemm<float, float>(g, kWgradSplits, s, &splits);
    if (rc) return rc;
    const long nw = long(dout) * din;
    hipLaunchKernelGGL(mlp_split_reduce_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, part, splits, nw, dw[l]);
    FNE_LAUNCH_CHECK();
    // db[j] = sum_i dz[i][j]
    hipLaunchKernelGGL(mlp_colsum_kernel, dim3(chunks, (dout + 255) / 256), dim3(256), 0, s, dz, long(n), dout,
                       cpart);
    FNE_LAUNCH_CHECK();
    hipLaunchKernelGGL(mlp_split_reduce_kernel, dim3((dout + 255) / 256), dim3(256), 0, s, cpart, int(chunks),
                       long(dout), db[l]);
    FNE_LAUNCH_CHECK();
    if (l == 0) break;
    // dz_prev[i][k] = (sum_j dz[i][j] W[j][k]) * (1 - h[i][k]^2)
    float* out = bufs[l & 1];
    GemmArgs h{};
    h.M = int(n);
    h.N = din;
    h.K = dout;
    h.A = dz;
    h.sam = dout;
    h.sak = 1;
    h.B = net->w[l];  // B(k=j, n=k) = W[j][k]
    h.sbk = din;
    h.sbn = 1;
    h.C = out;
    h.ldc = din;
    h.H = hin;
    h.ldh = din;
    h.epi = kEpiDTanh;
    rc = run_gemm<float, unsigned short>(h, 1, s);
    if (rc) return rc;
    dz = out;
  }
  return FNE_OK;
}

int fne_mlp_adam(int64_t n, float* w, uint16_t* w_bf16, const float* grad, float* m, float* v, float lr,
                 float beta1, float beta2, float eps, float weight_decay, int32_t step, fne_stream_t stream) {
  if (n < 0 || !w || !grad || !m || !v || step < 1) {
    fne_set_error("fne_mlp_adam: invalid arguments");
    return FNE_EINVAL;
  }
  if (n == 0) return FNE_OK;
  const float c1 = 1.f - powf(beta1, float(step));
  const float c2 = 1.f - powf(beta2, float(step));
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, long(n), w,
                     (unsigned short*)w_bf16, grad, m, v, lr, beta1, beta2, eps, c1, c2, weight_decay);
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

int fne_mlp_to_bf16(int64_t n, const float* w, uint16_t* w_bf16, fne_stream_t stream) {
  if (n < 0 || !w || !w_bf16) return FNE_EINVAL;
  if (n == 0) return FNE_OK;
  hipLaunchKernelGGL(mlp_to_bf16_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, long(n), w,
                     (unsigned short*)w_bf16);
  FNE_LAUNCH_CHECK();
  return FNE_OK;
}

}  // extern "C"
