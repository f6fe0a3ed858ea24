// Training of the learned limiter: hand-written FP32 kernels for the forward pass, the
// loss, the backward pass, the weight-gradient reductions and Adam, behind the C ABI of
// include/idefix_amd.h (idfx_mlp_*).  No counterpart in the reference.
//
// Parameters are packed as in oracle/mlp.py: W1 [48][64], b1 [64], W2 [64][64], b2 [64],
// W3 [64][16], b3 [16] (row-major [in][out]), 8336 floats.  A batch row of X holds
// MLP_ROW = 72 floats: the 48 features, the interface's 16 PPM weights wp and 8 per-variable
// loss weights om; TAU [B][8] holds the normalised targets (t - c) / s.  The network's 16
// outputs z3 shift both PPM states of a variable, wL = wpL + z3[2v], wR = wpR - z3[2v].  Loss = mean over rows of
// sum_v om_v^2 (eL_v^2 + eR_v^2) with eL = x2 + wL (x3 - x2) - tau,
// eR = x3 - wR (x3 - x2) - tau.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/idefix_amd.h"

namespace {

constexpr int NF = 48, NH = 64, NO = 16;
constexpr int P_W1 = 0, P_B1 = P_W1 + NF * NH, P_W2 = P_B1 + NH, P_B2 = P_W2 + NH * NH,
              P_W3 = P_B2 + NH, P_B3 = P_W3 + NH * NO, P_N = P_B3 + NO;
constexpr float GELU_K = 1.702f;
constexpr int ROW = 72, ROW_WP = 48, ROW_OM = 64;  // batch row: features | wp | om
constexpr int WG_ROWS = 512;  // rows per block of the weight-gradient reduction

__device__ __forceinline__ float sig(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ __forceinline__ float gelu(float z) { return z * sig(GELU_K * z); }
__device__ __forceinline__ float gelu_d(float z) {
  const float s = sig(GELU_K * z);
  return s + GELU_K * z * s * (1.0f - s);
}

// Z = A W + b, H = gelu(Z) (or Z when act == 0); A has K columns in rows of lda floats;
// one thread per (row, output unit).
__global__ __launch_bounds__(256) void k_dense_fwd(const float *A, int K, int lda, const float *W,
                                                   const float *b, int N, float *Z, float *H,
                                                   int act, int B) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)B * N) return;
  const int r = (int)(t / N), j = (int)(t % N);
  float z = b[j];
  for (int k = 0; k < K; k++) z += A[(long long)r * lda + k] * W[k * N + j];
  Z[t] = z;
  if (H) H[t] = act ? gelu(z) : z;
}

// Output layer: wL = wpL + z3[2v], wR = wpR - z3[2v], the row's loss, and dZ3 = dloss/dz3 (1/B included).
__global__ __launch_bounds__(256) void k_out_loss(const float *Z3, const float *X,
                                                  const float *TAU, int B, float *dZ3,
                                                  float *rowloss) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  float l = 0.0f;
  const float inv = 1.0f / (float)B;
  const float *xr = X + (long long)r * ROW;
  for (int v = 0; v < 8; v++) {
    const float x2 = xr[6 * v + 2], x3 = xr[6 * v + 3];
    const float d = x3 - x2;
    const float tau = TAU[(long long)r * 8 + v];
    const float om2 = xr[ROW_OM + v] * xr[ROW_OM + v];
    const float z = Z3[(long long)r * NO + 2 * v];
    const float wl = xr[ROW_WP + 2 * v] + z;
    const float wr = xr[ROW_WP + 2 * v + 1] - z;
    const float el = x2 + wl * d - tau;
    const float er = x3 - wr * d - tau;
    l += om2 * (el * el + er * er);
    dZ3[(long long)r * NO + 2 * v] = om2 * 2.0f * (el + er) * d * inv;
    dZ3[(long long)r * NO + 2 * v + 1] = 0.0f;
  }
  rowloss[r] = l;
}

// dZp = (dZ W^T) * gelu'(Zp); one thread per (row, input unit).
__global__ __launch_bounds__(256) void k_dense_bwd(const float *dZ, int N, const float *W,
                                                   int K, const float *Zp, float *dZp, int B) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)B * K) return;
  const int r = (int)(t / K), k = (int)(t % K);
  float s = 0.0f;
  for (int j = 0; j < N; j++) s += dZ[(long long)r * N + j] * W[k * N + j];
  dZp[t] = s * gelu_d(Zp[t]);
}

// Partial weight and bias gradients of one layer over WG_ROWS rows per block (A has K
// columns in rows of lda floats):
// out[blk][off + k N + n] = sum_r A[r][k] dZ[r][n] (k < K), out[blk][off + K N + n] =
// sum_r dZ[r][n].
__global__ __launch_bounds__(256) void k_wgrad(const float *A, int K, int lda, const float *dZ,
                                               int N, int B, float *partial, int off) {
  __shared__ float sa[16][NH + 1];
  __shared__ float sz[16][NH];
  const int total = (K + 1) * N;
  float acc[17];
#pragma unroll
  for (int m = 0; m < 17; m++) acc[m] = 0.0f;
  const int r0 = blockIdx.x * WG_ROWS;
  const int r1 = min(r0 + WG_ROWS, B);
  for (int rt = r0; rt < r1; rt += 16) {
    __syncthreads();
    for (int e = threadIdx.x; e < 16 * NH; e += 256) {
      const int rr = e / NH, c = e % NH;
      const int r = rt + rr;
      sa[rr][c] = (r < r1 && c < K) ? A[(long long)r * lda + c] : 0.0f;
      sz[rr][c] = (r < r1 && c < N) ? dZ[(long long)r * N + c] : 0.0f;
    }
    if (threadIdx.x < 16) sa[threadIdx.x][NH] = (rt + threadIdx.x < r1) ? 1.0f : 0.0f;
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 17; m++) {
      const int e = threadIdx.x + 256 * m;
      if (e < total) {
        const int k = e / N, n = e % N;
        const int kk = k < K ? k : NH;  // the bias row reads the ones column
        float s = acc[m];
        for (int rr = 0; rr < 16; rr++) s += sa[rr][kk] * sz[rr][n];
        acc[m] = s;
      }
    }
  }
#pragma unroll
  for (int m = 0; m < 17; m++) {
    const int e = threadIdx.x + 256 * m;
    if (e < total) partial[(long long)blockIdx.x * P_N + off + e] = acc[m];
  }
}

__global__ __launch_bounds__(256) void k_reduce_partials(const float *partial, int nblk,
                                                         float *G) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= P_N) return;
  float s = 0.0f;
  for (int b = 0; b < nblk; b++) s += partial[(long long)b * P_N + e];
  G[e] = s;
}

// Adam (Kingma & Ba 2015) with bias correction; step t is 1-based.
__global__ __launch_bounds__(256) void k_adam(float *P, const float *G, float *M, float *V, int n,
                                              float lr, float b1, float b2, float eps, float c1,
                                              float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float g = G[e];
  const float m = b1 * M[e] + (1.0f - b1) * g;
  const float v = b2 * V[e] + (1.0f - b2) * g * g;
  M[e] = m;
  V[e] = v;
  P[e] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

// Sum of the row losses into a device double (one block).
__global__ __launch_bounds__(256) void k_sum_loss(const float *rowloss, int B, double *out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int r = threadIdx.x; r < B; r += 256) s += rowloss[r];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0] / B;
}

unsigned nblocks(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct idfx_mlp {
  int cap = 0;
  long step = 0;
  std::vector<void *> allocs;
  float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;
  float *Z1, *H1, *Z2, *H2, *Z3, *dZ3, *dZ2, *dZ1, *partial, *rowloss;
  double *loss = nullptr;
};

namespace {

int mlp_alloc(idfx_mlp *h, float **p, long long n) {
  void *q = nullptr;
  if (hipMalloc(&q, sizeof(float) * (size_t)std::max(n, 1LL)) != hipSuccess) return IDFX_E_ALLOC;
  h->allocs.push_back(q);
  if (hipMemset(q, 0, sizeof(float) * (size_t)std::max(n, 1LL)) != hipSuccess) return IDFX_E_ALLOC;
  *p = static_cast<float *>(q);
  return IDFX_OK;
}

// forward + loss + backward + gradient into h->G (asynchronous on s)
int grad_enqueue(idfx_mlp *h, const float *X, const float *TAU, int B, hipStream_t s) {
  const float *P = h->P;
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblocks((long long)B * NH)), dim3(256), 0, s, X, NF, ROW,
                     P + P_W1, P + P_B1, NH, h->Z1, h->H1, 1, B);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblocks((long long)B * NH)), dim3(256), 0, s, h->H1, NH,
                     NH, P + P_W2, P + P_B2, NH, h->Z2, h->H2, 1, B);
  hipLaunchKernelGGL(k_dense_fwd, dim3(nblocks((long long)B * NO)), dim3(256), 0, s, h->H2, NH,
                     NH, P + P_W3, P + P_B3, NO, h->Z3, (float *)nullptr, 0, B);
  hipLaunchKernelGGL(k_out_loss, dim3(nblocks(B)), dim3(256), 0, s, h->Z3, X, TAU, B, h->dZ3,
                     h->rowloss);
  hipLaunchKernelGGL(k_dense_bwd, dim3(nblocks((long long)B * NH)), dim3(256), 0, s, h->dZ3, NO,
                     P + P_W3, NH, h->Z2, h->dZ2, B);
  hipLaunchKernelGGL(k_dense_bwd, dim3(nblocks((long long)B * NH)), dim3(256), 0, s, h->dZ2, NH,
                     P + P_W2, NH, h->Z1, h->dZ1, B);
  const int nblk = (B + WG_ROWS - 1) / WG_ROWS;
  hipLaunchKernelGGL(k_wgrad, dim3(nblk), dim3(256), 0, s, X, NF, ROW, h->dZ1, NH, B, h->partial,
                     P_W1);
  hipLaunchKernelGGL(k_wgrad, dim3(nblk), dim3(256), 0, s, h->H1, NH, NH, h->dZ2, NH, B,
                     h->partial, P_W2);
  hipLaunchKernelGGL(k_wgrad, dim3(nblk), dim3(256), 0, s, h->H2, NH, NH, h->dZ3, NO, B,
                     h->partial, P_W3);
  hipLaunchKernelGGL(k_reduce_partials, dim3(nblocks(P_N)), dim3(256), 0, s, h->partial, nblk, h->G);
  hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(256), 0, s, h->rowloss, B, h->loss);
  return hipGetLastError() == hipSuccess ? IDFX_OK : IDFX_E_LAUNCH;
}

}  // namespace

extern "C" {

int idfx_mlp_param_count(void) { return P_N; }

int idfx_mlp_create(int max_batch, idfx_mlp **out) {
  if (!out || max_batch < 1) return IDFX_E_ARGS;
  *out = nullptr;
  idfx_mlp *h = new idfx_mlp();
  h->cap = max_batch;
  const long long B = max_batch;
  const int nblk = (max_batch + WG_ROWS - 1) / WG_ROWS;
  int st = IDFX_OK;
  float **bufs[] = {&h->P, &h->G, &h->M, &h->V};
  for (float **b : bufs)
    if (st == IDFX_OK) st = mlp_alloc(h, b, P_N);
  if (st == IDFX_OK) st = mlp_alloc(h, &h->Z1, B * NH);
  if (st == IDFX_OK) st = mlp_alloc(h, &h->H1, B * NH);
  if (st == IDFX_OK) st = mlp_alloc(h, &h->Z2, B * NH);
  if (st == IDFX_OK) st = mlp_alloc(h, &h->H2, B * NH);
  if (st == IDFX_OK) st = mlp_alloc(h, &h->Z3, B * NO);
  if (st == IDFX_OK) st = mlp_alloc(h, &h->dZ3, B * NO);
  if (st == IDFX_OK) st = mlp_alloc(h, &h->dZ2, B * NH);
  if (st == IDFX_OK) st = mlp_alloc(h, &h->dZ1, B * NH);
  if (st == IDFX_OK) st = mlp_alloc(h, &h->partial, (long long)nblk * P_N);
  if (st == IDFX_OK) st = mlp_alloc(h, &h->rowloss, B);
  if (st == IDFX_OK) {
    float *l = nullptr;
    st = mlp_alloc(h, &l, 2);
    h->loss = reinterpret_cast<double *>(l);
  }
  if (st != IDFX_OK) {
    idfx_mlp_destroy(h);
    return st;
  }
  *out = h;
  return IDFX_OK;
}

int idfx_mlp_destroy(idfx_mlp *h) {
  if (!h) return IDFX_OK;
  hipDeviceSynchronize();
  for (void *p : h->allocs) hipFree(p);
  delete h;
  return IDFX_OK;
}

int idfx_mlp_set_params(idfx_mlp *h, const float *host) {
  if (!h || !host) return IDFX_E_ARGS;
  if (hipMemcpy(h->P, host, sizeof(float) * P_N, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(h->M, 0, sizeof(float) * P_N) != hipSuccess ||
      hipMemset(h->V, 0, sizeof(float) * P_N) != hipSuccess)
    return IDFX_E_LAUNCH;
  h->step = 0;
  return IDFX_OK;
}

int idfx_mlp_get_params(idfx_mlp *h, float *host) {
  if (!h || !host) return IDFX_E_ARGS;
  return hipMemcpy(host, h->P, sizeof(float) * P_N, hipMemcpyDeviceToHost) == hipSuccess
             ? IDFX_OK : IDFX_E_LAUNCH;
}

int idfx_mlp_loss_grad(idfx_mlp *h, const float *X, const float *TAU, int B, double *loss,
                       float *grad, void *stream) {
  if (!h || !X || !TAU || B < 1 || B > h->cap) return IDFX_E_ARGS;
  hipStream_t s = (hipStream_t)stream;
  int st = grad_enqueue(h, X, TAU, B, s);
  if (st != IDFX_OK) return st;
  if (hipStreamSynchronize(s) != hipSuccess) return IDFX_E_LAUNCH;
  if (loss && hipMemcpy(loss, h->loss, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return IDFX_E_LAUNCH;
  if (grad && hipMemcpy(grad, h->G, sizeof(float) * P_N, hipMemcpyDeviceToHost) != hipSuccess)
    return IDFX_E_LAUNCH;
  return IDFX_OK;
}

int idfx_mlp_train_step(idfx_mlp *h, const float *X, const float *TAU, int B, float lr,
                        void *stream) {
  if (!h || !X || !TAU || B < 1 || B > h->cap || !(lr > 0.0f)) return IDFX_E_ARGS;
  hipStream_t s = (hipStream_t)stream;
  int st = grad_enqueue(h, X, TAU, B, s);
  if (st != IDFX_OK) return st;
  h->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - (float)std::pow(0.9, (double)h->step);
  const float c2 = 1.0f - (float)std::pow(0.999, (double)h->step);
  hipLaunchKernelGGL(k_adam, dim3(nblocks(P_N)), dim3(256), 0, s, h->P, h->G, h->M, h->V, P_N, lr,
                     b1, b2, 1e-8f, c1, c2);
  return hipGetLastError() == hipSuccess ? IDFX_OK : IDFX_E_LAUNCH;
}

int idfx_mlp_loss(idfx_mlp *h, double *loss, void *stream) {
  if (!h || !loss) return IDFX_E_ARGS;
  hipStream_t s = (hipStream_t)stream;
  if (hipMemcpyAsync(loss, h->loss, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return IDFX_E_LAUNCH;
  return IDFX_OK;
}

}  // extern "C"
