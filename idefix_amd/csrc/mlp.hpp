// Learned reconstruction on gfx950: per-interface MLP inference on BF16 MFMA with FP32
// accumulation, and its FP32 training kernels.  No counterpart in the reference: this is
// the learned limiter of the FP32+learned mode.
//
// One interface x_{i-1/2} along DIR sees the 6-cell stencil i-3 .. i+2 of the 8
// primitives in the face frame (rho, v_n, v_t, v_b, p, B_n, B_t, B_b).  Features:
// per variable c = (V[i-1] + V[i]) / 2, s = max |V_j - c|, x_j = (V_j - c) / s (0 when
// s = 0), ordered [variable][cell].  Network 48 -> 64 -> 64 -> 16 with
// gelu(z) = z sigmoid(1.702 z) on the hidden layers and sigmoid outputs, the limiter
// weights w (2v: wL_v, 2v+1: wR_v); the interface states are
// vL = V[i-1] + wL (V[i] - V[i-1]), vR = V[i] - wR (V[i] - V[i-1]).  Per variable the
// weights must lie in the TVD region of slope limiters (Sweby 1984): each state's
// deviation from its cell value has the sign of the cell's other one-sided difference
// and does not exceed it, and 0 <= w <= 1.  A variable that fails, any variable of an
// interface with non-finite input, and every interface when no network is loaded take
// the weights of PPM (the member w = (v_PPM - V[i-1]) / (V[i] - V[i-1])).
//
// MFMA mapping (v_mfma_f32_16x16x16_bf16): a layer computes Z^T = W^T H^T with the
// weights as the A operand (rows = output units) and the activations as the B operand
// (columns = interfaces).  The 16x16 accumulator tile a lane holds (rows 4(l/16)+t,
// column l%16) is exactly the B-operand fragment of the next layer's K-step, so the
// three layers chain in registers; only the 48 input features pass through LDS.
#pragma once

#include "kernels.hpp"

namespace idfx {

constexpr int MLP_NF = 48, MLP_NH = 64, MLP_NO = 16;
constexpr float MLP_GELU_K = 1.702f;

typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));

// Network for inference: transposed BF16 weights ([out][in], row-major) and FP32 biases.
struct MlpNet {
  const short *w1t;  // [64][48]
  const short *w2t;  // [64][64]
  const short *w3t;  // [16][64]
  const float *b1, *b2, *b3;
  int valid;         // 0: no network, every interface takes the van Leer weights
};

__device__ __forceinline__ short mlp_bf16(float f) {
  return __builtin_bit_cast(short, (__bf16)f);  // round to nearest even
}
__device__ __forceinline__ float mlp_sigmoid(float z) {
  return 1.0f / (1.0f + __expf(-z));
}
__device__ __forceinline__ float mlp_gelu(float z) { return z * mlp_sigmoid(MLP_GELU_K * z); }

// face-frame variable q -> Vc index along DIR
template <class C, int DIR>
__device__ __forceinline__ constexpr int mlp_var(int q) {
  return q == 0 ? C::RHO : q == 1 ? C::template Xn<DIR>() : q == 2 ? C::template Xt<DIR>()
       : q == 3 ? C::template Xb<DIR>() : q == 4 ? C::PRS : q == 5 ? C::template BXn<DIR>()
       : q == 6 ? C::template BXt<DIR>() : C::template BXb<DIR>();
}
// Vc index -> face-frame variable q along DIR
template <class C, int DIR>
__device__ __forceinline__ constexpr int mlp_q(int n) {
  int q = 0;
  for (int k = 0; k < 8; k++)
    if (mlp_var<C, DIR>(k) == n) q = k;
  return q;
}

// PPM member of the family for one variable from the stencil s[0..5] = V[i-3..i+2]:
// the left state is cell i-1's right value, the right state cell i's left value
// (getPPMStates, slopeLimiter.hpp:142-191), with the van Leer fallback where rho or p
// would not be positive (extrapolateToFaces.hpp:260-304); w = 0 where V[i] = V[i-1].
__device__ __forceinline__ void mlp_ppm(const float *s, bool positive, float &wl, float &wr) {
  float vl_c, left, right, vr_d;
  ppm_states(s[0], s[1], s[2], s[3], s[4], vl_c, left);
  ppm_states(s[1], s[2], s[3], s[4], s[5], right, vr_d);
  if (positive) {
    if (left <= 0.0f) left = s[2] + 0.5f * vanleer(s[3] - s[2], s[2] - s[1]);
    if (right <= 0.0f) right = s[3] - 0.5f * vanleer(s[4] - s[3], s[3] - s[2]);
  }
  const float jump = s[3] - s[2];
  wl = jump != 0.0f ? (left - s[2]) / jump : 0.0f;
  wr = jump != 0.0f ? (s[3] - right) / jump : 0.0f;
}

// Interface weights W[q][face] (SoA, face index = index of the cell right of the face)
// for the faces [lo, hi) of direction DIR.  A wave handles 64 faces consecutive in i.
template <class C, int DIR>
__global__ __launch_bounds__(256) void k_mlp_infer(GridDesc g, const double *Vc, MlpNet net,
                                                   float *W, int lo0, int lo1, int lo2,
                                                   int n0, int n1, int n2) {
  static_assert(C::MHD && C::COMP == 3, "the learned reconstruction uses 8 primitives");
  constexpr int LDS_STRIDE = 52;  // bf16 per feature row (104 B: 8-byte aligned)
  __shared__ short feat[4][64 * LDS_STRIDE];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int nxb = (n0 + 63) / 64;
  const long long w = (long long)blockIdx.x * 4 + wv;
  const bool wave_ok = w < (long long)nxb * n1 * n2;
  const int xb = (int)(w % nxb);
  const long long r = w / nxb;
  const int j = lo1 + (int)(r % n1);
  const int k = lo2 + (int)(r / n1);
  const int i = lo0 + xb * 64 + lane;
  const bool face_ok = wave_ok && i < lo0 + n0;
  const int ic = face_ok ? i : lo0;
  const int jc = wave_ok ? j : lo1;
  const int kc = wave_ok ? k : lo2;
  const long long sd = DIR == 0 ? 1 : DIR == 1 ? (long long)g.nt[0] : (long long)g.nt[0] * g.nt[1];
  const long long pf = idx3(g, kc, jc, ic);  // the face's index (cell right of it)

  // ---- features -----------------------------------------------------------------
  bool finite_in = true;
  short *myfeat = &feat[wv][lane * LDS_STRIDE];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int n = mlp_var<C, DIR>(q);
    float S[6];
#pragma unroll
    for (int s = 0; s < 6; s++) S[s] = (float)Vc[n * g.sv + pf + (s - 3) * sd];
    const float c = 0.5f * (S[2] + S[3]);
    float sc = 0.0f;
#pragma unroll
    for (int s = 0; s < 6; s++) sc = fmaxf(sc, fabsf(S[s] - c));
    const float inv = sc > 0.0f ? 1.0f / sc : 0.0f;
#pragma unroll
    for (int s = 0; s < 6; s++) finite_in = finite_in && isfinite(S[s]);
    finite_in = finite_in && isfinite(inv);
#pragma unroll
    for (int s = 0; s < 6; s++) myfeat[q * 6 + s] = mlp_bf16((S[s] - c) * inv);
  }
  __syncthreads();

  float out[MLP_NO];
  bool ok = finite_in && net.valid;
  if (net.valid) {
    const int g4 = lane >> 4, rr = lane & 15;
    // layer 3 results land in lane (rr, g4) for interface 16 nt + rr, outputs 4 g4 + t;
    // they pass through LDS (the feature rows are no longer needed) to the face's lane
    float res[4][4];
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
      mlp_bf16x4 h1[4], h2[4];
      // layer 1: K = 48 in three steps of 16
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 3; ks++) {
          const mlp_bf16x4 a =
              *reinterpret_cast<const mlp_bf16x4 *>(&net.w1t[(16 * mt + rr) * 48 + 16 * ks + 4 * g4]);
          const mlp_bf16x4 b =
              *reinterpret_cast<const mlp_bf16x4 *>(&feat[wv][(16 * nt + rr) * LDS_STRIDE + 16 * ks + 4 * g4]);
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < 4; t++) h1[mt][t] = mlp_bf16(mlp_gelu(acc[t] + net.b1[16 * mt + 4 * g4 + t]));
      }
      // layer 2: K = 64, the K-step ks is layer 1's output tile ks
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
          const mlp_bf16x4 a =
              *reinterpret_cast<const mlp_bf16x4 *>(&net.w2t[(16 * mt + rr) * 64 + 16 * ks + 4 * g4]);
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h1[ks], acc, 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < 4; t++) h2[mt][t] = mlp_bf16(mlp_gelu(acc[t] + net.b2[16 * mt + 4 * g4 + t]));
      }
      // layer 3: 16 outputs
      mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 4; ks++) {
        const mlp_bf16x4 a =
            *reinterpret_cast<const mlp_bf16x4 *>(&net.w3t[rr * 64 + 16 * ks + 4 * g4]);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h2[ks], acc, 0, 0, 0);
      }
#pragma unroll
      for (int t = 0; t < 4; t++)
        res[nt][t] = mlp_sigmoid(acc[t] + net.b3[4 * g4 + t]);
    }
    __syncthreads();  // every wave is done reading its features
    float *outs = reinterpret_cast<float *>(&feat[wv][0]);  // [64][17] floats
#pragma unroll
    for (int nt = 0; nt < 4; nt++)
#pragma unroll
      for (int t = 0; t < 4; t++) outs[(16 * nt + rr) * 17 + 4 * g4 + t] = res[nt][t];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MLP_NO; q++) out[q] = outs[lane * 17 + q];
  }
  // per-variable TVD check, PPM where it fails
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int n = mlp_var<C, DIR>(q);
    float st[6];
#pragma unroll
    for (int t = 0; t < 6; t++) st[t] = (float)Vc[n * g.sv + pf + (t - 3) * sd];
    const float a = st[1], b = st[2], c = st[3], d = st[4];
    bool good = ok;
    if (good) {
      const float wl = out[2 * q], wr = out[2 * q + 1];
      const float jump = c - b;
      const float dl = wl * jump, dr = wr * jump;
      const bool okl = wl >= 0.0f && wl <= 1.0f &&
                       (dl == 0.0f || (dl * (b - a) > 0.0f && fabsf(dl) <= fabsf(b - a)));
      const bool okr = wr >= 0.0f && wr <= 1.0f &&
                       (dr == 0.0f || (dr * (d - c) > 0.0f && fabsf(dr) <= fabsf(d - c)));
      good = (okl && okr) || (jump == 0.0f && wl == wl && wr == wr);  // NaN fails
    }
    if (!good) mlp_ppm(st, q == 0 || q == 4, out[2 * q], out[2 * q + 1]);
  }
  if (face_ok) {
#pragma unroll
    for (int q = 0; q < MLP_NO; q++) W[q * g.sv + pf] = out[q];
  }
}

}  // namespace idfx
