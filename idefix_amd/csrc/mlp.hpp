// Learned reconstruction on gfx950: per-interface MLP inference on BF16 MFMA with FP32
// accumulation, and its FP32 training kernels.  No counterpart in the reference: this is
// the learned limiter of the FP32+learned mode.
//
// One interface x_{i-1/2} along DIR sees the 6-cell stencil i-3 .. i+2 of the 8
// primitives in the face frame (rho, v_n, v_t, v_b, p, B_n, B_t, B_b).  Features:
// per variable c = (V[i-1] + V[i]) / 2, s = max |V_j - c|, x_j = (V_j - c) / s (0 when
// s = 0), ordered [variable][cell].  Network 48 -> 64 -> 64 -> 16 with
// gelu(z) = z sigmoid(1.702 z) on the hidden layers and linear outputs z3 that correct
// the limiter weights of PPM: wL_v = wL_PPM + z3[2v], wR_v = wR_PPM - z3[2v] (output 2v+1
// unused), where the PPM member of the family is w_PPM = (v_PPM - V[i-1]) / (V[i] - V[i-1]);
// the interface states are vL = V[i-1] + wL (V[i] - V[i-1]), vR = V[i] - wR (V[i] - V[i-1]).
// The correction shifts both states of a variable by the same amount, so the jump
// vR - vL that sets the Riemann solver's upwind dissipation stays PPM's, and the BF16
// rounding of the inputs perturbs only the shift.  Per variable the weights must lie in the TVD region of slope
// limiters (Sweby 1984): each state's deviation from its cell value has the sign of the
// cell's other one-sided difference and does not exceed it, and 0 <= w <= 1.  A variable
// that fails, any variable of an interface with non-finite input, and every interface
// when no network is loaded take the PPM weights.
//
// MFMA mapping (v_mfma_f32_16x16x16_bf16): a layer computes Z^T = W^T H^T with the
// weights as the A operand (rows = output units) and the activations as the B operand
// (columns = interfaces).  The 16x16 accumulator tile a lane holds (rows 4(l/16)+t,
// column l%16) is exactly the B-operand fragment of the next layer's K-step, so the
// three layers chain in registers; the lane-to-variable mapping of k_mlp_infer makes the
// input features and the outputs register-resident too.  The weights (16 KB) are staged
// in LDS once per block: read from global memory by every tile of every wave, their
// fragment loads were the kernel's largest cost.
#pragma once

#include "kernels.hpp"

namespace idfx {

constexpr int MLP_NF = 48, MLP_NH = 64, MLP_NO = 16;
constexpr float MLP_GELU_K = 1.702f;

typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));

// Network for inference: transposed BF16 weights ([out][in], row-major) and FP32 biases.
struct MlpNet {
  const short *w1t;  // [64][48]; w2t and w3t follow it in the same allocation
  const short *w2t;  // [64][64]
  const short *w3t;  // [16][64]
  const float *b1, *b2, *b3;  // 64, 64, 16; b2 and b3 follow b1
  int valid;         // 0: no network, every interface takes the PPM weights
};

__device__ __forceinline__ short mlp_bf16(float f) {
  return __builtin_bit_cast(short, (__bf16)f);  // round to nearest even
}
__device__ __forceinline__ float mlp_sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-z));  // v_rcp_f32: 1 ulp
}
__device__ __forceinline__ float mlp_gelu(float z) { return z * mlp_sigmoid(MLP_GELU_K * z); }

// face-frame variable q -> Vc index along DIR
template <class C, int DIR>
__device__ __forceinline__ constexpr int mlp_var(int q) {
  return q == 0 ? C::RHO : q == 1 ? C::template Xn<DIR>() : q == 2 ? C::template Xt<DIR>()
       : q == 3 ? C::template Xb<DIR>() : q == 4 ? C::PRS : q == 5 ? C::template BXn<DIR>()
       : q == 6 ? C::template BXt<DIR>() : C::template BXb<DIR>();
}
// PPM member of the family at one face from the stencil s[0..5] = V[i-3..i+2] and the
// PPM values around the face: left = cell i-1's right value, right = cell i's left value
// (getPPMStates, slopeLimiter.hpp:142-191), with the van Leer fallback where rho or p
// would not be positive (extrapolateToFaces.hpp:260-304); w = 0 where V[i] = V[i-1].
__device__ __forceinline__ void mlp_ppm_weights(const float *s, bool positive, float left,
                                                float right, float &wl, float &wr) {
  if (positive) {
    if (left <= 0.0f) left = s[2] + 0.5f * vanleer(s[3] - s[2], s[2] - s[1]);
    if (right <= 0.0f) right = s[3] - 0.5f * vanleer(s[4] - s[3], s[3] - s[2]);
  }
  const float jump = s[3] - s[2];
  wl = jump != 0.0f ? (left - s[2]) / jump : 0.0f;
  wr = jump != 0.0f ? (s[3] - right) / jump : 0.0f;
}

// The same from the stencil alone: the PPM values of cells i-1 (s[0..4]) and i (s[1..5]).
__device__ __forceinline__ void mlp_ppm(const float *s, bool positive, float &wl, float &wr) {
  float unused, left, right;
  ppm_states<float, true>(s[0], s[1], s[2], s[3], s[4], unused, left);
  ppm_states<float, true>(s[1], s[2], s[3], s[4], s[5], right, unused);
  mlp_ppm_weights(s, positive, left, right, wl, wr);
}

// Features of one variable at one interface from its stencil S[0..5] = V[i-3..i+2]:
// x_j = (S_j - c) / s with c = (S[2] + S[3]) / 2, s = max |S_j - c| (x = 0 when s = 0).
// Returns false for non-finite input.
__device__ __forceinline__ bool mlp_features6(const float *S, float *x) {
  const float c = 0.5f * (S[2] + S[3]);
  float sc = 0.0f;
#pragma unroll
  for (int s = 0; s < 6; s++) sc = fmaxf(sc, fabsf(S[s] - c));
  const float inv = sc > 0.0f ? __builtin_amdgcn_rcpf(sc) : 0.0f;
  bool fin = isfinite(inv);
#pragma unroll
  for (int s = 0; s < 6; s++) fin = fin && isfinite(S[s]);
#pragma unroll
  for (int s = 0; s < 6; s++) x[s] = (S[s] - c) * inv;
  return fin;
}

// TVD-region check of one variable's weights: 0 <= w <= 1, and each state's deviation
// from its cell value has the sign of the cell's other one-sided difference and does not
// exceed it.  False also for NaN weights.
__device__ __forceinline__ bool mlp_tvd_ok(const float *st, float wl, float wr) {
  const float a = st[1], b = st[2], c = st[3], d = st[4];
  const float jump = c - b;
  const float dl = wl * jump, dr = wr * jump;
  const bool okl = wl >= 0.0f && wl <= 1.0f &&
                   (dl == 0.0f || (dl * (b - a) > 0.0f && fabsf(dl) <= fabsf(b - a)));
  const bool okr = wr >= 0.0f && wr <= 1.0f &&
                   (dr == 0.0f || (dr * (d - c) > 0.0f && fabsf(dr) <= fabsf(d - c)));
  return (okl && okr) || (jump == 0.0f && wl == wl && wr == wr);
}

// Interface weights W[q][face] (SoA, face index = index of the cell right of the face)
// for the faces [lo, hi) of direction DIR.  A wave handles 64 faces consecutive in i, in
// four tiles of 16.  Lane l = 16 g + r works on face 16 nt + r of tile nt and on the
// face-frame variables 2g and 2g+1: it loads their two stencils, builds their 12
// features and, after the three layers, holds exactly their four outputs (wL, wR of
// each; the accumulator rows 4g..4g+3), so features enter the MFMA and outputs leave it
// without LDS, and the range check runs on the stencils already in registers.  The K
// order of layer 1 is permuted to match: K index 16 ks + 4 g + t carries feature
// 12 g + 4 ks + t (variable 2g + (4ks+t)/6, cell (4ks+t)%6).
template <class C, int DIR>
__global__ __launch_bounds__(256) void k_mlp_infer(GridDesc g, const double *Vc, MlpNet net,
                                                   float *W, int lo0, int lo1, int lo2,
                                                   int n0, int n1, int n2) {
  static_assert(C::MHD && C::COMP == 3, "the learned reconstruction uses 8 primitives");
  const int lane = threadIdx.x & 63;
  const int g4 = lane >> 4, rr = lane & 15;
  const int nxb = (n0 + 63) / 64;
  const long long w = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  // the network (16 KB of BF16 weights, 144 biases), staged in LDS once per block: every
  // tile of every wave reads all of it as MFMA A fragments
  constexpr int NWT = 64 * 48 + 64 * 64 + 16 * 64, W2 = 64 * 48, W3 = 64 * 48 + 64 * 64;
  __shared__ uint4 wlds[NWT / 8];
  __shared__ float blds[64 + 64 + 16];
  if (net.valid) {
    const uint4 *src = reinterpret_cast<const uint4 *>(net.w1t);
    for (int q = threadIdx.x; q < NWT / 8; q += blockDim.x) wlds[q] = src[q];
    if (threadIdx.x < 144) blds[threadIdx.x] = net.b1[threadIdx.x];
  }
  __syncthreads();
  const short *wsh = reinterpret_cast<const short *>(wlds);
  if (w >= (long long)nxb * n1 * n2) return;  // wave-uniform: the MFMAs need whole waves
  const int xb = (int)(w % nxb);
  const long long r = w / nxb;
  const int j = lo1 + (int)(r % n1);
  const int k = lo2 + (int)(r / n1);
  const long long sd = DIR == 0 ? 1 : DIR == 1 ? (long long)g.nt[0] : (long long)g.nt[0] * g.nt[1];
  const int na = mlp_var<C, DIR>(2 * g4), nb = mlp_var<C, DIR>(2 * g4 + 1);
  const bool pos_a = g4 == 0 || g4 == 2;  // rho (q = 0) and p (q = 4) stay positive
  const double *Va = Vc + na * g.sv, *Vb = Vc + nb * g.sv;
  float carry_a = 0.0f, carry_b = 0.0f;  // PPM right value of the cell left of the tile

#pragma unroll 1
  for (int nt = 0; nt < 4; nt++) {
    const int i = lo0 + xb * 64 + 16 * nt + rr;
    const bool face_ok = i < lo0 + n0;
    const long long pf = idx3(g, k, j, face_ok ? i : lo0);  // the face's index
    float sa[6], sb[6];
#pragma unroll
    for (int s = 0; s < 6; s++) {
      sa[s] = (float)Va[pf + (s - 3) * sd];
      sb[s] = (float)Vb[pf + (s - 3) * sd];
    }
    float x[12];
    // both feature sets are always written (no short circuit): the BF16 conversion below
    // reads all twelve even when a stencil is not finite
    const bool fa = mlp_features6(sa, x);
    const bool fb = mlp_features6(sb, x + 6);
    int fin = fa && fb;
    // the face's eight variables sit in lanes r, r + 16, r + 32, r + 48
    fin &= __shfl_xor(fin, 16);
    fin &= __shfl_xor(fin, 32);
    float wal = 0.0f, war = 0.0f, wbl = 0.0f, wbr = 0.0f;
    if (net.valid) {
      mlp_bf16x4 bx[3];
#pragma unroll
      for (int q = 0; q < 12; q++) bx[q >> 2][q & 3] = mlp_bf16(x[q]);
      mlp_bf16x4 h1[4], h2[4];
      // layer 1: K = 48 in three steps of 16
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 3; ks++) {
          const mlp_bf16x4 a =
              *reinterpret_cast<const mlp_bf16x4 *>(&wsh[(16 * mt + rr) * 48 + 12 * g4 + 4 * ks]);
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bx[ks], acc, 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < 4; t++) h1[mt][t] = mlp_bf16(mlp_gelu(acc[t] + blds[16 * mt + 4 * g4 + t]));
      }
      // layer 2: K = 64, the K-step ks is layer 1's output tile ks
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
          const mlp_bf16x4 a =
              *reinterpret_cast<const mlp_bf16x4 *>(&wsh[W2 + (16 * mt + rr) * 64 + 16 * ks + 4 * g4]);
          acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h1[ks], acc, 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < 4; t++) h2[mt][t] = mlp_bf16(mlp_gelu(acc[t] + blds[64 + 16 * mt + 4 * g4 + t]));
      }
      // layer 3: 16 outputs; this lane holds outputs 4 g4 .. 4 g4 + 3
      mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 4; ks++) {
        const mlp_bf16x4 a =
            *reinterpret_cast<const mlp_bf16x4 *>(&wsh[W3 + rr * 64 + 16 * ks + 4 * g4]);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, h2[ks], acc, 0, 0, 0);
      }
      // the network's corrections of the four PPM weights this lane holds
      wal = acc[0] + blds[128 + 4 * g4];
      war = acc[1] + blds[128 + 4 * g4 + 1];
      wbl = acc[2] + blds[128 + 4 * g4 + 2];
      wbr = acc[3] + blds[128 + 4 * g4 + 3];
    }
    const bool use_net = fin && net.valid;
    // PPM weights of both variables.  The limiter's quotients use v_rcp_f32 (1 ulp); a
    // non-finite one only skips a limiter step, since every comparison with NaN or inf
    // there keeps the unlimited value.
    float pal, par, pbl, pbr;
    if constexpr (DIR == IDIR) {
      // x faces: the lanes of a group hold consecutive faces along x, so the PPM values
      // of a cell serve both of its faces: every lane computes them for its own cell
      // (right of its face, s[1..5]) and hands the right value to the next lane, whose
      // face it bounds; the tile's first face takes it from the previous tile's last
      // lane (tile 0: its lane computes it from s[0..4])
      if (nt == 0 && rr == 0) {
        float unused;
        ppm_states<float, true>(sa[0], sa[1], sa[2], sa[3], sa[4], unused, carry_a);
        ppm_states<float, true>(sb[0], sb[1], sb[2], sb[3], sb[4], unused, carry_b);
      }
      float la, ra, lb, rb;
      ppm_states<float, true>(sa[1], sa[2], sa[3], sa[4], sa[5], la, ra);
      ppm_states<float, true>(sb[1], sb[2], sb[3], sb[4], sb[5], lb, rb);
      float pa = __shfl_up(ra, 1, 16), pb = __shfl_up(rb, 1, 16);
      if (rr == 0) {
        pa = carry_a;
        pb = carry_b;
      }
      carry_a = __shfl(ra, 16 * g4 + 15);
      carry_b = __shfl(rb, 16 * g4 + 15);
      mlp_ppm_weights(sa, pos_a, pa, la, pal, par);
      mlp_ppm_weights(sb, false, pb, lb, pbl, pbr);
    } else {
      // y and z faces: neighbouring lanes are neighbours across the faces, not along
      // them, so each lane computes the PPM values of both cells of each variable
      mlp_ppm(sa, pos_a, pal, par);
      mlp_ppm(sb, false, pbl, pbr);
    }
    // corrected weights where they pass the check, else PPM's: the correction of variable
    // v is output 2v, a common shift of both states (wL + z, wR - z), which keeps PPM's
    // jump vR - vL at the interface; output 2v + 1 is not used
    const float cal = pal + wal, car = par - wal, cbl = pbl + wbl, cbr = pbr - wbl;
    const bool ok_a = use_net && mlp_tvd_ok(sa, cal, car);
    const bool ok_b = use_net && mlp_tvd_ok(sb, cbl, cbr);
    wal = ok_a ? cal : pal;
    war = ok_a ? car : par;
    wbl = ok_b ? cbl : pbl;
    wbr = ok_b ? cbr : pbr;
    if (face_ok) {
      W[(4 * g4) * g.sv + pf] = wal;
      W[(4 * g4 + 1) * g.sv + pf] = war;
      W[(4 * g4 + 2) * g.sv + pf] = wbl;
      W[(4 * g4 + 3) * g.sv + pf] = wbr;
    }
  }
}

}  // namespace idfx
