// Fused second half of a 2-D MHD stage: y sweep + corner EMF + face-B update + RK
// combination + ConsToPrim in one marching kernel.
//
// A lane owns one x column c and walks up a segment of cells.  At face row f it
//   1. reconstructs cell f and solves the y face f (as k_sweep_march does);
//   2. computes the corner EMF ez(c, f) at (x_{c-1/2}, y_{f-1/2}): the y-face data of
//      the neighbouring columns come from the neighbouring lanes by cross-lane shuffles,
//      the x-face data (written by the x sweep) are read from memory;
//   3. updates BX2s(c, f) (needs ez(c+1, f) from the next lane) and BX1s(c, f-1) (needs
//      the previous row's corner, carried in registers), with the RK combination;
//   4. finishes cell (c, f-1): y flux difference into Uc in FP64, InvDt, RK combination,
//      cell-centred B from the new face fields, ConsToPrim in FP64.
// Reads Vc/Vs of the stage start and writes the new state into separate output arrays
// (the next stage reads them), so no lane ever sees a partially updated neighbour.
// The arithmetic of every step follows the same reference expressions as the unfused
// kernels (calcRiemannEmf.hpp, calcCornerEmf.hpp, evolveMagField.hpp:61-74,
// stateContainer.cpp:91-98, convertConsToPrim.hpp:15-67).
#pragma once

#include "kernels.hpp"

namespace idfx {

struct Fused2dArgs {
  const double *Vs_in;  // face field at the stage start (ghosts filled)
  double *Vc_out;       // new primitive state (interior cells)
  double *Vs_out;       // new face field (interior faces)
  double *Vs0;          // Runge-Kutta begin state of the face field
  int rk;               // 0: plain, 1: save Vs0 (stage 0 of a multi-stage scheme), 2: combine
  double wc, w0;        // AddAndStore weights of this stage
  int write_uc;         // write the combined conserved state back to Uc
  int hl;               // halo lanes on the left of the output columns
  int w;                // output columns per wave
  int ext0;             // x extension the reference computes y faces over
};

template <typename T>
__device__ __forceinline__ T lane_at(T v, int lane, int off) {
  return __shfl(v, lane + off);
}

// 3 waves per SIMD for the 6-variable state; the 8-variable one (three vector components)
// needs about 220 VGPRs and runs at 2
template <class C>
constexpr int fused2d_waves_per_eu() {
  return C::NVAR >= 8 ? 2 : 3;
}
template <class C, typename T, int EMF, bool XTRA>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(fused2d_waves_per_eu<C>())))
void k_fused2d_y(GridDesc g, SweepArgs a, EmfArrays<T> e, Fused2dArgs s, MarchGeom m) {
  static_assert(C::DIM == 2 && C::MHD, "2-D MHD only");
  constexpr int NV = C::NVAR;
  constexpr int NW = window<C>();
  constexpr int H = NW / 2;
  constexpr int BX2 = C::BX2;
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  unsigned long long red = 0;
  if (w < m.nxb * m.nseg) {
    const int xb = w % m.nxb;
    const int seg = w / m.nxb;
    const int c = g.beg[0] + xb * s.w - s.hl + lane;
    const int cc = min(max(c, 0), g.nt[0] - 1);
    const bool out_col = lane >= s.hl && lane < s.hl + s.w && c <= g.end[0];
    const bool cell_col = out_col && c < g.end[0];
    // columns whose y faces the reference computes (outside: never-written zeros)
    const bool yface_col = c >= g.beg[0] - s.ext0 && c < g.end[0] + s.ext0;
    const int c0 = g.beg[1] + seg * m.ts;
    const int c1 = min(c0 + m.ts, g.end[1]);
    // 32-bit element offsets from uniform bases (arrays < 2^29 elements)
    const unsigned sd = g.nt[0];
    const unsigned col = cc;                  // idx3(g, 0, 0, cc)
    const unsigned vsd = g.vsn[0];
    const double *Vs1_in = s.Vs_in + g.svs;   // BX2s component
    double *Vs1_out = s.Vs_out + g.svs;
    double *Vs1_0 = s.Vs0 + g.svs;
    const double dtdx1 = a.dt / g.dx[0][cc];  // evolveMagField.hpp:67 (dt/dx1(i))
    constexpr int emf = EMF;  // the corner-EMF scheme is a compile-time choice here

    // PLM: the stencil rows a marching step shares with the next ones are carried in LDS
    // (slot q of thread t at lds[q][t]) instead of being fetched again: each step then
    // reads every input row once.  PPM reloads them (its wider window does not fit).
    constexpr bool CARRY = C::ORDER != RECON_PPM;
    constexpr bool HLLE = emf == EMF_UCT_HLL || emf == EMF_UCT_HLLD;
    constexpr bool NEED_E3 = emf == EMF_UCT0 || emf == EMF_UCT_CONTACT;
    constexpr int RW = NW - 1;                  // window rows carried (PLM: f-1, f)
    constexpr int S_BX = RW * NV;               // uct_hll(d): BX1s rows f-2 .. f (ring of 4)
    constexpr int S_EZ = S_BX + 4;              //             ezi rows f-2 .. f (ring of 4)
    constexpr int S_AD = S_EZ + 4;              //             axL axR dxL dxR of row f-1
    constexpr int S_E1 = RW * NV;               // others: ezi of row f-1
    constexpr int S_S1 = S_E1 + 1;              //         svx of row f-1
    constexpr int NSLOT = CARRY ? (HLLE ? S_AD + 4 : S_S1 + 1) : 1;
    __shared__ T lds[NSLOT][256];
    const int tid = threadIdx.x;
    auto wslot = [&](int row, int n) -> T & { return lds[(row & (RW - 1)) * NV + n][tid]; };
    auto ld_row = [&](const T *arr, int row) { return ld32(arr, (unsigned)row * sd + col); };

    T carry[NV], Fprev[NV];
    T cprev = 0, ezprev = 0;
    double b2prev = 0.0;
    // right-face value of cell c0-1: the left state of the first face
#pragma unroll
    for (int n = 0; n < NV; n++) {
      Fprev[n] = 0;
      carry[n] = 0;
      const bool keep = n != BX2 || NEED_E3;
      if (n == BX2 && !(CARRY && keep)) continue;  // the normal field comes from Vs
      T wv[NW];
#pragma unroll
      for (int q = 0; q < NW; q++)
        wv[q] = (T)ld32(a.Vc + n * g.sv, col + (unsigned)(c0 - 1 - H + q) * sd);
      if constexpr (CARRY) {
        if (keep) {
#pragma unroll
          for (int q = 1; q < NW; q++) wslot(c0 - 1 - H + q, n) = wv[q];
        }
      }
      if (n == BX2) continue;
      T dummy;
      if constexpr (C::ORDER == RECON_MLP)
        recon_mlp<C, JDIR>(wv, n, a.mlpw, g.sv, col + (unsigned)(c0 - 1) * sd, sd, dummy, carry[n]);
      else
        recon_cell<C>(wv, n, dummy, carry[n], recon_dx<C, T>(g, 1, c0 - 1),
                      XTRA && recon_flat<C>(a, col + (unsigned)(c0 - 1) * sd));
    }
    if constexpr (CARRY) {
      if constexpr (HLLE) {
#pragma unroll
        for (int r = c0 - 2; r <= c0; r++) {
          lds[S_BX + (r & 3)][tid] = (T)ld32(s.Vs_in, (unsigned)r * vsd + col);
          lds[S_EZ + (r & 3)][tid] = ld_row(e.ezi, r);
        }
        lds[S_AD + 0][tid] = ld_row(e.axL, c0 - 1);
        lds[S_AD + 1][tid] = ld_row(e.axR, c0 - 1);
        lds[S_AD + 2][tid] = ld_row(e.dxL, c0 - 1);
        lds[S_AD + 3][tid] = ld_row(e.dxR, c0 - 1);
      } else {
        lds[S_E1][tid] = ld_row(e.ezi, c0 - 1);
        if constexpr (emf == EMF_UCT_CONTACT) lds[S_S1][tid] = sv_val<T>(ld32(e.svx, (unsigned)(c0 - 1) * sd + col));
      }
    }
    for (int f = c0; f <= c1; f++) {
      // ---- 1. y face f -------------------------------------------------------------
      T win[NW][NV];
#pragma unroll
      for (int n = 0; n < NV; n++) {
        const bool keep = n != BX2 || NEED_E3;
        if (!keep) continue;
        if constexpr (CARRY) {
#pragma unroll
          for (int q = 0; q < NW - 1; q++) win[q][n] = wslot(f - H + q, n);
          win[NW - 1][n] = (T)ld32(a.Vc + n * g.sv, col + (unsigned)(f + H) * sd);
          wslot(f + H, n) = win[NW - 1][n];  // replaces row f-H, read above
        } else {
#pragma unroll
          for (int q = 0; q < NW; q++)
            win[q][n] = (T)ld32(a.Vc + n * g.sv, col + (unsigned)(f - H + q) * sd);
        }
      }
      T vL[NV], vR[NV], vr_next[NV];
      const bool flat = XTRA && recon_flat<C>(a, col + (unsigned)f * sd);
#pragma unroll
      for (int n = 0; n < NV; n++) {
        vL[n] = carry[n];
        if (n == BX2) {
          vR[n] = vr_next[n] = 0;
          continue;
        }
        T wv[NW];
#pragma unroll
        for (int q = 0; q < NW; q++) wv[q] = win[q][n];
        if constexpr (C::ORDER == RECON_MLP)
          recon_mlp<C, JDIR>(wv, n, a.mlpw, g.sv, col + (unsigned)f * sd, sd, vR[n], vr_next[n]);
        else
          recon_cell<C>(wv, n, vR[n], vr_next[n], recon_dx<C, T>(g, 1, f), flat);
      }
      const double b2old = ld32(Vs1_in, (unsigned)f * vsd + col);
      vL[BX2] = (T)b2old;
      vR[BX2] = vL[BX2];
      T F[NV], cmax;
      FaceEmf<T> E;
      riemann<C, JDIR>(vL, vR, F, cmax, E, emf,
                       eos_face<C, T>(g, col + (unsigned)(f - 1) * sd, col + (unsigned)f * sd));
      if (!yface_col) {
        E.Et = 0;
        E.sv = 0;
        E.aL = E.aR = E.dL = E.dR = 0;
      }

      // ---- 2. corner EMF ez(c, f) ----------------------------------------------------
      T ez;
      auto xf = [&](const T *arr, int row) { return ld32(arr, (unsigned)row * sd + col); };
      double b1old;  // BX1s of row f-1
      if constexpr (HLLE) {
        constexpr int SLO = C::ORDER == RECON_PPM ? 1 : 2;
        constexpr int SHI = C::ORDER == RECON_PPM ? 6 : 5;
        T b[7], v[7];
        // x faces: B = BX2s of y face f and v = ezj along x (columns c-4+s)
        const T b2T = (T)b2old;
#pragma unroll
        for (int q = SLO; q <= SHI; q++) {
          b[q] = lane_at(b2T, lane, q - 4);
          v[q] = lane_at(E.Et, lane, q - 4);
        }
        T aL, aR, dL, dR;
        {
          const T axL = xf(e.axL, f), axR = xf(e.axR, f), dxL = xf(e.dxL, f), dxR = xf(e.dxR, f);
          if constexpr (CARRY) {
            aL = T(0.5) * (lds[S_AD + 0][tid] + axL);
            aR = T(0.5) * (lds[S_AD + 1][tid] + axR);
            dL = T(0.5) * (lds[S_AD + 2][tid] + dxL);
            dR = T(0.5) * (lds[S_AD + 3][tid] + dxR);
            lds[S_AD + 0][tid] = axL;
            lds[S_AD + 1][tid] = axR;
            lds[S_AD + 2][tid] = dxL;
            lds[S_AD + 3][tid] = dxR;
          } else {
            aL = T(0.5) * (xf(e.axL, f - 1) + axL);
            aR = T(0.5) * (xf(e.axR, f - 1) + axR);
            dL = T(0.5) * (xf(e.dxL, f - 1) + dxL);
            dR = T(0.5) * (xf(e.dxR, f - 1) + dxR);
          }
        }
        ez = edge_term<C>(b, v, aL, aR, dL, dR, true);
        // y faces: B = BX1s of x face c and v = ezi along y (rows f-4+s)
        if constexpr (CARRY) {
#pragma unroll
          for (int q = SLO; q < SHI; q++) {
            b[q] = lds[S_BX + ((f - 4 + q) & 3)][tid];
            v[q] = lds[S_EZ + ((f - 4 + q) & 3)][tid];
          }
          b[SHI] = (T)ld32(s.Vs_in, (unsigned)(f - 4 + SHI) * vsd + col);
          v[SHI] = xf(e.ezi, f - 4 + SHI);
          lds[S_BX + ((f - 4 + SHI) & 3)][tid] = b[SHI];
          lds[S_EZ + ((f - 4 + SHI) & 3)][tid] = v[SHI];
        } else {
#pragma unroll
          for (int q = SLO; q <= SHI; q++) {
            b[q] = (T)ld32(s.Vs_in, (unsigned)(f - 4 + q) * vsd + col);
            v[q] = xf(e.ezi, f - 4 + q);
          }
        }
        aL = T(0.5) * (lane_at(E.aL, lane, -1) + E.aL);
        aR = T(0.5) * (lane_at(E.aR, lane, -1) + E.aR);
        dL = T(0.5) * (lane_at(E.dL, lane, -1) + E.dL);
        dR = T(0.5) * (lane_at(E.dR, lane, -1) + E.dR);
        ez += edge_term<C>(b, v, aL, aR, dL, dR, false);
        // BX1s of row f-1 is slot 3 of the stencil (exact when T is double)
        if constexpr (CARRY && std::is_same<T, double>::value) b1old = b[3];
        else b1old = ld32(s.Vs_in, (unsigned)(f - 1) * vsd + col);
      } else {
        b1old = ld32(s.Vs_in, (unsigned)(f - 1) * vsd + col);
        const T ezj_c = E.Et;
        const T ezj_m = lane_at(E.Et, lane, -1);
        const T ezi_f = xf(e.ezi, f);
        T ezi_m;
        if constexpr (CARRY) {
          ezi_m = lds[S_E1][tid];
          lds[S_E1][tid] = ezi_f;
        } else {
          ezi_m = xf(e.ezi, f - 1);
        }
        if constexpr (emf == EMF_ARITHMETIC) {
          ez = T(0.25) * (ezi_f + ezi_m + ezj_c + ezj_m);
        } else {
          // cell-centred E3 = vx2 Bx1 - vx1 Bx2 of cells (c, f-1), (c, f) and column c-1
          auto e3 = [&](int q) {
            return win[q][C::MX2] * win[q][C::BX1] - win[q][C::MX1] * win[q][C::BX2];
          };
          const T E3_cm = e3(H - 1);  // cell (c, f-1)
          const T E3_c = e3(H);       // cell (c, f)
          const T E3_mm = lane_at(E3_cm, lane, -1);
          const T E3_m = lane_at(E3_c, lane, -1);
          if constexpr (emf == EMF_UCT0) {
            const T a1 = T(2) * ezi_f - T(0.5) * (E3_m + E3_c);
            const T a2 = T(2) * ezi_m - T(0.5) * (E3_mm + E3_cm);
            const T a3 = T(2) * ezj_c - T(0.5) * (E3_cm + E3_c);
            const T a4 = T(2) * ezj_m - T(0.5) * (E3_mm + E3_m);
            ez = T(0.25) * (a1 + a2 + a3 + a4);
          } else {
            const T svx_c = sv_val<T>(ld32(e.svx, (unsigned)f * sd + col));
            T svx_m;
            if constexpr (CARRY) {
              svx_m = lds[S_S1][tid];
              lds[S_S1][tid] = svx_c;
            } else {
              svx_m = sv_val<T>(ld32(e.svx, (unsigned)(f - 1) * sd + col));
            }
            const T svy_c = E.sv, svy_m = lane_at(E.sv, lane, -1);
            const T ez_l2 = (1 - svx_m) * (ezj_c - E3_cm) + (svx_m) * (ezj_m - E3_mm);
            const T ez_r2 = (1 - svx_c) * (ezj_c - E3_c) + (svx_c) * (ezj_m - E3_m);
            const T ez_l1 = (1 - svy_m) * (ezi_f - E3_m) + (svy_m) * (ezi_m - E3_mm);
            const T ez_r1 = (1 - svy_c) * (ezi_f - E3_c) + (svy_c) * (ezi_m - E3_cm);
            ez = T(0.25) * (ez_l2 + ez_r2 + ez_l1 + ez_r1 + ezi_f + ezi_m + ezj_c + ezj_m);
          }
        }
      }

      // ---- 3. face fields ------------------------------------------------------------
      const T ezp = lane_at(ez, lane, 1);
      double b2new = b2old + dtdx1 * ((double)ezp - (double)ez);
      if (s.rk == 2) b2new = s.wc * b2new + s.w0 * ld32((const double *)Vs1_0, (unsigned)f * vsd + col);
      if (cell_col && (f < c1 || c1 == g.end[1])) {
        const unsigned q = (unsigned)f * vsd + col;
        if (s.rk == 1) st32(Vs1_0, q, b2old);
        st32(Vs1_out, q, b2new);
      }
      if (f > c0) {
        const int fr = f - 1;
        const unsigned q1 = (unsigned)fr * vsd + col;
        double b1new = b1old + (0.0 - a.dt / g.dx[1][fr] * ((double)ez - (double)ezprev));
        if (s.rk == 2) b1new = s.wc * b1new + s.w0 * ld32((const double *)s.Vs0, q1);
        const double b1next = __shfl(b1new, lane + 1);
        if (out_col) {
          if (s.rk == 1) st32(s.Vs0, q1, b1old);
          st32(s.Vs_out, q1, b1new);
        }
        // ---- 4. cell (c, f-1) ---------------------------------------------------------
        if (cell_col) {
          const unsigned p = (unsigned)fr * sd + col;
          const double dtdV = a.uniform ? a.dtdV : a.dt / cell_dV<2>(g, 0, fr, c);
          const double A = face_A<2, JDIR>(g, 0, fr, c);
          double u[NV], v[NV];
#pragma unroll
          for (int n = 0; n < NV; n++) {
            if (C::is_ct_field(n)) continue;
            const double rhs = -dtdV * ((double)F[n] * A - (double)Fprev[n] * A);
            u[n] = ld32((const double *)a.Uc + n * g.sv, p) + rhs;
          }
          const double term = 0.5 * ((double)cmax + (double)cprev) / (g.dx[1][fr]);
          if (a.flags & SW_INVDT_ADD) {
            const double inv = ld32((const double *)a.InvDt, p) + term;
            st32(a.InvDt, p, inv);
            const unsigned long long bb = __double_as_longlong(inv);
            red = bb > red ? bb : red;
          }
          if (s.rk == 2) {
#pragma unroll
            for (int n = 0; n < NV; n++)
              if (!C::is_ct_field(n))
                u[n] = s.wc * u[n] + s.w0 * ld32((const double *)a.Uc0 + n * g.sv, p);
          }
          u[C::BX1] = 0.5 * (b1new + b1next);
          u[C::BX2] = 0.5 * (b2prev + b2new);
          cons2prim<C>(u, v, g.gamma);
#pragma unroll
          for (int n = 0; n < NV; n++) st32(s.Vc_out + n * g.sv, p, v[n]);
          if (s.write_uc) {
#pragma unroll
            for (int n = 0; n < NV; n++) st32(a.Uc + n * g.sv, p, u[n]);
          }
        }
      }
#pragma unroll
      for (int n = 0; n < NV; n++) {
        carry[n] = vr_next[n];
        Fprev[n] = F[n];
      }
      cprev = cmax;
      ezprev = ez;
      b2prev = b2new;
    }
  }
  if (a.flags & SW_DT_REDUCE) block_max_atomic(red, a.dtmax);
}

}  // namespace idfx
