// Device-side physics of the per-stage finite-volume update, written for gfx950.
//
// Every function is templated on the compile-time configuration `C` (dimensions,
// components, MHD, reconstruction order, Riemann solver) and on the compute type `T`
// (double for the plain FP64 method, float for the FP32-stencil mode).  The
// floating-point operation order of each expression follows the reference function
// cited beside it, so that FP64 results match the reference to the last bit when the
// code is compiled without FMA contraction (-ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace idfx {

constexpr int IDIR = 0, JDIR = 1, KDIR = 2;
// RECON_MLP: the learned limiter (no counterpart in the reference), 3 ghost cells
enum Recon { RECON_CONSTANT = 1, RECON_PLM = 2, RECON_LIMO3 = 3, RECON_PPM = 4, RECON_MLP = 8 };
enum Solver { SOLVER_HLLC = 1, SOLVER_HLLD = 2, SOLVER_ROE = 3, SOLVER_HLL = 4,
              SOLVER_TVDLF = 5 };
enum EmfScheme { EMF_ARITHMETIC = 0, EMF_UCT0 = 1, EMF_UCT_CONTACT = 2, EMF_UCT_HLL = 3,
                 EMF_UCT_HLLD = 4 };

// Compile-time configuration; variable indices follow src/idefix.hpp:74-108.
// ISO: isothermal equation of state (ISOTHERMAL builds: no energy variable, ENG = -1).
template <int DIM_, int COMP_, bool MHD_, int ORDER_, int SOLVER_, bool ISO_ = false>
struct Cfg {
  static constexpr int DIM = DIM_;
  static constexpr int COMP = COMP_;
  static constexpr bool MHD = MHD_;
  static constexpr int ORDER = ORDER_;
  static constexpr int SOLVER = SOLVER_;
  static constexpr bool ISO = ISO_;
  static constexpr int HE = ISO ? 0 : 1;  // HAVE_ENERGY
  static constexpr int RHO = 0, MX1 = 1;
  static constexpr int MX2 = COMP >= 2 ? 2 : -1;
  static constexpr int MX3 = COMP == 3 ? 3 : -1;
  static constexpr int ENG = ISO ? -1 : COMP + 1, PRS = ENG;
  static constexpr int BX1 = MHD ? COMP + 1 + HE : -1;
  static constexpr int BX2 = (MHD && COMP >= 2) ? BX1 + 1 : -1;
  static constexpr int BX3 = (MHD && COMP == 3) ? BX1 + 2 : -1;
  static constexpr int NVAR = MHD ? 2 * COMP + 1 + HE : COMP + 1 + HE;
  static constexpr int NG = ORDER < 4 ? 2 : 3;  // grid.cpp:60-65
  // per-direction indices (hlldMHD.hpp:153-159)
  template <int DIR> static constexpr int Xn() { return MX1 + DIR; }
  template <int DIR> static constexpr int Xt() { return DIR == IDIR ? MX2 : MX1; }
  template <int DIR> static constexpr int Xb() { return DIR == KDIR ? MX2 : MX3; }
  template <int DIR> static constexpr int BXn() { return MHD ? BX1 + DIR : -1; }
  template <int DIR> static constexpr int BXt() { return DIR == IDIR ? BX2 : BX1; }
  template <int DIR> static constexpr int BXb() { return DIR == KDIR ? BX2 : BX3; }
  // a variable that the cell-update loop skips (magnetic components handled by CT)
  static constexpr bool is_ct_field(int n) {
    return MHD && ((n == BX1) || (DIM >= 2 && n == BX2) || (DIM >= 3 && n == BX3));
  }
};

// FMIN/FMAX/FABS of the reference are fmin/fmax/fabs (real_types.hpp:49-51)
template <typename T> __device__ __forceinline__ T fmn(T a, T b) { return fmin(a, b); }
template <typename T> __device__ __forceinline__ T fmx(T a, T b) { return fmax(a, b); }
template <typename T> __device__ __forceinline__ T fab(T a) { return fabs(a); }
__device__ __forceinline__ float fmn(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float fmx(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ float fab(float a) { return fabsf(a); }
template <typename T> __device__ __forceinline__ T sqrt_(T a) { return sqrt(a); }
__device__ __forceinline__ float sqrt_(float a) { return sqrtf(a); }

// ------------------------------------------------------------------------------------
// Slope limiters: src/fluid/RiemannSolver/slopeLimiter.hpp
// ------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T vanleer(T dvp, T dvm) {  // slopeLimiter.hpp:54-57
  return (dvp * dvm > T(0) ? T(2) * dvp * dvm / (dvp + dvm) : T(0));
}
template <typename T>
__device__ __forceinline__ T mclim(T dvp, T dvm) {  // slopeLimiter.hpp:67-75
  T dq = 0;
  if (dvp * dvm > T(0)) {
    T dqc = T(0.5) * (dvp + dvm);
    T d2q = T(2) * (fab(dvp) < fab(dvm) ? dvp : dvm);
    dq = fab(d2q) < fab(dqc) ? d2q : dqc;
  }
  return dq;
}
// LimO3Lim, slopeLimiter.hpp:25-51: the third-order limiter of Cada & Torrilhon (2009)
// with r = 0.1; the limited slope of a cell is dvp * limo3_lim(dvp, dvm, dx).
template <typename T>
__device__ __forceinline__ T limo3_lim(T dvp, T dvm, T dx) {
  const T r = T(0.1), eps = T(1.e-12);
  const T th = dvm / (dvp + T(1.e-16));
  const T q = (T(2) + th) / T(3);
  T a = fmn(T(1.5), T(2) * th);
  a = fmn(q, a);
  const T b = fmx(T(-0.5) * th, a);
  const T c = fmn(q, b);
  T psi = fmx(T(0), c);
  T eta = r * dx;
  eta = (dvm * dvm + dvp * dvp) / (eta * eta);
  if (eta <= T(1) - eps) return q;
  if (eta >= T(1) + eps) return psi;
  psi = (T(1) - (eta - T(1)) / eps) * q + (T(1) + (eta - T(1)) / eps) * psi;
  return T(0.5) * psi;
}
template <typename T>
__device__ __forceinline__ int sgn(T v) { return (T(0) < v) - (v < T(0)); }

// The PPM helpers work on values, not references: conditional stores through reference
// parameters let the compiler merge them into one store through a selected address,
// which keeps the caller's state arrays in scratch memory.
// Quotient of the PPM helpers: correctly rounded, or, with FAST (the learned limiter's
// FP32 fallback only), one v_rcp_f32 (1 ulp) and a product.
template <typename T, bool FAST>
__device__ __forceinline__ T ppm_div(T a, T b) {
  if constexpr (FAST) return a * __builtin_amdgcn_rcpf(b);
  else return a / b;
}
template <typename T, bool FAST = false>
__device__ __forceinline__ T limit_ppm_face(T vm1, T v0, T vp1, T vp2, T vph) {
  // limitPPMFaceValues, slopeLimiter.hpp:121-140
  if ((vp1 - vph) * (vph - v0) < T(0)) {
    const T deltaL = (vm1 - 2 * v0 + vp1);
    const T deltaC = 3 * (v0 - 2 * vph + vp1);
    const T deltaR = (v0 - 2 * vp1 + vp2);
    T delta = 0;
    if (sgn(deltaL) == sgn(deltaC) && sgn(deltaR) == sgn(deltaC)) {
      const T Cc = T(1.25);
      delta = Cc * fmn(fab(deltaL), fab(deltaR));
      delta = sgn(deltaC) * fmn(delta, fab(deltaC));
    }
    return T(0.5) * (v0 + vp1) - ppm_div<T, FAST>(delta, T(6.0));
  }
  return vph;
}
template <typename T, bool FAST = false>
__device__ __forceinline__ void ppm_states(T vm2, T vm1, T v0, T vp1, T vp2, T &vl_out,
                                           T &vr_out) {
  // getPPMStates, slopeLimiter.hpp:142-191
  const int n = 2;
  T vr = T(7.0) / T(12.0) * (v0 + vp1) - T(1.0) / T(12.0) * (vm1 + vp2);
  T vl = T(7.0) / T(12.0) * (vm1 + v0) - T(1.0) / T(12.0) * (vm2 + vp1);
  vl = limit_ppm_face<T, FAST>(vm2, vm1, v0, vp1, vl);
  vr = limit_ppm_face<T, FAST>(vm1, v0, vp1, vp2, vr);
  T d2qf = T(6.0) * (vl + vr - T(2.0) * v0);
  T d2qc0 = vm1 + vp1 - T(2.0) * v0;
  T d2qcp1 = v0 + vp2 - T(2.0) * vp1;
  T d2qcm1 = vm2 + v0 - T(2.0) * vm1;
  T d2q = 0;
  if (sgn(d2qf) == sgn(d2qc0) && sgn(d2qf) == sgn(d2qcp1) && sgn(d2qf) == sgn(d2qcm1)) {
    const T Cc = T(1.25);
    d2q = fmn(fab(d2qc0), fab(d2qcp1));
    d2q = Cc * fmn(fab(d2qcm1), d2q);
    d2q = sgn(d2qf) * fmn(fab(d2qf), d2q);
  }
  T qmax = fmx(fmx(fab(vm1), fab(v0)), fab(vp1));
  T rho = 0;
  if (fab(d2qf) > T(1e-12) * qmax) rho = ppm_div<T, FAST>(d2q, d2qf);
  T l = vl, r = vr;
  if (((vr - v0) * (v0 - vl) <= T(0)) || (vm1 - v0) * (v0 - vp1) <= T(0)) {
    if (rho <= (T(1.0) - T(1e-12))) {
      l = v0 - rho * (v0 - vl);
      r = v0 + rho * (vr - v0);
    }
  } else {
    // sequential as in the reference: the second test sees the updated right state
    if (fab(vr - v0) >= n * fab(v0 - vl)) r = v0 + n * (v0 - vl);
    if (fab(vl - v0) >= n * fab(v0 - r)) l = v0 + n * (v0 - r);
  }
  vl_out = l;
  vr_out = r;
}

// ------------------------------------------------------------------------------------
// Pointwise conversions: src/fluid/convertConsToPrim.hpp, eos_adiabatic.hpp:39-45
// ------------------------------------------------------------------------------------
template <class C, typename T>
__device__ __forceinline__ void prim2cons(const T *v, T *u, T gamma) {
  // K_PrimToCons, convertConsToPrim.hpp:69-100
  u[C::RHO] = v[C::RHO];
  u[C::MX1] = v[C::MX1] * v[C::RHO];
  if constexpr (C::COMP >= 2) u[C::MX2] = v[C::MX2] * v[C::RHO];
  if constexpr (C::COMP >= 3) u[C::MX3] = v[C::MX3] * v[C::RHO];
  if constexpr (C::ISO) {
    if constexpr (C::MHD) {
      u[C::BX1] = v[C::BX1];
      if constexpr (C::COMP >= 2) u[C::BX2] = v[C::BX2];
      if constexpr (C::COMP >= 3) u[C::BX3] = v[C::BX3];
    }
  } else {
    T v2 = v[C::MX1] * v[C::MX1];
    if constexpr (C::COMP >= 2) v2 = v2 + v[C::MX2] * v[C::MX2];
    if constexpr (C::COMP >= 3) v2 = v2 + v[C::MX3] * v[C::MX3];
    const T eint = v[C::ENG] / (gamma - T(1.0));
    if constexpr (C::MHD) {
      u[C::BX1] = v[C::BX1];
      if constexpr (C::COMP >= 2) u[C::BX2] = v[C::BX2];
      if constexpr (C::COMP >= 3) u[C::BX3] = v[C::BX3];
      T b2 = u[C::BX1] * u[C::BX1];
      if constexpr (C::COMP >= 2) b2 = b2 + u[C::BX2] * u[C::BX2];
      if constexpr (C::COMP >= 3) b2 = b2 + u[C::BX3] * u[C::BX3];
      u[C::ENG] = eint + T(0.5) * v[C::RHO] * v2 + T(0.5) * b2;
    } else {
      u[C::ENG] = eint + T(0.5) * v[C::RHO] * v2;
    }
  }
}

// K_ConsToPrim, convertConsToPrim.hpp:15-67: always evaluated in FP64 (the
// energy-to-pressure recovery suffers cancellation at low plasma beta).
// SMALL_PRESSURE_FIX = 1e-5 (fluid_defs.hpp:23); the re-synced energy is local.
template <class C>
__device__ __forceinline__ void cons2prim(const double *u, double *v, double gamma) {
  v[C::RHO] = u[C::RHO];
  v[C::MX1] = u[C::MX1] / u[C::RHO];
  if constexpr (C::COMP >= 2) v[C::MX2] = u[C::MX2] / u[C::RHO];
  if constexpr (C::COMP >= 3) v[C::MX3] = u[C::MX3] / u[C::RHO];
  if constexpr (C::ISO) {
    if constexpr (C::MHD) {
      v[C::BX1] = u[C::BX1];
      if constexpr (C::COMP >= 2) v[C::BX2] = u[C::BX2];
      if constexpr (C::COMP >= 3) v[C::BX3] = u[C::BX3];
    }
  } else {
    double m2 = u[C::MX1] * u[C::MX1];
    if constexpr (C::COMP >= 2) m2 = m2 + u[C::MX2] * u[C::MX2];
    if constexpr (C::COMP >= 3) m2 = m2 + u[C::MX3] * u[C::MX3];
    const double kin = 0.5 / u[C::RHO] * m2;
    if constexpr (C::MHD) {
      v[C::BX1] = u[C::BX1];
      if constexpr (C::COMP >= 2) v[C::BX2] = u[C::BX2];
      if constexpr (C::COMP >= 3) v[C::BX3] = u[C::BX3];
      double b2 = u[C::BX1] * u[C::BX1];
      if constexpr (C::COMP >= 2) b2 = b2 + u[C::BX2] * u[C::BX2];
      if constexpr (C::COMP >= 3) b2 = b2 + u[C::BX3] * u[C::BX3];
      const double mag = 0.5 * b2;
      v[C::ENG] = (u[C::ENG] - kin - mag) * (gamma - 1.0);
    } else {
      v[C::ENG] = (u[C::ENG] - kin) * (gamma - 1.0);
    }
    if (v[C::ENG] <= 0.0) v[C::ENG] = 1.0e-5;
  }
}

// cons2prim above split in two for a schedule that knows the face field only later
// (adiabatic MHD): cons2prim_part gives rho, v and E - kinetic energy (in the energy
// slot) with cons2prim's operations, and prs_from_part finishes the pressure once the
// cell field is known: ((E - kin) - mag) (gamma - 1) and the floor, as cons2prim does.
template <class C>
__device__ __forceinline__ void cons2prim_part(const double *u, double *v) {
  static_assert(C::MHD && !C::ISO, "the split conversion is for adiabatic MHD");
  v[C::RHO] = u[C::RHO];
  v[C::MX1] = u[C::MX1] / u[C::RHO];
  if constexpr (C::COMP >= 2) v[C::MX2] = u[C::MX2] / u[C::RHO];
  if constexpr (C::COMP >= 3) v[C::MX3] = u[C::MX3] / u[C::RHO];
  double m2 = u[C::MX1] * u[C::MX1];
  if constexpr (C::COMP >= 2) m2 = m2 + u[C::MX2] * u[C::MX2];
  if constexpr (C::COMP >= 3) m2 = m2 + u[C::MX3] * u[C::MX3];
  const double kin = 0.5 / u[C::RHO] * m2;
  v[C::ENG] = u[C::ENG] - kin;
}
template <class C>
__device__ __forceinline__ double prs_from_part(double e_part, const double *b, double gamma) {
  double b2 = b[0] * b[0];
  if constexpr (C::COMP >= 2) b2 = b2 + b[1] * b[1];
  if constexpr (C::COMP >= 3) b2 = b2 + b[2] * b[2];
  const double mag = 0.5 * b2;
  double p = (e_part - mag) * (gamma - 1.0);
  if (p <= 0.0) p = 1.0e-5;
  return p;
}

// ------------------------------------------------------------------------------------
// K_Flux<Phys,DIR>, src/fluid/RiemannSolver/flux.hpp:28-95
// ------------------------------------------------------------------------------------
// c2 is Cs2Iso, the squared isothermal sound speed (isothermal configurations only)
template <class C, int DIR, typename T>
__device__ __forceinline__ void kflux(T *F, const T *v, const T *u, T c2 = T(0)) {
  constexpr int Xn = C::template Xn<DIR>();
  F[C::RHO] = u[Xn];
  if constexpr (C::MHD) {
    constexpr int BXn = C::template BXn<DIR>();
    constexpr int BXt = C::template BXt<DIR>();
    constexpr int BXb = C::template BXb<DIR>();
    constexpr int Xt = C::template Xt<DIR>();
    constexpr int Xb = C::template Xb<DIR>();
    F[C::MX1] = u[C::MX1] * v[Xn] - v[BXn] * v[C::BX1];
    if constexpr (C::COMP >= 2) F[C::MX2] = u[C::MX2] * v[Xn] - v[BXn] * v[C::BX2];
    if constexpr (C::COMP >= 3) F[C::MX3] = u[C::MX3] * v[Xn] - v[BXn] * v[C::BX3];
    F[BXn] = T(0);
    if constexpr (C::COMP >= 2) F[BXt] = v[Xn] * v[BXt] - v[BXn] * v[Xt];
    if constexpr (C::COMP >= 3) F[BXb] = v[Xn] * v[BXb] - v[BXn] * v[Xb];
    T b2 = v[C::BX1] * v[C::BX1];
    if constexpr (C::COMP >= 2) b2 = b2 + v[C::BX2] * v[C::BX2];
    if constexpr (C::COMP >= 3) b2 = b2 + v[C::BX3] * v[C::BX3];
    if constexpr (C::ISO) {
      F[Xn] += c2 * v[C::RHO] + T(0.5) * b2;
    } else {
      const T ptot = v[C::ENG] + T(0.5) * b2;
      T vb = v[C::MX1] * v[C::BX1];
      if constexpr (C::COMP >= 2) vb = vb + v[C::MX2] * v[C::BX2];
      if constexpr (C::COMP >= 3) vb = vb + v[C::MX3] * v[C::BX3];
      F[C::ENG] = (u[C::ENG] + ptot) * v[Xn] - v[BXn] * vb;
      F[Xn] += ptot;
    }
  } else {
    F[C::MX1] = u[C::MX1] * v[Xn];
    if constexpr (C::COMP >= 2) F[C::MX2] = u[C::MX2] * v[Xn];
    if constexpr (C::COMP >= 3) F[C::MX3] = u[C::MX3] * v[Xn];
    if constexpr (C::ISO) {
      F[Xn] += c2 * v[C::RHO];
    } else {
      const T ptot = v[C::ENG];
      F[C::ENG] = (u[C::ENG] + ptot) * v[Xn];
      F[Xn] += ptot;
    }
  }
}

// Face data the constrained-transport step needs (storeFlux.hpp:14-208).
template <typename T>
struct FaceEmf {
  T Et, Eb, sv, aL, aR, dL, dR;
};

// ------------------------------------------------------------------------------------
// HllcHD, src/fluid/RiemannSolver/HDsolvers/hllcHD.hpp:35-163 (HAVE_ENERGY branch)
// ------------------------------------------------------------------------------------
template <class C, int DIR, typename T>
__device__ __forceinline__ void hllc_hd(const T *vL, const T *vR, T *flux, T &cmax, T gamma) {
  constexpr int NV = C::NVAR;
  constexpr int Xn = C::template Xn<DIR>();
  constexpr int Xt = C::template Xt<DIR>();
  constexpr int Xb = C::template Xb<DIR>();
  T uL[NV], uR[NV], fL[NV], fR[NV];
  const T cL = sqrt_(gamma * (vL[C::ENG] / vL[C::RHO]));
  const T cR = sqrt_(gamma * (vR[C::ENG] / vR[C::RHO]));
  const T cminL = vL[Xn] - cL, cmaxL = vL[Xn] + cL;
  const T cminR = vR[Xn] - cR, cmaxR = vR[Xn] + cR;
  const T SL = fmn(cminL, cminR);
  const T SR = fmx(cmaxL, cmaxR);
  cmax = fmx(fab(SL), fab(SR));
  prim2cons<C>(vL, uL, gamma);
  prim2cons<C>(vR, uR, gamma);
  kflux<C, DIR>(fL, vL, uL);
  kflux<C, DIR>(fR, vR, uR);
  if (SL > T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fL[n];
  } else if (SR < T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fR[n];
  } else {
    T usL[NV], usR[NV];
    const T qL = vL[C::ENG] + uL[Xn] * (vL[Xn] - SL);
    const T qR = vR[C::ENG] + uR[Xn] * (vR[Xn] - SR);
    const T wL = vL[C::RHO] * (vL[Xn] - SL);
    const T wR = vR[C::RHO] * (vR[Xn] - SR);
    const T vs = (qR - qL) / (wR - wL);
    usL[C::RHO] = uL[C::RHO] * (SL - vL[Xn]) / (SL - vs);
    usR[C::RHO] = uR[C::RHO] * (SR - vR[Xn]) / (SR - vs);
    usL[Xn] = usL[C::RHO] * vs;
    usR[Xn] = usR[C::RHO] * vs;
    if constexpr (C::COMP >= 2) {
      usL[Xt] = usL[C::RHO] * vL[Xt];
      usR[Xt] = usR[C::RHO] * vR[Xt];
    }
    if constexpr (C::COMP >= 3) {
      usL[Xb] = usL[C::RHO] * vL[Xb];
      usR[Xb] = usR[C::RHO] * vR[Xb];
    }
    usL[C::ENG] = uL[C::ENG] / vL[C::RHO] +
                  (vs - vL[Xn]) * (vs + vL[C::ENG] / (vL[C::RHO] * (SL - vL[Xn])));
    usR[C::ENG] = uR[C::ENG] / vR[C::RHO] +
                  (vs - vR[Xn]) * (vs + vR[C::ENG] / (vR[C::RHO] * (SR - vR[Xn])));
    usL[C::ENG] *= usL[C::RHO];
    usR[C::ENG] *= usR[C::RHO];
    if (vs >= T(0)) {
#pragma unroll
      for (int n = 0; n < NV; n++) flux[n] = fL[n] + SL * (usL[n] - uL[n]);
    } else {
#pragma unroll
      for (int n = 0; n < NV; n++) flux[n] = fR[n] + SR * (usR[n] - uR[n]);
    }
  }
}

// ------------------------------------------------------------------------------------
// HllHD, src/fluid/RiemannSolver/HDsolvers/hllHD.hpp:35-120 (HAVE_ENERGY branch)
// ------------------------------------------------------------------------------------
template <class C, int DIR, typename T>
__device__ __forceinline__ void hll_hd(const T *vL, const T *vR, T *flux, T &cmax, T gamma) {
  constexpr int NV = C::NVAR;
  constexpr int Xn = C::template Xn<DIR>();
  T uL[NV], uR[NV], fL[NV], fR[NV];
  const T cL = sqrt_(gamma * (vL[C::ENG] / vL[C::RHO]));
  const T cR = sqrt_(gamma * (vR[C::ENG] / vR[C::RHO]));
  const T SL = fmn(vL[Xn] - cL, vR[Xn] - cR);
  const T SR = fmx(vL[Xn] + cL, vR[Xn] + cR);
  cmax = fmx(fab(SL), fab(SR));
  prim2cons<C>(vL, uL, gamma);
  prim2cons<C>(vR, uR, gamma);
  kflux<C, DIR>(fL, vL, uL);
  kflux<C, DIR>(fR, vR, uR);
  if (SL > T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fL[n];
  } else if (SR < T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fR[n];
  } else {
    const T sls = SL * SR, dS = SR - SL;
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = (sls * uR[n] - sls * uL[n] + SR * fL[n] - SL * fR[n]) / dS;
  }
}

// ------------------------------------------------------------------------------------
// TvdlfHD, src/fluid/RiemannSolver/HDsolvers/tvdlfHD.hpp:37-96 (HAVE_ENERGY branch):
// Rusanov flux with the sound speed of the arithmetic mean state.
// ------------------------------------------------------------------------------------
template <class C, int DIR, typename T>
__device__ __forceinline__ void tvdlf_hd(const T *vL, const T *vR, T *flux, T &cmax, T gamma) {
  constexpr int NV = C::NVAR;
  constexpr int Xn = C::template Xn<DIR>();
  T uL[NV], uR[NV], fL[NV], fR[NV];
  const T rho = T(0.5) * (vL[C::RHO] + vR[C::RHO]);
  const T prs = T(0.5) * (vL[C::ENG] + vR[C::ENG]);
  const T vn = T(0.5) * (vL[Xn] + vR[Xn]);
  const T cRL = sqrt_(gamma * (prs / rho));
  cmax = fmx(fab(vn + cRL), fab(vn - cRL));
  prim2cons<C>(vL, uL, gamma);
  prim2cons<C>(vR, uR, gamma);
  kflux<C, DIR>(fL, vL, uL);
  kflux<C, DIR>(fR, vR, uR);
#pragma unroll
  for (int n = 0; n < NV; n++) flux[n] = T(0.5) * (fL[n] + fR[n] - cmax * (uR[n] - uL[n]));
}

// ------------------------------------------------------------------------------------
// RoeHD, src/fluid/RiemannSolver/HDsolvers/roeHD.hpp:55-364 (HAVE_ENERGY,
// ROE_AVERAGE == 0).  Returns false where the reference leaves the flux unwritten
// (strong shock in 1-D: the HLL switch exists only under DIMENSIONS > 1).
// ------------------------------------------------------------------------------------
template <class C, int DIR, typename T>
__device__ __forceinline__ bool roe_hd(const T *vL, const T *vR, T *flux, T &cmax, T gamma) {
  constexpr int NV = C::NVAR;
  constexpr int Xn = C::template Xn<DIR>();
  constexpr int Xt = C::template Xt<DIR>();
  constexpr int Xb = C::template Xb<DIR>();
  T dv[NV], uL[NV], uR[NV], fL[NV], fR[NV], um[NV];
#pragma unroll
  for (int n = 0; n < NV; n++) dv[n] = vR[n] - vL[n];
  T a2L = sqrt_(gamma * (vL[C::ENG] / vL[C::RHO]));
  T a2R = sqrt_(gamma * (vR[C::ENG] / vR[C::RHO]));
  a2L = a2L * a2L;
  a2R = a2R * a2R;
  (void)a2L;
  (void)a2R;
  prim2cons<C>(vL, uL, gamma);
  prim2cons<C>(vR, uR, gamma);
  kflux<C, DIR>(fL, vL, uL);
  kflux<C, DIR>(fR, vR, uR);
  const T gamma_m1 = gamma - 1;
#pragma unroll
  for (int n = 0; n < NV; n++) um[n] = T(0.5) * (vR[n] + vL[n]);
  const T a2 = gamma * um[C::ENG] / um[C::RHO];
  const T a = sqrt_(a2);
  T vel2 = um[C::MX1] * um[C::MX1];
  if constexpr (C::COMP >= 2) vel2 = vel2 + um[C::MX2] * um[C::MX2];
  if constexpr (C::COMP >= 3) vel2 = vel2 + um[C::MX3] * um[C::MX3];
  const T h = T(0.5) * vel2 + a2 / gamma_m1;

  // modes I0, I1, IE, I2, I3 = 0..4 (roeHD.hpp:20-27); Rc[var][mode]
  T lambda[NV], eta[NV], Rc[NV][NV];
#pragma unroll
  for (int p = 0; p < NV; p++)
#pragma unroll
    for (int q = 0; q < NV; q++) Rc[p][q] = 0;
  lambda[0] = um[Xn] - a;
  eta[0] = T(0.5) / a2 * (dv[C::ENG] - dv[Xn] * um[C::RHO] * a);
  Rc[C::RHO][0] = 1;
  Rc[Xn][0] = um[Xn] - a;
  if constexpr (C::COMP >= 2) Rc[Xt][0] = um[Xt];
  if constexpr (C::COMP >= 3) Rc[Xb][0] = um[Xb];
  Rc[C::ENG][0] = h - um[Xn] * a;
  lambda[1] = um[Xn] + a;
  eta[1] = T(0.5) / a2 * (dv[C::ENG] + dv[Xn] * um[C::RHO] * a);
  Rc[C::RHO][1] = 1;
  Rc[Xn][1] = um[Xn] + a;
  if constexpr (C::COMP >= 2) Rc[Xt][1] = um[Xt];
  if constexpr (C::COMP >= 3) Rc[Xb][1] = um[Xb];
  Rc[C::ENG][1] = h + um[Xn] * a;
  lambda[2] = um[Xn];
  eta[2] = dv[C::RHO] - dv[C::ENG] / a2;
  Rc[C::RHO][2] = 1;
  Rc[C::MX1][2] = um[C::MX1];
  if constexpr (C::COMP >= 2) Rc[C::MX2][2] = um[C::MX2];
  if constexpr (C::COMP >= 3) Rc[C::MX3][2] = um[C::MX3];
  Rc[C::ENG][2] = T(0.5) * vel2;
  if constexpr (C::COMP > 1) {
    lambda[3] = um[Xn];
    eta[3] = um[C::RHO] * dv[Xt];
    Rc[Xt][3] = 1;
    Rc[C::ENG][3] = um[Xt];
  }
  if constexpr (C::COMP > 2) {
    lambda[4] = um[Xn];
    eta[4] = um[C::RHO] * dv[Xb];
    Rc[Xb][4] = 1;
    Rc[C::ENG][4] = um[Xb];
  }
  cmax = fab(um[Xn]) + a;
  T scrh = fab(vL[C::ENG] - vR[C::ENG]);
  scrh /= fmn(vL[C::ENG], vR[C::ENG]);
  if (scrh > T(0.5) && (vR[Xn] < vL[Xn])) {
    if constexpr (C::DIM > 1) {
      const T bmin = fmn(T(0), lambda[0]);
      const T bmax = fmx(T(0), lambda[1]);
      const T scrh1 = T(1) / (bmax - bmin);
#pragma unroll
      for (int n = 0; n < NV; n++) {
        flux[n] = bmin * bmax * (uR[n] - uL[n]) + bmax * fL[n] - bmin * fR[n];
        flux[n] *= scrh1;
      }
      return true;
    } else {
      return false;
    }
  }
  T alambda[NV];
#pragma unroll
  for (int n = 0; n < NV; n++) alambda[n] = fab(lambda[n]);
  const T delta = T(1.e-7);
  if (alambda[0] <= delta) alambda[0] = T(0.5) * lambda[0] * lambda[0] / delta + T(0.5) * delta;
  if (alambda[1] <= delta) alambda[1] = T(0.5) * lambda[1] * lambda[1] / delta + T(0.5) * delta;
#pragma unroll
  for (int n = 0; n < NV; n++) {
    T f = fL[n] + fR[n];
#pragma unroll
    for (int m = 0; m < NV; m++) f -= alambda[m] * eta[m] * Rc[n][m];
    flux[n] = f * T(0.5);
  }
  return true;
}

// Star-region quantities that HllD and K_StoreHLLD both compute from the same inputs
// with the same expressions (hlldMHD.hpp:292-306, storeFlux.hpp:121-135): computed once.
template <typename T>
struct StarPre {
  bool valid;
  T duL, duR, scrh, Sc, usLRHO, usRRHO, sqrL, sqrR;
};

template <class C, int DIR, typename T>
__device__ __forceinline__ void star_pre(const T *vL, const T *vR, const T *uL, const T *uR,
                                         T sl, T sr, T ptL, T ptR, StarPre<T> &P) {
  constexpr int Xn = C::template Xn<DIR>();
  P.valid = true;
  P.duL = sl - vL[Xn];
  P.duR = sr - vR[Xn];
  P.scrh = T(1) / (P.duR * uR[C::RHO] - P.duL * uL[C::RHO]);
  P.Sc = (P.duR * uR[Xn] - P.duL * uL[Xn] - ptR + ptL) * P.scrh;
  P.usLRHO = uL[C::RHO] * P.duL / (sl - P.Sc);
  P.usRRHO = uR[C::RHO] * P.duR / (sr - P.Sc);
  P.sqrL = sqrt_(P.usLRHO);
  P.sqrR = sqrt_(P.usRRHO);
}

// Face data for constrained transport (storeFlux.hpp:14-208) with the signs of
// hlldMHD.hpp:79-142 / roeMHD.hpp:106-170; shared by HLLD and Roe MHD.
template <class C, int DIR, typename T>
__device__ __forceinline__ void store_face_emf(const T *vL, const T *vR, const T *uL,
                                               const T *uR, const T *flux, T sl, T sr, T ptL,
                                               T ptR, FaceEmf<T> &E, int emf,
                                               const StarPre<T> &pre) {
  constexpr int CO = C::COMP;
  constexpr int Xn = C::template Xn<DIR>();
  constexpr int Xt = C::template Xt<DIR>();
  constexpr int Xb = C::template Xb<DIR>();
  constexpr int BXn = C::template BXn<DIR>();
  constexpr int BXt = C::template BXt<DIR>();
  constexpr int BXb = C::template BXb<DIR>();
  T st, sb = T(1);  // hlldMHD.hpp:79-142
  if constexpr (DIR == IDIR) { st = T(-1); sb = T(1); }
  if constexpr (DIR == JDIR) { st = T(1); sb = (C::DIM == 3) ? T(-1) : T(1); }
  if constexpr (DIR == KDIR) { st = T(-1); sb = T(1); }
  if (emf <= EMF_UCT_CONTACT) {
    // K_StoreEMF (arithmetic, uct0) and K_StoreContact
    if constexpr (CO > 1) {
      E.Et = st * flux[BXt];
      if constexpr (C::DIM == 3) E.Eb = sb * flux[BXb];
    }
    if (emf == EMF_UCT_CONTACT) {
      T s = T(0.5);
      if (flux[C::RHO] > T(1.0e-6)) s = T(1);
      if (flux[C::RHO] < -T(1.0e-6)) s = T(0);
      E.sv = s;
    }
  } else if (emf == EMF_UCT_HLL) {
    const T ar = fmx(T(0), sr);
    const T al = fmn(T(0), sl);
    const T scrh = T(1) / (ar - al);
    if constexpr (CO > 1) {
      E.Et = -st * (ar * vL[Xt] - al * vR[Xt]) * scrh;
      if constexpr (CO >= 3) E.Eb = -sb * (ar * vL[Xb] - al * vR[Xb]) * scrh;
    }
    E.aL = ar * scrh;
    E.aR = -al * scrh;
    E.dR = -al * ar * scrh;
    E.dL = E.dR;
  } else {
    // K_StoreHLLD, storeFlux.hpp:77-208 (HAVE_ENERGY branch)
    bool revert_to_hll = false;
    const T Bn = (sr * vR[BXn] - sl * vL[BXn]) / (sr - sl);
    T scrh, Sc, SaL, SaR;
    if constexpr (C::ISO) {
      // storeFlux.hpp:139-147: one averaged density across the fan
      const T duL = sl - vL[Xn], duR = sr - vR[Xn];
      const T scrh0 = T(1) / (sr - sl);
      const T rho_h = (uR[C::RHO] * duR - uL[C::RHO] * duL) * scrh0;
      Sc = (sl * uR[C::RHO] * duR - sr * uL[C::RHO] * duL) * scrh0 / rho_h;
      const T sqrho_h = sqrt_(rho_h);
      SaL = Sc - fab(Bn) / sqrho_h;
      SaR = Sc + fab(Bn) / sqrho_h;
    } else {
      StarPre<T> P = pre;
      if (!P.valid) star_pre<C, DIR>(vL, vR, uL, uR, sl, sr, ptL, ptR, P);
      Sc = P.Sc;
      SaL = Sc - fab(Bn) / P.sqrL;
      SaR = Sc + fab(Bn) / P.sqrR;
    }
    if ((SaL - sl) < T(1e-4) * (Sc - sl)) revert_to_hll = true;
    if ((SaR - sr) > -T(1e-4) * (sr - Sc)) revert_to_hll = true;
    const T chiL = (vL[Xn] - Sc) * (sl - Sc) / (SaL + sl - T(2) * Sc);
    const T chiR = (vR[Xn] - Sc) * (sr - Sc) / (SaR + sr - T(2) * Sc);
    const T nuL = (SaL + sl) / (fab(SaL) + fab(sl));
    const T nuR = (SaR + sr) / (fab(SaR) + fab(sr));
    const T nuLR = (SaL + SaR) / (fab(SaL) + fab(SaR));
    if (fab(SaR - SaL) > T(1.e-9) * fab(sr - sl)) {
      E.dL = T(0.5) * (chiL * nuL - chiL * nuLR) + T(0.5) * (fab(SaL) - nuLR * SaL);
      E.dR = T(0.5) * (chiR * nuR - chiR * nuLR) + T(0.5) * (fab(SaR) - nuLR * SaR);
      E.aL = T(0.5) * (T(1) + nuLR);
      E.aR = T(0.5) * (T(1) - nuLR);
    } else {
      E.dL = T(0.5) * chiL * nuL + T(0.5) * fab(SaL);
      E.dR = T(0.5) * chiR * nuR + T(0.5) * fab(SaR);
      E.aL = T(0.5);
      E.aR = T(0.5);
    }
    const T ar = fmx(T(0), sr);
    const T al = fmn(T(0), sl);
    scrh = T(1) / (ar - al);
    if (revert_to_hll) {
      E.aL = ar * scrh;
      E.aR = -al * scrh;
      E.dR = -al * ar * scrh;
      E.dL = E.dR;
    }
    if constexpr (CO > 1) {
      E.Et = -st * (ar * vL[Xt] - al * vR[Xt]) * scrh;
      if constexpr (CO >= 3) E.Eb = -sb * (ar * vL[Xb] - al * vR[Xb]) * scrh;
    }
  }
}

// ------------------------------------------------------------------------------------
// HlldMHD, src/fluid/RiemannSolver/MHDsolvers/hlldMHD.hpp:147-589 (HAVE_ENERGY branch)
// followed by the face-EMF stores of storeFlux.hpp:14-208.
// ------------------------------------------------------------------------------------
template <class C, int DIR, typename T>
__device__ __forceinline__ void hlld_mhd(const T *vL, const T *vR, T *flux, T &cmax,
                                         FaceEmf<T> &E, int emf, T gamma) {
  constexpr int NV = C::NVAR;
  constexpr int CO = C::COMP;
  constexpr int Xn = C::template Xn<DIR>();
  constexpr int Xt = C::template Xt<DIR>();
  constexpr int Xb = C::template Xb<DIR>();
  constexpr int BXn = C::template BXn<DIR>();
  constexpr int BXt = C::template BXt<DIR>();
  constexpr int BXb = C::template BXb<DIR>();
  T uL[NV], uR[NV], fluxL[NV], fluxR[NV];

  // fast magnetosonic speeds, hlldMHD.hpp:183-226
  T cL, cR;
  {
    T gpr = gamma * vL[C::ENG];
    T b1 = vL[BXn], b2 = 0, b3 = 0;
    if constexpr (CO >= 2) b2 = vL[BXt];
    if constexpr (CO >= 3) b3 = vL[BXb];
    T Btmag2 = b2 * b2 + b3 * b3;
    T Bmag2 = b1 * b1 + Btmag2;
    cL = gpr - Bmag2;
    cL = gpr + Bmag2 + sqrt_(cL * cL + T(4) * gpr * Btmag2);
    cL = sqrt_(T(0.5) * cL / vL[C::RHO]);
  }
  {
    T gpr = gamma * vR[C::ENG];
    T b1 = vR[BXn], b2 = 0, b3 = 0;
    if constexpr (CO >= 2) b2 = vR[BXt];
    if constexpr (CO >= 3) b3 = vR[BXb];
    T Btmag2 = b2 * b2 + b3 * b3;
    T Bmag2 = b1 * b1 + Btmag2;
    cR = gpr - Bmag2;
    cR = gpr + Bmag2 + sqrt_(cR * cR + T(4) * gpr * Btmag2);
    cR = sqrt_(T(0.5) * cR / vR[C::RHO]);
  }
  const T cminL = vL[Xn] - cL, cmaxL = vL[Xn] + cL;
  const T cminR = vR[Xn] - cR, cmaxR = vR[Xn] + cR;
  const T sl = fmn(cminL, cminR);
  const T sr = fmx(cmaxL, cmaxR);
  cmax = fmx(fab(sl), fab(sr));

  prim2cons<C>(vL, uL, gamma);
  prim2cons<C>(vR, uR, gamma);
  kflux<C, DIR>(fluxL, vL, uL);
  kflux<C, DIR>(fluxR, vR, uR);

  T b2L = vL[C::BX1] * vL[C::BX1];
  if constexpr (CO >= 2) b2L = b2L + vL[C::BX2] * vL[C::BX2];
  if constexpr (CO >= 3) b2L = b2L + vL[C::BX3] * vL[C::BX3];
  T b2R = vR[C::BX1] * vR[C::BX1];
  if constexpr (CO >= 2) b2R = b2R + vR[C::BX2] * vR[C::BX2];
  if constexpr (CO >= 3) b2R = b2R + vR[C::BX3] * vR[C::BX3];
  const T ptL = vL[C::ENG] + T(0.5) * b2L;
  const T ptR = vR[C::ENG] + T(0.5) * b2R;
  StarPre<T> pre;
  pre.valid = false;
  if (emf == EMF_UCT_HLLD) star_pre<C, DIR>(vL, vR, uL, uR, sl, sr, ptL, ptR, pre);

  if (sl > T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fluxL[n];
  } else if (sr < T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fluxR[n];
  } else {
    T usL[NV], usR[NV];
    T scrh, scrhL, scrhR, duL, duR, sBx, Bx, SM, S1L, S1R, pts, sqrL, sqrR;
    T vsL = 0, vsR = 0, wsL = 0, wsR = 0;
    bool revert_to_hllc = false;
    scrh = T(1) / (sr - sl);
    Bx = (sr * vR[BXn] - sl * vL[BXn]) * scrh;
    sBx = (Bx > T(0) ? T(1) : T(-1));
    StarPre<T> P = pre;
    if (!P.valid) star_pre<C, DIR>(vL, vR, uL, uR, sl, sr, ptL, ptR, P);
    duL = P.duL;
    duR = P.duR;
    scrh = P.scrh;
    SM = P.Sc;
    pts = duR * uR[C::RHO] * ptL - duL * uL[C::RHO] * ptR +
          vL[C::RHO] * vR[C::RHO] * duR * duL * (vR[Xn] - vL[Xn]);
    pts *= scrh;
    usL[C::RHO] = P.usLRHO;
    usR[C::RHO] = P.usRRHO;
    sqrL = P.sqrL;
    sqrR = P.sqrR;
    S1L = SM - fab(Bx) / sqrL;
    S1R = SM + fab(Bx) / sqrR;
    if ((S1L - sl) < T(1.e-4) * (SM - sl)) revert_to_hllc = true;
    if ((S1R - sr) > -T(1.e-4) * (sr - SM)) revert_to_hllc = true;
    if (revert_to_hllc) {
      scrh = T(1) / (sr - sl);
      T Uhll_n, Uhll_t = 0, Uhll_b = 0;
      Uhll_n = (sr * uR[BXn] - sl * uL[BXn] + fluxL[BXn] - fluxR[BXn]) * scrh;
      if constexpr (CO >= 2) Uhll_t = (sr * uR[BXt] - sl * uL[BXt] + fluxL[BXt] - fluxR[BXt]) * scrh;
      if constexpr (CO >= 3) Uhll_b = (sr * uR[BXb] - sl * uL[BXb] + fluxL[BXb] - fluxR[BXb]) * scrh;
      usL[BXn] = usR[BXn] = Uhll_n;
      if constexpr (CO >= 2) usL[BXt] = usR[BXt] = Uhll_t;
      if constexpr (CO >= 3) usL[BXb] = usR[BXb] = Uhll_b;
      S1L = S1R = SM;
    } else {
      scrhL = (uL[C::RHO] * duL * duL - Bx * Bx) / (uL[C::RHO] * duL * (sl - SM) - Bx * Bx);
      scrhR = (uR[C::RHO] * duR * duR - Bx * Bx) / (uR[C::RHO] * duR * (sr - SM) - Bx * Bx);
      usL[BXn] = Bx;
      if constexpr (CO >= 2) usL[BXt] = uL[BXt] * scrhL;
      if constexpr (CO >= 3) usL[BXb] = uL[BXb] * scrhL;
      usR[BXn] = Bx;
      if constexpr (CO >= 2) usR[BXt] = uR[BXt] * scrhR;
      if constexpr (CO >= 3) usR[BXb] = uR[BXb] * scrhR;
    }
    scrhL = Bx / (uL[C::RHO] * duL);
    scrhR = Bx / (uR[C::RHO] * duR);
    if constexpr (CO >= 2) {
      vsL = vL[Xt] - scrhL * (usL[BXt] - uL[BXt]);
      vsR = vR[Xt] - scrhR * (usR[BXt] - uR[BXt]);
    }
    if constexpr (CO >= 3) {
      wsL = vL[Xb] - scrhL * (usL[BXb] - uL[BXb]);
      wsR = vR[Xb] - scrhR * (usR[BXb] - uR[BXb]);
    }
    usL[Xn] = usL[C::RHO] * SM;
    usR[Xn] = usR[C::RHO] * SM;
    if constexpr (CO >= 2) {
      usL[Xt] = usL[C::RHO] * vsL;
      usR[Xt] = usR[C::RHO] * vsR;
    }
    if constexpr (CO >= 3) {
      usL[Xb] = usL[C::RHO] * wsL;
      usR[Xb] = usR[C::RHO] * wsR;
    }
    scrhL = vL[Xn] * Bx;
    if constexpr (CO >= 2) scrhL = scrhL + vL[Xt] * uL[BXt];
    if constexpr (CO >= 3) scrhL = scrhL + vL[Xb] * uL[BXb];
    {
      T s = SM * Bx;
      if constexpr (CO >= 2) s = s + vsL * usL[BXt];
      if constexpr (CO >= 3) s = s + wsL * usL[BXb];
      scrhL -= s;
    }
    usL[C::ENG] = duL * uL[C::ENG] - ptL * vL[Xn] + pts * SM + Bx * scrhL;
    usL[C::ENG] /= sl - SM;
    scrhR = vR[Xn] * Bx;
    if constexpr (CO >= 2) scrhR = scrhR + vR[Xt] * uR[BXt];
    if constexpr (CO >= 3) scrhR = scrhR + vR[Xb] * uR[BXb];
    {
      T s = SM * Bx;
      if constexpr (CO >= 2) s = s + vsR * usR[BXt];
      if constexpr (CO >= 3) s = s + wsR * usR[BXb];
      scrhR -= s;
    }
    usR[C::ENG] = duR * uR[C::ENG] - ptR * vR[Xn] + pts * SM + Bx * scrhR;
    usR[C::ENG] /= sr - SM;

    if (S1L >= T(0)) {
#pragma unroll
      for (int n = 0; n < NV; n++) flux[n] = fluxL[n] + sl * (usL[n] - uL[n]);
    } else if (S1R <= T(0)) {
#pragma unroll
      for (int n = 0; n < NV; n++) flux[n] = fluxR[n] + sr * (usR[n] - uR[n]);
    } else {
      T vss = 0, wss = 0, ussl[NV], ussr[NV];
      ussl[C::RHO] = usL[C::RHO];
      ussr[C::RHO] = usR[C::RHO];
      if constexpr (CO >= 2) {
        vss = sqrL * vsL + sqrR * vsR + (usR[BXt] - usL[BXt]) * sBx;
        vss /= sqrL + sqrR;
      }
      if constexpr (CO >= 3) {
        wss = sqrL * wsL + sqrR * wsR + (usR[BXb] - usL[BXb]) * sBx;
        wss /= sqrL + sqrR;
      }
      ussl[Xn] = ussl[C::RHO] * SM;
      ussr[Xn] = ussr[C::RHO] * SM;
      if constexpr (CO >= 2) {
        ussl[Xt] = ussl[C::RHO] * vss;
        ussr[Xt] = ussr[C::RHO] * vss;
      }
      if constexpr (CO >= 3) {
        ussl[Xb] = ussl[C::RHO] * wss;
        ussr[Xb] = ussr[C::RHO] * wss;
      }
      ussl[BXn] = ussr[BXn] = Bx;
      if constexpr (CO >= 2) {
        ussl[BXt] = sqrL * usR[BXt] + sqrR * usL[BXt] + sqrL * sqrR * (vsR - vsL) * sBx;
        ussl[BXt] /= sqrL + sqrR;
        ussr[BXt] = ussl[BXt];
      }
      if constexpr (CO >= 3) {
        ussl[BXb] = sqrL * usR[BXb] + sqrR * usL[BXb] + sqrL * sqrR * (wsR - wsL) * sBx;
        ussl[BXb] /= sqrL + sqrR;
        ussr[BXb] = ussl[BXb];
      }
      scrhL = SM * Bx;
      if constexpr (CO >= 2) scrhL = scrhL + vsL * usL[BXt];
      if constexpr (CO >= 3) scrhL = scrhL + wsL * usL[BXb];
      {
        T s = SM * Bx;
        if constexpr (CO >= 2) s = s + vss * ussl[BXt];
        if constexpr (CO >= 3) s = s + wss * ussl[BXb];
        scrhL -= s;
      }
      scrhR = SM * Bx;
      if constexpr (CO >= 2) scrhR = scrhR + vsR * usR[BXt];
      if constexpr (CO >= 3) scrhR = scrhR + wsR * usR[BXb];
      {
        T s = SM * Bx;
        if constexpr (CO >= 2) s = s + vss * ussr[BXt];
        if constexpr (CO >= 3) s = s + wss * ussr[BXb];
        scrhR -= s;
      }
      ussl[C::ENG] = usL[C::ENG] - sqrL * scrhL * sBx;
      ussr[C::ENG] = usR[C::ENG] + sqrR * scrhR * sBx;
      if (SM >= T(0)) {
#pragma unroll
        for (int n = 0; n < NV; n++)
          flux[n] = fluxL[n] + S1L * (ussl[n] - usL[n]) + sl * (usL[n] - uL[n]);
      } else {
#pragma unroll
        for (int n = 0; n < NV; n++)
          flux[n] = fluxR[n] + S1R * (ussr[n] - usR[n]) + sr * (usR[n] - uR[n]);
      }
    }
  }

  store_face_emf<C, DIR>(vL, vR, uL, uR, flux, sl, sr, ptL, ptR, E, emf, pre);
}

// ------------------------------------------------------------------------------------
// RoeMHD, src/fluid/RiemannSolver/MHDsolvers/roeMHD.hpp:175-651 (HAVE_ENERGY,
// ROE_AVERAGE 0; waves fast-, fast+, div B, entropy, slow-, slow+, Alfven-, Alfven+).
// ------------------------------------------------------------------------------------
template <class C, int DIR, typename T>
__device__ __forceinline__ void roe_mhd(const T *vL, const T *vR, T *flux, T &cmax,
                                        FaceEmf<T> &E, int emf, T gamma) {
  constexpr int NV = C::NVAR;
  constexpr int CO = C::COMP;
  constexpr int Xn = C::template Xn<DIR>();
  constexpr int Xt = C::template Xt<DIR>();
  constexpr int Xb = C::template Xb<DIR>();
  constexpr int BXn = C::template BXn<DIR>();
  constexpr int BXt = C::template BXt<DIR>();
  constexpr int BXb = C::template BXb<DIR>();
  constexpr int KFASTM = 0, KFASTP = 1, KDIVB = 2, KENTRP = 3, KSLOWM = 4, KSLOWP = 5,
                KALFVM = 6, KALFVP = 7;
  const T delta = T(1.e-6);
  T dV[NV], uL[NV], uR[NV], dU[NV], fluxL[NV], fluxR[NV], Rc[NV][NV];
#pragma unroll
  for (int n = 0; n < NV; n++) dV[n] = vR[n] - vL[n];
  prim2cons<C>(vL, uL, gamma);
  prim2cons<C>(vR, uR, gamma);
#pragma unroll
  for (int n = 0; n < NV; n++) dU[n] = uR[n] - uL[n];
  kflux<C, DIR>(fluxL, vL, uL);
  kflux<C, DIR>(fluxR, vR, uR);
#pragma unroll
  for (int p = 0; p < NV; p++)
#pragma unroll
    for (int q = 0; q < NV; q++) Rc[p][q] = 0;
  const T sqr_rho_L = sqrt_(vL[C::RHO]);
  const T sqr_rho_R = sqrt_(vR[C::RHO]);
  T sl = sqr_rho_L / (sqr_rho_L + sqr_rho_R);
  T sr = sqr_rho_R / (sqr_rho_L + sqr_rho_R);
  const T rho = sr * vL[C::RHO] + sl * vR[C::RHO];
  const T sqrt_rho = sqrt_(rho);
  T u, v = 0, w = 0, Bx, By = 0, Bz = 0, bx, by = 0, bz = 0;
  u = sl * vL[Xn] + sr * vR[Xn];
  if constexpr (CO >= 2) v = sl * vL[Xt] + sr * vR[Xt];
  if constexpr (CO >= 3) w = sl * vL[Xb] + sr * vR[Xb];
  Bx = sr * vL[BXn] + sl * vR[BXn];
  if constexpr (CO >= 2) By = sr * vL[BXt] + sl * vR[BXt];
  if constexpr (CO >= 3) Bz = sr * vL[BXb] + sl * vR[BXb];
  const T sBx = (Bx >= T(0) ? T(1) : T(-1));
  bx = Bx / sqrt_rho;
  if constexpr (CO >= 2) by = By / sqrt_rho;
  if constexpr (CO >= 3) bz = Bz / sqrt_rho;
  T bt2 = T(0);
  if constexpr (CO >= 2) bt2 = bt2 + by * by;
  if constexpr (CO >= 3) bt2 = bt2 + bz * bz;
  const T b2 = bx * bx + bt2;
  const T Btmag = sqrt_(bt2 * rho);
  T X = dV[BXn] * dV[BXn];
  if constexpr (CO >= 2) X = X + dV[BXt] * dV[BXt];
  if constexpr (CO >= 3) X = X + dV[BXb] * dV[BXb];
  X /= (sqr_rho_L + sqr_rho_R) * (sqr_rho_L + sqr_rho_R) * T(2);
  T Bmag2L = vL[C::BX1] * vL[C::BX1];
  if constexpr (CO >= 2) Bmag2L = Bmag2L + vL[C::BX2] * vL[C::BX2];
  if constexpr (CO >= 3) Bmag2L = Bmag2L + vL[C::BX3] * vL[C::BX3];
  T Bmag2R = vR[C::BX1] * vR[C::BX1];
  if constexpr (CO >= 2) Bmag2R = Bmag2R + vR[C::BX2] * vR[C::BX2];
  if constexpr (CO >= 3) Bmag2R = Bmag2R + vR[C::BX3] * vR[C::BX3];
  const T pL = vL[C::ENG] + T(0.5) * Bmag2L;
  const T pR = vR[C::ENG] + T(0.5) * Bmag2R;
  T vdm = u * dU[Xn];
  if constexpr (CO >= 2) vdm = vdm + v * dU[Xt];
  if constexpr (CO >= 3) vdm = vdm + w * dU[Xb];
  T BdB = Bx * dU[BXn];
  if constexpr (CO >= 2) BdB = BdB + By * dU[BXt];
  if constexpr (CO >= 3) BdB = BdB + Bz * dU[BXb];
  T vel2 = u * u;
  if constexpr (CO >= 2) vel2 = vel2 + v * v;
  if constexpr (CO >= 3) vel2 = vel2 + w * w;
  dV[C::ENG] = (gamma - T(1)) * ((T(0.5) * vel2 - X) * dV[C::RHO] - vdm + dU[C::ENG] - BdB);
  const T HL = (uL[C::ENG] + pL) / vL[C::RHO];
  const T HR = (uR[C::ENG] + pR) / vR[C::RHO];
  const T H = sl * HL + sr * HR;
  const T Hgas = H - b2;
  const T a2 = (T(2) - gamma) * X + (gamma - T(1)) * (Hgas - T(0.5) * vel2);
  T scrh = a2 - b2;
  const T ca2 = bx * bx;
  scrh = scrh * scrh + T(4) * bt2 * a2;
  scrh = sqrt_(scrh);
  const T cf2 = T(0.5) * (a2 + b2 + scrh);
  const T cs2 = a2 * ca2 / cf2;
  const T cf = sqrt_(cf2);
  const T cs = sqrt_(cs2);
  const T ca = sqrt_(ca2);
  const T a = sqrt_(a2);
  (void)ca;
  T alpha_f, alpha_s;
  if (cf == cs) {
    alpha_f = T(1);
    alpha_s = T(0);
  } else if (a <= cs) {
    alpha_f = T(0);
    alpha_s = T(1);
  } else if (cf <= a) {
    alpha_f = T(1);
    alpha_s = T(0);
  } else {
    scrh = T(1) / (cf2 - cs2);
    alpha_f = (a2 - cs2) * scrh;
    alpha_s = (cf2 - a2) * scrh;
    alpha_f = fmx(T(0), alpha_f);
    alpha_s = fmx(T(0), alpha_s);
    alpha_f = sqrt_(alpha_f);
    alpha_s = sqrt_(alpha_s);
  }
  T beta_y = 0, beta_z = 0;
  if (Btmag > T(1.e-9)) {
    if constexpr (CO == 2) beta_y = (By >= T(0) ? T(1) : T(-1));
    if constexpr (CO == 3) {
      beta_y = By / Btmag;
      beta_z = Bz / Btmag;
    }
  } else {
    if constexpr (CO == 2) beta_y = T(1);
    if constexpr (CO == 3) beta_z = beta_y = T(1);
  }
  T lambda[NV], eta[NV];
  // fast wave u - c_f
  lambda[KFASTM] = u - cf;
  scrh = alpha_s * cs * sBx;
  T beta_dv = T(0), beta_dB = T(0), beta_v = T(0);
  if constexpr (CO >= 2) {
    beta_dv = beta_dv + beta_y * dV[Xt];
    beta_dB = beta_dB + beta_y * dV[BXt];
    beta_v = beta_v + beta_y * v;
  }
  if constexpr (CO >= 3) {
    beta_dv = beta_dv + beta_z * dV[Xb];
    beta_dB = beta_dB + beta_z * dV[BXb];
    beta_v = beta_v + beta_z * w;
  }
  Rc[C::RHO][KFASTM] = alpha_f;
  Rc[Xn][KFASTM] = alpha_f * lambda[KFASTM];
  if constexpr (CO >= 2) Rc[Xt][KFASTM] = alpha_f * v + scrh * beta_y;
  if constexpr (CO >= 3) Rc[Xb][KFASTM] = alpha_f * w + scrh * beta_z;
  if constexpr (CO >= 2) Rc[BXt][KFASTM] = alpha_s * a * beta_y / sqrt_rho;
  if constexpr (CO >= 3) Rc[BXb][KFASTM] = alpha_s * a * beta_z / sqrt_rho;
  Rc[C::ENG][KFASTM] = alpha_f * (Hgas - u * cf) + scrh * beta_v + alpha_s * a * Btmag / sqrt_rho;
  eta[KFASTM] = alpha_f * (X * dV[C::RHO] + dV[C::ENG]);
  eta[KFASTM] += rho * scrh * beta_dv - rho * alpha_f * cf * dV[Xn] + sqrt_rho * alpha_s * a * beta_dB;
  eta[KFASTM] *= T(0.5) / a2;
  // fast wave u + c_f
  lambda[KFASTP] = u + cf;
  Rc[C::RHO][KFASTP] = alpha_f;
  Rc[Xn][KFASTP] = alpha_f * lambda[KFASTP];
  if constexpr (CO >= 2) Rc[Xt][KFASTP] = alpha_f * v - scrh * beta_y;
  if constexpr (CO >= 3) Rc[Xb][KFASTP] = alpha_f * w - scrh * beta_z;
  if constexpr (CO >= 2) Rc[BXt][KFASTP] = Rc[BXt][KFASTM];
  if constexpr (CO >= 3) Rc[BXb][KFASTP] = Rc[BXb][KFASTM];
  Rc[C::ENG][KFASTP] = alpha_f * (Hgas + u * cf) - scrh * beta_v + alpha_s * a * Btmag / sqrt_rho;
  eta[KFASTP] = alpha_f * (X * dV[C::RHO] + dV[C::ENG]) - rho * scrh * beta_dv +
                rho * alpha_f * cf * dV[Xn] + sqrt_rho * alpha_s * a * beta_dB;
  eta[KFASTP] *= T(0.5) / a2;
  // entropy wave u
  lambda[KENTRP] = u;
  Rc[C::RHO][KENTRP] = T(1);
  Rc[Xn][KENTRP] = u;
  if constexpr (CO >= 2) Rc[Xt][KENTRP] = v;
  if constexpr (CO >= 3) Rc[Xb][KENTRP] = w;
  Rc[C::ENG][KENTRP] = T(0.5) * vel2 + (gamma - T(2)) / (gamma - T(1)) * X;
  eta[KENTRP] = ((a2 - X) * dV[C::RHO] - dV[C::ENG]) / a2;
  // div B wave u (zero jump)
  lambda[KDIVB] = u;
  eta[KDIVB] = T(0);
  if constexpr (CO > 1) {
    scrh = alpha_f * cf * sBx;
    lambda[KSLOWM] = u - cs;
    Rc[C::RHO][KSLOWM] = alpha_s;
    Rc[Xn][KSLOWM] = alpha_s * lambda[KSLOWM];
    Rc[Xt][KSLOWM] = alpha_s * v - scrh * beta_y;
    if constexpr (CO >= 3) Rc[Xb][KSLOWM] = alpha_s * w - scrh * beta_z;
    Rc[BXt][KSLOWM] = -alpha_f * a * beta_y / sqrt_rho;
    if constexpr (CO >= 3) Rc[BXb][KSLOWM] = -alpha_f * a * beta_z / sqrt_rho;
    Rc[C::ENG][KSLOWM] = alpha_s * (Hgas - u * cs) - scrh * beta_v - alpha_f * a * Btmag / sqrt_rho;
    eta[KSLOWM] = alpha_s * (X * dV[C::RHO] + dV[C::ENG]) - rho * scrh * beta_dv -
                  rho * alpha_s * cs * dV[Xn] - sqrt_rho * alpha_f * a * beta_dB;
    eta[KSLOWM] *= T(0.5) / a2;
    lambda[KSLOWP] = u + cs;
    Rc[C::RHO][KSLOWP] = alpha_s;
    Rc[Xn][KSLOWP] = alpha_s * lambda[KSLOWP];
    Rc[Xt][KSLOWP] = alpha_s * v + scrh * beta_y;
    if constexpr (CO >= 3) Rc[Xb][KSLOWP] = alpha_s * w + scrh * beta_z;
    Rc[BXt][KSLOWP] = Rc[BXt][KSLOWM];
    if constexpr (CO >= 3) Rc[BXb][KSLOWP] = Rc[BXb][KSLOWM];
    Rc[C::ENG][KSLOWP] = alpha_s * (Hgas + u * cs) + scrh * beta_v - alpha_f * a * Btmag / sqrt_rho;
    eta[KSLOWP] = alpha_s * (X * dV[C::RHO] + dV[C::ENG]) + rho * scrh * beta_dv +
                  rho * alpha_s * cs * dV[Xn] - sqrt_rho * alpha_f * a * beta_dB;
    eta[KSLOWP] *= T(0.5) / a2;
  }
  if constexpr (CO == 3) {
    lambda[KALFVM] = u - ca;
    Rc[Xt][KALFVM] = -rho * beta_z;
    Rc[Xb][KALFVM] = +rho * beta_y;
    Rc[BXt][KALFVM] = -sBx * sqrt_rho * beta_z;
    Rc[BXb][KALFVM] = sBx * sqrt_rho * beta_y;
    Rc[C::ENG][KALFVM] = -rho * (v * beta_z - w * beta_y);
    eta[KALFVM] = +beta_y * dV[Xb] - beta_z * dV[Xt] +
                  sBx / sqrt_rho * (beta_y * dV[BXb] - beta_z * dV[BXt]);
    eta[KALFVM] *= T(0.5);
    lambda[KALFVP] = u + ca;
    Rc[Xt][KALFVP] = -Rc[Xt][KALFVM];
    Rc[Xb][KALFVP] = -Rc[Xb][KALFVM];
    Rc[BXt][KALFVP] = Rc[BXt][KALFVM];
    Rc[BXb][KALFVP] = Rc[BXb][KALFVM];
    Rc[C::ENG][KALFVP] = -Rc[C::ENG][KALFVM];
    eta[KALFVP] = -beta_y * dV[Xb] + beta_z * dV[Xt] +
                  sBx / sqrt_rho * (beta_y * dV[BXb] - beta_z * dV[BXt]);
    eta[KALFVP] *= T(0.5);
  }
  cmax = fab(u) + cf;
  sl = lambda[KFASTM];
  sr = lambda[KFASTP];
  T alambda[NV];
#pragma unroll
  for (int n = 0; n < NV; n++) alambda[n] = fab(lambda[n]);
  if (alambda[KFASTM] < T(0.5) * delta)
    alambda[KFASTM] = lambda[KFASTM] * lambda[KFASTM] / delta + T(0.25) * delta;
  if (alambda[KFASTP] < T(0.5) * delta)
    alambda[KFASTP] = lambda[KFASTP] * lambda[KFASTP] / delta + T(0.25) * delta;
  if constexpr (CO > 1) {
    if (alambda[KSLOWM] < T(0.5) * delta)
      alambda[KSLOWM] = lambda[KSLOWM] * lambda[KSLOWM] / delta + T(0.25) * delta;
    if (alambda[KSLOWP] < T(0.5) * delta)
      alambda[KSLOWP] = lambda[KSLOWP] * lambda[KSLOWP] / delta + T(0.25) * delta;
  }
#pragma unroll
  for (int n1 = 0; n1 < NV; n1++) {
    T sc = T(0);
#pragma unroll
    for (int n2 = 0; n2 < NV; n2++) sc += alambda[n2] * eta[n2] * Rc[n1][n2];
    flux[n1] = T(0.5) * (fluxL[n1] + fluxR[n1] - sc);
  }
  StarPre<T> pre;
  pre.valid = false;
  store_face_emf<C, DIR>(vL, vR, uL, uR, flux, sl, sr, pL, pR, E, emf, pre);
}

// Fast magnetosonic speed of one state (hlldMHD.hpp:183-226, hllMHD.hpp:205-231).
template <class C, int DIR, typename T>
__device__ __forceinline__ T fast_speed(const T *v, T gamma) {
  constexpr int CO = C::COMP;
  const T gpr = gamma * v[C::ENG];
  T b1 = v[C::template BXn<DIR>()], b2 = 0, b3 = 0;
  if constexpr (CO >= 2) b2 = v[C::template BXt<DIR>()];
  if constexpr (CO >= 3) b3 = v[C::template BXb<DIR>()];
  const T Btmag2 = b2 * b2 + b3 * b3;
  const T Bmag2 = b1 * b1 + Btmag2;
  T c = gpr - Bmag2;
  c = gpr + Bmag2 + sqrt_(c * c + T(4) * gpr * Btmag2);
  return sqrt_(T(0.5) * c / v[C::RHO]);
}

// ------------------------------------------------------------------------------------
// HllMHD, src/fluid/RiemannSolver/MHDsolvers/hllMHD.hpp:170-453 (HAVE_ENERGY, no Hall).
// Density and momentum fluxes divide by (sr - sl); field and energy fluxes multiply by
// 1/(sr - sl).  The reference has no uct_hlld face store for this solver, so the
// context rejects that pairing.
// ------------------------------------------------------------------------------------
template <class C, int DIR, typename T>
__device__ __forceinline__ void hll_mhd(const T *vL, const T *vR, T *flux, T &cmax,
                                        FaceEmf<T> &E, int emf, T gamma) {
  constexpr int NV = C::NVAR;
  constexpr int Xn = C::template Xn<DIR>();
  T uL[NV], uR[NV], fluxL[NV], fluxR[NV];
  const T cL = fast_speed<C, DIR>(vL, gamma);
  const T cR = fast_speed<C, DIR>(vR, gamma);
  const T sl = fmn(vL[Xn] - cL, vR[Xn] - cR);
  const T sr = fmx(vL[Xn] + cL, vR[Xn] + cR);
  cmax = fmx(fab(sl), fab(sr));
  prim2cons<C>(vL, uL, gamma);
  prim2cons<C>(vR, uR, gamma);
  kflux<C, DIR>(fluxL, vL, uL);
  kflux<C, DIR>(fluxR, vR, uR);
  if (sl > T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fluxL[n];
  } else if (sr < T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fluxR[n];
  } else {
    const T sls = sl * sr, dS = sr - sl, rdS = T(1) / dS;
#pragma unroll
    for (int n = 0; n < NV; n++) {
      const T f = sls * uR[n] - sls * uL[n] + sr * fluxL[n] - sl * fluxR[n];
      flux[n] = (n < C::ENG) ? f / dS : f * rdS;
    }
  }
  StarPre<T> pre;
  pre.valid = false;
  // uct_hlld is rejected at context creation; the total pressures are unused below
  store_face_emf<C, DIR>(vL, vR, uL, uR, flux, sl, sr, T(0), T(0), E,
                         emf == EMF_UCT_HLLD ? EMF_UCT_HLL : emf, pre);
}

// ------------------------------------------------------------------------------------
// TvdlfMHD, src/fluid/RiemannSolver/MHDsolvers/tvdlfMHD.hpp:138-234 (HAVE_ENERGY, no
// Hall): Rusanov flux with the fast speed of the arithmetic mean state; the uct_hll
// store sees sl = -cmax, sr = cmax.  No uct_hlld store (rejected at context creation).
// ------------------------------------------------------------------------------------
template <class C, int DIR, typename T>
__device__ __forceinline__ void tvdlf_mhd(const T *vL, const T *vR, T *flux, T &cmax,
                                          FaceEmf<T> &E, int emf, T gamma) {
  constexpr int NV = C::NVAR;
  constexpr int CO = C::COMP;
  constexpr int Xn = C::template Xn<DIR>();
  T uL[NV], uR[NV], fluxL[NV], fluxR[NV];
  const T rho = T(0.5) * (vL[C::RHO] + vR[C::RHO]);
  const T gpr = gamma * (T(0.5) * (vL[C::ENG] + vR[C::ENG]));
  const T bn = vL[C::template BXn<DIR>()];  // vL and vR share the face value
  T Bt2 = T(0);
  if constexpr (CO >= 2) {
    constexpr int BXt = C::template BXt<DIR>();
    const T bt = T(0.5) * (vL[BXt] + vR[BXt]);
    Bt2 = Bt2 + bt * bt;
  }
  if constexpr (CO >= 3) {
    constexpr int BXb = C::template BXb<DIR>();
    const T bb = T(0.5) * (vL[BXb] + vR[BXb]);
    Bt2 = Bt2 + bb * bb;
  }
  const T B2 = Bt2 + bn * bn;
  T cRL = gpr - B2;
  cRL = cRL + B2 + sqrt_(cRL * cRL + T(4) * gpr * Bt2);
  cRL = sqrt_(T(0.5) * cRL / rho);
  const T vn = T(0.5) * (vL[Xn] + vR[Xn]);
  cmax = fmx(fab(vn + cRL), fab(vn - cRL));
  prim2cons<C>(vL, uL, gamma);
  prim2cons<C>(vR, uR, gamma);
  kflux<C, DIR>(fluxL, vL, uL);
  kflux<C, DIR>(fluxR, vR, uR);
#pragma unroll
  for (int n = 0; n < NV; n++) flux[n] = T(0.5) * (fluxL[n] + fluxR[n] + cmax * (uL[n] - uR[n]));
  StarPre<T> pre;
  pre.valid = false;
  store_face_emf<C, DIR>(vL, vR, uL, uR, flux, -cmax, cmax, T(0), T(0), E,
                         emf == EMF_UCT_HLLD ? EMF_UCT_HLL : emf, pre);
}

// ====================================================================================
// Isothermal branches of the solvers (HAVE_ENERGY == 0): the argument `cs` takes the
// place of gamma and is the sound speed of the face, the average of its two cells' speeds
// (computed by the caller, eos_face); Cs2Iso is its square.
// ====================================================================================

// HllHD (hllHD.hpp:67-70) and TvdlfHD (tvdlfHD.hpp:72-75), isothermal
template <class C, int DIR, typename T>
__device__ __forceinline__ void hll_hd_iso(const T *vL, const T *vR, T *flux, T &cmax, T cs) {
  constexpr int NV = C::NVAR;
  constexpr int Xn = C::template Xn<DIR>();
  T uL[NV], uR[NV], fL[NV], fR[NV];
  const T cL = cs, cR = cL;
  const T SL = fmn(vL[Xn] - cL, vR[Xn] - cR);
  const T SR = fmx(vL[Xn] + cL, vR[Xn] + cR);
  cmax = fmx(fab(SL), fab(SR));
  prim2cons<C>(vL, uL, cs);
  prim2cons<C>(vR, uR, cs);
  kflux<C, DIR>(fL, vL, uL, cL * cL);
  kflux<C, DIR>(fR, vR, uR, cR * cR);
  if (SL > T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fL[n];
  } else if (SR < T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fR[n];
  } else {
    const T sls = SL * SR, dS = SR - SL;
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = (sls * uR[n] - sls * uL[n] + SR * fL[n] - SL * fR[n]) / dS;
  }
}

template <class C, int DIR, typename T>
__device__ __forceinline__ void tvdlf_hd_iso(const T *vL, const T *vR, T *flux, T &cmax, T cs) {
  constexpr int NV = C::NVAR;
  constexpr int Xn = C::template Xn<DIR>();
  T uL[NV], uR[NV], fL[NV], fR[NV];
  const T cRL = cs;
  const T vn = T(0.5) * (vL[Xn] + vR[Xn]);
  cmax = fmx(fab(vn + cRL), fab(vn - cRL));
  prim2cons<C>(vL, uL, cs);
  prim2cons<C>(vR, uR, cs);
  kflux<C, DIR>(fL, vL, uL, cRL * cRL);
  kflux<C, DIR>(fR, vR, uR, cRL * cRL);
#pragma unroll
  for (int n = 0; n < NV; n++) flux[n] = T(0.5) * (fL[n] + fR[n] - cmax * (uR[n] - uL[n]));
}

// HllcHD isothermal star states (hllcHD.hpp:129-143)
template <class C, int DIR, typename T>
__device__ __forceinline__ void hllc_hd_iso(const T *vL, const T *vR, T *flux, T &cmax, T cs) {
  constexpr int NV = C::NVAR;
  constexpr int Xn = C::template Xn<DIR>();
  constexpr int Xt = C::template Xt<DIR>();
  constexpr int Xb = C::template Xb<DIR>();
  T uL[NV], uR[NV], fL[NV], fR[NV];
  const T cL = cs, cR = cL;
  const T SL = fmn(vL[Xn] - cL, vR[Xn] - cR);
  const T SR = fmx(vL[Xn] + cL, vR[Xn] + cR);
  cmax = fmx(fab(SL), fab(SR));
  prim2cons<C>(vL, uL, cs);
  prim2cons<C>(vR, uR, cs);
  kflux<C, DIR>(fL, vL, uL, cL * cL);
  kflux<C, DIR>(fR, vR, uR, cR * cR);
  if (SL > T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fL[n];
  } else if (SR < T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fR[n];
  } else {
    T usL[NV], usR[NV];
    const T scrh = T(1) / (SR - SL);
    const T rho = (SR * uR[C::RHO] - SL * uL[C::RHO] - fR[C::RHO] + fL[C::RHO]) * scrh;
    const T mx = (SR * uR[Xn] - SL * uL[Xn] - fR[Xn] + fL[Xn]) * scrh;
    usL[C::RHO] = usR[C::RHO] = rho;
    usL[Xn] = usR[Xn] = mx;
    T vs = (SR * fL[C::RHO] - SL * fR[C::RHO] + SR * SL * (uR[C::RHO] - uL[C::RHO]));
    vs *= scrh;
    vs /= rho;
    if constexpr (C::COMP >= 2) {
      usL[Xt] = rho * vL[Xt];
      usR[Xt] = rho * vR[Xt];
    }
    if constexpr (C::COMP >= 3) {
      usL[Xb] = rho * vL[Xb];
      usR[Xb] = rho * vR[Xb];
    }
    if (vs >= T(0)) {
#pragma unroll
      for (int n = 0; n < NV; n++) flux[n] = fL[n] + SL * (usL[n] - uL[n]);
    } else {
#pragma unroll
      for (int n = 0; n < NV; n++) flux[n] = fR[n] + SR * (usR[n] - uR[n]);
    }
  }
}

// RoeHD isothermal branches (roeHD.hpp:29-36, :95-99, :164-167, :186-194, :254-264)
template <class C, int DIR, typename T>
__device__ __forceinline__ bool roe_hd_iso(const T *vL, const T *vR, T *flux, T &cmax, T cs) {
  constexpr int NV = C::NVAR;
  constexpr int Xn = C::template Xn<DIR>();
  constexpr int Xt = C::template Xt<DIR>();
  constexpr int Xb = C::template Xb<DIR>();
  constexpr int I0 = 0, I1 = 1, I2 = 2, I3 = 3;
  T dv[NV], uL[NV], uR[NV], fL[NV], fR[NV], um[NV], Rc[NV][NV];
  T lambda[NV], alambda[NV], eta[NV];
#pragma unroll
  for (int n = 0; n < NV; n++) dv[n] = vR[n] - vL[n];
  T a2L = cs;
  T a2R = a2L;
  a2L = a2L * a2L;
  a2R = a2R * a2R;
  prim2cons<C>(vL, uL, cs);
  prim2cons<C>(vR, uR, cs);
  kflux<C, DIR>(fL, vL, uL, a2L);
  kflux<C, DIR>(fR, vR, uR, a2R);
#pragma unroll
  for (int n = 0; n < NV; n++) um[n] = T(0.5) * (vR[n] + vL[n]);
  const T a2 = T(0.5) * (a2L + a2R);
  const T a = sqrt_(a2);
#pragma unroll
  for (int n1 = 0; n1 < NV; n1++)
#pragma unroll
    for (int n2 = 0; n2 < NV; n2++) Rc[n1][n2] = T(0);
  lambda[I0] = um[Xn] - a;
  eta[I0] = T(0.5) * (dv[C::RHO] - um[C::RHO] * dv[Xn] / a);
  Rc[C::RHO][I0] = T(1);
  Rc[Xn][I0] = um[Xn] - a;
  if constexpr (C::COMP >= 2) Rc[Xt][I0] = um[Xt];
  if constexpr (C::COMP >= 3) Rc[Xb][I0] = um[Xb];
  lambda[I1] = um[Xn] + a;
  eta[I1] = T(0.5) * (dv[C::RHO] + um[C::RHO] * dv[Xn] / a);
  Rc[C::RHO][I1] = T(1);
  Rc[Xn][I1] = um[Xn] + a;
  if constexpr (C::COMP >= 2) Rc[Xt][I1] = um[Xt];
  if constexpr (C::COMP >= 3) Rc[Xb][I1] = um[Xb];
  if constexpr (C::COMP > 1) {
    lambda[I2] = um[Xn];
    eta[I2] = um[C::RHO] * dv[Xt];
    Rc[Xt][I2] = T(1);
  }
  if constexpr (C::COMP > 2) {
    lambda[I3] = um[Xn];
    eta[I3] = um[C::RHO] * dv[Xb];
    Rc[Xb][I3] = T(1);
  }
  cmax = fab(um[Xn]) + a;
  T scrh = fab(vL[C::RHO] - vR[C::RHO]);
  scrh /= fmn(vL[C::RHO], vR[C::RHO]);
  scrh *= a * a;
  if (scrh > T(0.5) && (vR[Xn] < vL[Xn])) {
    if constexpr (C::DIM > 1) {
      const T bmin = fmn(T(0), lambda[0]);
      const T bmax = fmx(T(0), lambda[1]);
      const T scrh1 = T(1) / (bmax - bmin);
#pragma unroll
      for (int n = 0; n < NV; n++) {
        flux[n] = bmin * bmax * (uR[n] - uL[n]) + bmax * fL[n] - bmin * fR[n];
        flux[n] *= scrh1;
      }
      return true;
    } else {
      return false;
    }
  }
#pragma unroll
  for (int n = 0; n < NV; n++) alambda[n] = fab(lambda[n]);
  const T delta = T(1.e-7);
  if (alambda[0] <= delta) alambda[0] = T(0.5) * lambda[0] * lambda[0] / delta + T(0.5) * delta;
  if (alambda[1] <= delta) alambda[1] = T(0.5) * lambda[1] * lambda[1] / delta + T(0.5) * delta;
#pragma unroll
  for (int n = 0; n < NV; n++) {
    T f = fL[n] + fR[n];
#pragma unroll
    for (int n2 = 0; n2 < NV; n2++) f -= alambda[n2] * eta[n2] * Rc[n][n2];
    flux[n] = f * T(0.5);
  }
  return true;
}

// Fast speed of one state with gpr = Cs2Iso rho (hlldMHD.hpp:188-226 isothermal)
template <class C, int DIR, typename T>
__device__ __forceinline__ T fast_speed_iso(const T *v, T c2) {
  constexpr int CO = C::COMP;
  const T gpr = c2 * v[C::RHO];
  T b1 = v[C::template BXn<DIR>()], b2 = 0, b3 = 0;
  if constexpr (CO >= 2) b2 = v[C::template BXt<DIR>()];
  if constexpr (CO >= 3) b3 = v[C::template BXb<DIR>()];
  const T Btmag2 = b2 * b2 + b3 * b3;
  const T Bmag2 = b1 * b1 + Btmag2;
  T c = gpr - Bmag2;
  c = gpr + Bmag2 + sqrt_(c * c + T(4) * gpr * Btmag2);
  return sqrt_(T(0.5) * c / v[C::RHO]);
}

template <class C, typename T>
__device__ __forceinline__ T total_pressure_iso(const T *v, T c2) {
  T b2 = v[C::BX1] * v[C::BX1];
  if constexpr (C::COMP >= 2) b2 = b2 + v[C::BX2] * v[C::BX2];
  if constexpr (C::COMP >= 3) b2 = b2 + v[C::BX3] * v[C::BX3];
  return c2 * v[C::RHO] + T(0.5) * b2;
}

// HllMHD isothermal (hllMHD.hpp:213-231)
template <class C, int DIR, typename T>
__device__ __forceinline__ void hll_mhd_iso(const T *vL, const T *vR, T *flux, T &cmax,
                                            FaceEmf<T> &E, int emf, T cs) {
  constexpr int NV = C::NVAR;
  constexpr int Xn = C::template Xn<DIR>();
  T uL[NV], uR[NV], fluxL[NV], fluxR[NV];
  const T c2 = cs * cs;
  const T cL = fast_speed_iso<C, DIR>(vL, c2);
  const T cR = fast_speed_iso<C, DIR>(vR, c2);
  const T sl = fmn(vL[Xn] - cL, vR[Xn] - cR);
  const T sr = fmx(vL[Xn] + cL, vR[Xn] + cR);
  cmax = fmx(fab(sl), fab(sr));
  prim2cons<C>(vL, uL, cs);
  prim2cons<C>(vR, uR, cs);
  kflux<C, DIR>(fluxL, vL, uL, c2);
  kflux<C, DIR>(fluxR, vR, uR, c2);
  if (sl > T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fluxL[n];
  } else if (sr < T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fluxR[n];
  } else {
    const T sls = sl * sr, dS = sr - sl, rdS = T(1) / dS;
#pragma unroll
    for (int n = 0; n < NV; n++) {
      const T f = sls * uR[n] - sls * uL[n] + sr * fluxL[n] - sl * fluxR[n];
      flux[n] = (n < C::BX1) ? f / dS : f * rdS;
    }
  }
  StarPre<T> pre;
  pre.valid = false;
  store_face_emf<C, DIR>(vL, vR, uL, uR, flux, sl, sr, total_pressure_iso<C>(vL, c2),
                         total_pressure_iso<C>(vR, c2), E,
                         emf == EMF_UCT_HLLD ? EMF_UCT_HLL : emf, pre);
}

// TvdlfMHD isothermal (tvdlfMHD.hpp:179-185)
template <class C, int DIR, typename T>
__device__ __forceinline__ void tvdlf_mhd_iso(const T *vL, const T *vR, T *flux, T &cmax,
                                              FaceEmf<T> &E, int emf, T cs) {
  constexpr int NV = C::NVAR;
  constexpr int CO = C::COMP;
  constexpr int Xn = C::template Xn<DIR>();
  T uL[NV], uR[NV], fluxL[NV], fluxR[NV];
  const T c2 = cs * cs;
  const T rho = T(0.5) * (vL[C::RHO] + vR[C::RHO]);
  const T gpr = c2 * rho;
  const T bn = vL[C::template BXn<DIR>()];
  T Bt2 = T(0);
  if constexpr (CO >= 2) {
    constexpr int BXt = C::template BXt<DIR>();
    const T bt = T(0.5) * (vL[BXt] + vR[BXt]);
    Bt2 = Bt2 + bt * bt;
  }
  if constexpr (CO >= 3) {
    constexpr int BXb = C::template BXb<DIR>();
    const T bb = T(0.5) * (vL[BXb] + vR[BXb]);
    Bt2 = Bt2 + bb * bb;
  }
  const T B2 = Bt2 + bn * bn;
  T cRL = gpr - B2;
  cRL = cRL + B2 + sqrt_(cRL * cRL + T(4) * gpr * Bt2);
  cRL = sqrt_(T(0.5) * cRL / rho);
  const T vn = T(0.5) * (vL[Xn] + vR[Xn]);
  cmax = fmx(fab(vn + cRL), fab(vn - cRL));
  prim2cons<C>(vL, uL, cs);
  prim2cons<C>(vR, uR, cs);
  kflux<C, DIR>(fluxL, vL, uL, c2);
  kflux<C, DIR>(fluxR, vR, uR, c2);
#pragma unroll
  for (int n = 0; n < NV; n++) flux[n] = T(0.5) * (fluxL[n] + fluxR[n] + cmax * (uL[n] - uR[n]));
  StarPre<T> pre;
  pre.valid = false;
  store_face_emf<C, DIR>(vL, vR, uL, uR, flux, -cmax, cmax, T(0), T(0), E,
                         emf == EMF_UCT_HLLD ? EMF_UCT_HLL : emf, pre);
}

// HlldMHD isothermal (hlldMHD.hpp:459-571, Mignone 2007)
template <class C, int DIR, typename T>
__device__ __forceinline__ void hlld_mhd_iso(const T *vL, const T *vR, T *flux, T &cmax,
                                             FaceEmf<T> &E, int emf, T cs) {
  constexpr int NV = C::NVAR;
  constexpr int CO = C::COMP;
  constexpr int Xn = C::template Xn<DIR>();
  constexpr int Xt = C::template Xt<DIR>();
  constexpr int Xb = C::template Xb<DIR>();
  constexpr int BXn = C::template BXn<DIR>();
  constexpr int BXt = C::template BXt<DIR>();
  constexpr int BXb = C::template BXb<DIR>();
  T uL[NV], uR[NV], fluxL[NV], fluxR[NV];
  const T c2 = cs * cs;
  const T cL = fast_speed_iso<C, DIR>(vL, c2);
  const T cR = fast_speed_iso<C, DIR>(vR, c2);
  const T sl = fmn(vL[Xn] - cL, vR[Xn] - cR);
  const T sr = fmx(vL[Xn] + cL, vR[Xn] + cR);
  cmax = fmx(fab(sl), fab(sr));
  prim2cons<C>(vL, uL, cs);
  prim2cons<C>(vR, uR, cs);
  kflux<C, DIR>(fluxL, vL, uL, c2);
  kflux<C, DIR>(fluxR, vR, uR, c2);
  if (sl > T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fluxL[n];
  } else if (sr < T(0)) {
#pragma unroll
    for (int n = 0; n < NV; n++) flux[n] = fluxR[n];
  } else {
    T scrh = T(1) / (sr - sl);
    const T duL = sl - vL[Xn];
    const T duR = sr - vR[Xn];
    const T Bx = (sr * vR[BXn] - sl * vL[BXn]) * scrh;
    const T rho = (uR[C::RHO] * duR - uL[C::RHO] * duL) * scrh;
    flux[C::RHO] = (sl * uR[C::RHO] * duR - sr * uL[C::RHO] * duL) * scrh;
    const T sqrho = sqrt_(rho);
    const T SM = flux[C::RHO] / rho;
    const T S1L = SM - fab(Bx) / sqrho;
    const T S1R = SM + fab(Bx) / sqrho;
    bool revert_to_hll = false;
    if ((S1L - sl) < T(1.e-4) * (sr - sl)) revert_to_hll = true;
    if ((S1R - sr) > -T(1.e-4) * (sr - sl)) revert_to_hll = true;
    if (revert_to_hll) {
      scrh = T(1) / (sr - sl);
#pragma unroll
      for (int n = 0; n < NV; n++) {
        flux[n] = sl * sr * (uR[n] - uL[n]) + sr * fluxL[n] - sl * fluxR[n];
        flux[n] *= scrh;
      }
    } else {
      flux[Xn] = (sr * fluxL[Xn] - sl * fluxR[Xn] + sr * sl * (uR[Xn] - uL[Xn])) * scrh;
      flux[BXn] = sr * sl * (uR[BXn] - uL[BXn]) * scrh;
      const T scrhL = T(1) / ((sl - S1L) * (sl - S1R));
      const T scrhR = T(1) / ((sr - S1L) * (sr - S1R));
      T usLt = 0, usRt = 0, usLbt = 0, usRbt = 0, usLb = 0, usRb = 0, usLbb = 0, usRbb = 0;
      if constexpr (CO >= 2) {
        usLt = rho * vL[Xt] - Bx * uL[BXt] * (SM - vL[Xn]) * scrhL;
        usRt = rho * vR[Xt] - Bx * uR[BXt] * (SM - vR[Xn]) * scrhR;
        usLbt = uL[BXt] / rho * (uL[C::RHO] * duL * duL - Bx * Bx) * scrhL;
        usRbt = uR[BXt] / rho * (uR[C::RHO] * duR * duR - Bx * Bx) * scrhR;
      }
      if constexpr (CO >= 3) {
        usLb = rho * vL[Xb] - Bx * uL[BXb] * (SM - vL[Xn]) * scrhL;
        usRb = rho * vR[Xb] - Bx * uR[BXb] * (SM - vR[Xn]) * scrhR;
        usLbb = uL[BXb] / rho * (uL[C::RHO] * duL * duL - Bx * Bx) * scrhL;
        usRbb = uR[BXb] / rho * (uR[C::RHO] * duR * duR - Bx * Bx) * scrhR;
      }
      if (S1L >= T(0)) {
        if constexpr (CO >= 2) {
          flux[Xt] = fluxL[Xt] + sl * (usLt - uL[Xt]);
          flux[BXt] = fluxL[BXt] + sl * (usLbt - uL[BXt]);
        }
        if constexpr (CO >= 3) {
          flux[Xb] = fluxL[Xb] + sl * (usLb - uL[Xb]);
          flux[BXb] = fluxL[BXb] + sl * (usLbb - uL[BXb]);
        }
      } else if (S1R <= T(0)) {
        if constexpr (CO >= 2) {
          flux[Xt] = fluxR[Xt] + sr * (usRt - uR[Xt]);
          flux[BXt] = fluxR[BXt] + sr * (usRbt - uR[BXt]);
        }
        if constexpr (CO >= 3) {
          flux[Xb] = fluxR[Xb] + sr * (usRb - uR[Xb]);
          flux[BXb] = fluxR[BXb] + sr * (usRbb - uR[BXb]);
        }
      } else {
        const T sBx = (Bx > T(0) ? T(1) : T(-1));
        if constexpr (CO >= 2) {
          const T uct = T(0.5) * (usRt + usLt + (usRbt - usLbt) * sBx * sqrho);
          const T ucbt = T(0.5) * (usRbt + usLbt + (usRt - usLt) * sBx / sqrho);
          flux[Xt] = uct * SM - Bx * ucbt;
          flux[BXt] = ucbt * SM - Bx * uct / rho;
        }
        if constexpr (CO >= 3) {
          const T ucb = T(0.5) * (usRb + usLb + (usRbb - usLbb) * sBx * sqrho);
          const T ucbb = T(0.5) * (usRbb + usLbb + (usRb - usLb) * sBx / sqrho);
          flux[Xb] = ucb * SM - Bx * ucbb;
          flux[BXb] = ucbb * SM - Bx * ucb / rho;
        }
      }
    }
  }
  StarPre<T> pre;
  pre.valid = false;
  store_face_emf<C, DIR>(vL, vR, uL, uR, flux, sl, sr, total_pressure_iso<C>(vL, c2),
                         total_pressure_iso<C>(vR, c2), E, emf, pre);
}

// RoeMHD without an energy equation (roeMHD.hpp:175-651, the !HAVE_ENERGY branches
// :21-41, :225-232, :296-302, :325-328, :427-431, :454-461, :515-520, :541-546): 7 waves
// fast-, fast+, div B, slow-, slow+, Alfven-, Alfven+ (no entropy wave), the sound speed
// squared a2 = Cs2Iso + X, fast and slow wave strengths built on a2 dV[RHO].
template <class C, int DIR, typename T>
__device__ __forceinline__ void roe_mhd_iso(const T *vL, const T *vR, T *flux, T &cmax,
                                            FaceEmf<T> &E, int emf, T csiso) {
  constexpr int NV = C::NVAR;
  constexpr int CO = C::COMP;
  constexpr int Xn = C::template Xn<DIR>();
  constexpr int Xt = C::template Xt<DIR>();
  constexpr int Xb = C::template Xb<DIR>();
  constexpr int BXn = C::template BXn<DIR>();
  constexpr int BXt = C::template BXt<DIR>();
  constexpr int BXb = C::template BXb<DIR>();
  constexpr int KFASTM = 0, KFASTP = 1, KDIVB = 2, KSLOWM = 3, KSLOWP = 4, KALFVM = 5,
                KALFVP = 6;
  const T delta = T(1.e-6);
  T dV[NV], uL[NV], uR[NV], fluxL[NV], fluxR[NV], Rc[NV][NV];
#pragma unroll
  for (int n = 0; n < NV; n++) dV[n] = vR[n] - vL[n];
  prim2cons<C>(vL, uL, csiso);
  prim2cons<C>(vR, uR, csiso);
  T a2L = csiso;
  a2L = a2L * a2L;
  const T a2R = a2L;
  kflux<C, DIR>(fluxL, vL, uL, a2L);
  kflux<C, DIR>(fluxR, vR, uR, a2R);
#pragma unroll
  for (int p = 0; p < NV; p++)
#pragma unroll
    for (int q = 0; q < NV; q++) Rc[p][q] = 0;
  const T sqr_rho_L = sqrt_(vL[C::RHO]);
  const T sqr_rho_R = sqrt_(vR[C::RHO]);
  T sl = sqr_rho_L / (sqr_rho_L + sqr_rho_R);
  T sr = sqr_rho_R / (sqr_rho_L + sqr_rho_R);
  const T rho = sr * vL[C::RHO] + sl * vR[C::RHO];
  const T sqrt_rho = sqrt_(rho);
  T u, v = 0, w = 0, Bx, By = 0, Bz = 0, bx, by = 0, bz = 0;
  u = sl * vL[Xn] + sr * vR[Xn];
  if constexpr (CO >= 2) v = sl * vL[Xt] + sr * vR[Xt];
  if constexpr (CO >= 3) w = sl * vL[Xb] + sr * vR[Xb];
  Bx = sr * vL[BXn] + sl * vR[BXn];
  if constexpr (CO >= 2) By = sr * vL[BXt] + sl * vR[BXt];
  if constexpr (CO >= 3) Bz = sr * vL[BXb] + sl * vR[BXb];
  (void)v;
  (void)w;
  const T sBx = (Bx >= T(0) ? T(1) : T(-1));
  bx = Bx / sqrt_rho;
  if constexpr (CO >= 2) by = By / sqrt_rho;
  if constexpr (CO >= 3) bz = Bz / sqrt_rho;
  T bt2 = T(0);
  if constexpr (CO >= 2) bt2 = bt2 + by * by;
  if constexpr (CO >= 3) bt2 = bt2 + bz * bz;
  const T b2 = bx * bx + bt2;
  const T Btmag = sqrt_(bt2 * rho);
  T X = dV[BXn] * dV[BXn];
  if constexpr (CO >= 2) X = X + dV[BXt] * dV[BXt];
  if constexpr (CO >= 3) X = X + dV[BXb] * dV[BXb];
  X /= (sqr_rho_L + sqr_rho_R) * (sqr_rho_L + sqr_rho_R) * T(2);
  const T a2 = T(0.5) * (a2L + a2R) + X;
  T scrh = a2 - b2;
  const T ca2 = bx * bx;
  scrh = scrh * scrh + T(4) * bt2 * a2;
  scrh = sqrt_(scrh);
  const T cf2 = T(0.5) * (a2 + b2 + scrh);
  const T cs2 = a2 * ca2 / cf2;
  const T cf = sqrt_(cf2);
  const T cs = sqrt_(cs2);
  const T ca = sqrt_(ca2);
  const T a = sqrt_(a2);
  (void)ca;
  T alpha_f, alpha_s;
  if (cf == cs) {
    alpha_f = T(1);
    alpha_s = T(0);
  } else if (a <= cs) {
    alpha_f = T(0);
    alpha_s = T(1);
  } else if (cf <= a) {
    alpha_f = T(1);
    alpha_s = T(0);
  } else {
    scrh = T(1) / (cf2 - cs2);
    alpha_f = (a2 - cs2) * scrh;
    alpha_s = (cf2 - a2) * scrh;
    alpha_f = fmx(T(0), alpha_f);
    alpha_s = fmx(T(0), alpha_s);
    alpha_f = sqrt_(alpha_f);
    alpha_s = sqrt_(alpha_s);
  }
  T beta_y = 0, beta_z = 0;
  if (Btmag > T(1.e-9)) {
    if constexpr (CO == 2) beta_y = (By >= T(0) ? T(1) : T(-1));
    if constexpr (CO == 3) {
      beta_y = By / Btmag;
      beta_z = Bz / Btmag;
    }
  } else {
    if constexpr (CO == 2) beta_y = T(1);
    if constexpr (CO == 3) beta_z = beta_y = T(1);
  }
  T lambda[NV], eta[NV];
  // fast wave u - c_f
  lambda[KFASTM] = u - cf;
  scrh = alpha_s * cs * sBx;
  T beta_dv = T(0), beta_dB = T(0);
  if constexpr (CO >= 2) {
    beta_dv = beta_dv + beta_y * dV[Xt];
    beta_dB = beta_dB + beta_y * dV[BXt];
  }
  if constexpr (CO >= 3) {
    beta_dv = beta_dv + beta_z * dV[Xb];
    beta_dB = beta_dB + beta_z * dV[BXb];
  }
  Rc[C::RHO][KFASTM] = alpha_f;
  Rc[Xn][KFASTM] = alpha_f * lambda[KFASTM];
  if constexpr (CO >= 2) Rc[Xt][KFASTM] = alpha_f * v + scrh * beta_y;
  if constexpr (CO >= 3) Rc[Xb][KFASTM] = alpha_f * w + scrh * beta_z;
  if constexpr (CO >= 2) Rc[BXt][KFASTM] = alpha_s * a * beta_y / sqrt_rho;
  if constexpr (CO >= 3) Rc[BXb][KFASTM] = alpha_s * a * beta_z / sqrt_rho;
  eta[KFASTM] = alpha_f * a2 * dV[C::RHO];
  eta[KFASTM] += rho * scrh * beta_dv - rho * alpha_f * cf * dV[Xn] + sqrt_rho * alpha_s * a * beta_dB;
  eta[KFASTM] *= T(0.5) / a2;
  // fast wave u + c_f
  lambda[KFASTP] = u + cf;
  Rc[C::RHO][KFASTP] = alpha_f;
  Rc[Xn][KFASTP] = alpha_f * lambda[KFASTP];
  if constexpr (CO >= 2) Rc[Xt][KFASTP] = alpha_f * v - scrh * beta_y;
  if constexpr (CO >= 3) Rc[Xb][KFASTP] = alpha_f * w - scrh * beta_z;
  if constexpr (CO >= 2) Rc[BXt][KFASTP] = Rc[BXt][KFASTM];
  if constexpr (CO >= 3) Rc[BXb][KFASTP] = Rc[BXb][KFASTM];
  eta[KFASTP] = alpha_f * (T(0) * X + a2) * dV[C::RHO] - rho * scrh * beta_dv +
                rho * alpha_f * cf * dV[Xn] + sqrt_rho * alpha_s * a * beta_dB;
  eta[KFASTP] *= T(0.5) / a2;
  // div B wave u (zero jump)
  lambda[KDIVB] = u;
  eta[KDIVB] = T(0);
  if constexpr (CO > 1) {
    scrh = alpha_f * cf * sBx;
    lambda[KSLOWM] = u - cs;
    Rc[C::RHO][KSLOWM] = alpha_s;
    Rc[Xn][KSLOWM] = alpha_s * lambda[KSLOWM];
    Rc[Xt][KSLOWM] = alpha_s * v - scrh * beta_y;
    if constexpr (CO >= 3) Rc[Xb][KSLOWM] = alpha_s * w - scrh * beta_z;
    Rc[BXt][KSLOWM] = -alpha_f * a * beta_y / sqrt_rho;
    if constexpr (CO >= 3) Rc[BXb][KSLOWM] = -alpha_f * a * beta_z / sqrt_rho;
    eta[KSLOWM] = alpha_s * (T(0) * X + a2) * dV[C::RHO] - rho * scrh * beta_dv -
                  rho * alpha_s * cs * dV[Xn] - sqrt_rho * alpha_f * a * beta_dB;
    eta[KSLOWM] *= T(0.5) / a2;
    lambda[KSLOWP] = u + cs;
    Rc[C::RHO][KSLOWP] = alpha_s;
    Rc[Xn][KSLOWP] = alpha_s * lambda[KSLOWP];
    Rc[Xt][KSLOWP] = alpha_s * v + scrh * beta_y;
    if constexpr (CO >= 3) Rc[Xb][KSLOWP] = alpha_s * w + scrh * beta_z;
    Rc[BXt][KSLOWP] = Rc[BXt][KSLOWM];
    if constexpr (CO >= 3) Rc[BXb][KSLOWP] = Rc[BXb][KSLOWM];
    eta[KSLOWP] = alpha_s * (T(0) * X + a2) * dV[C::RHO] + rho * scrh * beta_dv +
                  rho * alpha_s * cs * dV[Xn] - sqrt_rho * alpha_f * a * beta_dB;
    eta[KSLOWP] *= T(0.5) / a2;
  }
  if constexpr (CO == 3) {
    lambda[KALFVM] = u - ca;
    Rc[Xt][KALFVM] = -rho * beta_z;
    Rc[Xb][KALFVM] = +rho * beta_y;
    Rc[BXt][KALFVM] = -sBx * sqrt_rho * beta_z;
    Rc[BXb][KALFVM] = sBx * sqrt_rho * beta_y;
    eta[KALFVM] = +beta_y * dV[Xb] - beta_z * dV[Xt] +
                  sBx / sqrt_rho * (beta_y * dV[BXb] - beta_z * dV[BXt]);
    eta[KALFVM] *= T(0.5);
    lambda[KALFVP] = u + ca;
    Rc[Xt][KALFVP] = -Rc[Xt][KALFVM];
    Rc[Xb][KALFVP] = -Rc[Xb][KALFVM];
    Rc[BXt][KALFVP] = Rc[BXt][KALFVM];
    Rc[BXb][KALFVP] = Rc[BXb][KALFVM];
    eta[KALFVP] = -beta_y * dV[Xb] + beta_z * dV[Xt] +
                  sBx / sqrt_rho * (beta_y * dV[BXb] - beta_z * dV[BXt]);
    eta[KALFVP] *= T(0.5);
  }
  cmax = fab(u) + cf;
  sl = lambda[KFASTM];
  sr = lambda[KFASTP];
  T alambda[NV];
#pragma unroll
  for (int n = 0; n < NV; n++) alambda[n] = fab(lambda[n]);
  if (alambda[KFASTM] < T(0.5) * delta)
    alambda[KFASTM] = lambda[KFASTM] * lambda[KFASTM] / delta + T(0.25) * delta;
  if (alambda[KFASTP] < T(0.5) * delta)
    alambda[KFASTP] = lambda[KFASTP] * lambda[KFASTP] / delta + T(0.25) * delta;
  if constexpr (CO > 1) {
    if (alambda[KSLOWM] < T(0.5) * delta)
      alambda[KSLOWM] = lambda[KSLOWM] * lambda[KSLOWM] / delta + T(0.25) * delta;
    if (alambda[KSLOWP] < T(0.5) * delta)
      alambda[KSLOWP] = lambda[KSLOWP] * lambda[KSLOWP] / delta + T(0.25) * delta;
  }
#pragma unroll
  for (int n1 = 0; n1 < NV; n1++) {
    T sc = T(0);
#pragma unroll
    for (int n2 = 0; n2 < NV; n2++) sc += alambda[n2] * eta[n2] * Rc[n1][n2];
    flux[n1] = T(0.5) * (fluxL[n1] + fluxR[n1] - sc);
  }
  StarPre<T> pre;
  pre.valid = false;
  store_face_emf<C, DIR>(vL, vR, uL, uR, flux, sl, sr, total_pressure_iso<C>(vL, a2L),
                         total_pressure_iso<C>(vR, a2R), E, emf, pre);
}

// Dispatch to the configured Riemann solver.  Returns false only in the 1-D Roe case
// where the reference leaves the flux unwritten.
template <class C, int DIR, typename T>
__device__ __forceinline__ bool riemann(const T *vL, const T *vR, T *flux, T &cmax,
                                        FaceEmf<T> &E, int emf, T gamma) {
  if constexpr (C::ISO) {
    // gamma carries the isothermal sound speed in isothermal configurations
    if constexpr (C::MHD) {
      if constexpr (C::SOLVER == SOLVER_HLL) hll_mhd_iso<C, DIR>(vL, vR, flux, cmax, E, emf, gamma);
      else if constexpr (C::SOLVER == SOLVER_TVDLF) tvdlf_mhd_iso<C, DIR>(vL, vR, flux, cmax, E, emf, gamma);
      else if constexpr (C::SOLVER == SOLVER_ROE) roe_mhd_iso<C, DIR>(vL, vR, flux, cmax, E, emf, gamma);
      else {
        static_assert(C::SOLVER == SOLVER_HLLD, "isothermal MHD solver not built");
        hlld_mhd_iso<C, DIR>(vL, vR, flux, cmax, E, emf, gamma);
      }
      return true;
    } else if constexpr (C::SOLVER == SOLVER_HLLC) {
      hllc_hd_iso<C, DIR>(vL, vR, flux, cmax, gamma);
      return true;
    } else if constexpr (C::SOLVER == SOLVER_HLL) {
      hll_hd_iso<C, DIR>(vL, vR, flux, cmax, gamma);
      return true;
    } else if constexpr (C::SOLVER == SOLVER_TVDLF) {
      tvdlf_hd_iso<C, DIR>(vL, vR, flux, cmax, gamma);
      return true;
    } else {
      return roe_hd_iso<C, DIR>(vL, vR, flux, cmax, gamma);
    }
  } else if constexpr (C::MHD) {
    if constexpr (C::SOLVER == SOLVER_ROE) {
      roe_mhd<C, DIR>(vL, vR, flux, cmax, E, emf, gamma);
    } else if constexpr (C::SOLVER == SOLVER_HLL) {
      hll_mhd<C, DIR>(vL, vR, flux, cmax, E, emf, gamma);
    } else if constexpr (C::SOLVER == SOLVER_TVDLF) {
      tvdlf_mhd<C, DIR>(vL, vR, flux, cmax, E, emf, gamma);
    } else {
      static_assert(C::SOLVER == SOLVER_HLLD, "MHD solver not built");
      hlld_mhd<C, DIR>(vL, vR, flux, cmax, E, emf, gamma);
    }
    return true;
  } else if constexpr (C::SOLVER == SOLVER_HLLC) {
    hllc_hd<C, DIR>(vL, vR, flux, cmax, gamma);
    return true;
  } else if constexpr (C::SOLVER == SOLVER_HLL) {
    hll_hd<C, DIR>(vL, vR, flux, cmax, gamma);
    return true;
  } else if constexpr (C::SOLVER == SOLVER_TVDLF) {
    tvdlf_hd<C, DIR>(vL, vR, flux, cmax, gamma);
    return true;
  } else {
    return roe_hd<C, DIR>(vL, vR, flux, cmax, gamma);
  }
}

}  // namespace idfx
