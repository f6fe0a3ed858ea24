// CPU oracle for the per-stage finite-volume update.  TEST INFRASTRUCTURE ONLY.
//
// This file restates, on the CPU and in plain loops, the reference's algorithm for the
// hot path: ExtrapolatePrimVar (PLM/PPM), the HLLC/HLLD/Roe Riemann fluxes with the
// face-EMF stores, CalcRightHandSide, ConstrainedTransport's CalcCornerEMF and
// EvolveMagField, ConvertConsToPrim/PrimToCons, the CFL dt reduction, the periodic /
// outflow / reflective boundaries and the RK driver.  It mirrors the reference's kernel
// sequence one loop per reference kernel, so that every intermediate array can be
// compared with the GPU path.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, as the results the GPU path is compared with.
//
// Runtime configuration replaces the reference's compile-time macros: `dim` is
// DIMENSIONS, `comps` is COMPONENTS, `order` is ORDER.  A term that the reference
// removes through EXPAND()/D_EXPAND() is evaluated here as an exact zero, which leaves
// every sum bit-identical (x + 0.0 == x).
//
// Compile with -ffp-contract=off (the reference's own CPU/GPU parity setting,
// pytools/idfx_test.py:132-135).
//
// Pinning status (checks that pin each function independently of this restatement;
// the tests that run them are named in tests/; test_oracle_pins.py also runs each of its
// checks against a planted mistake, orc_plant below, and requires it to fail):
//   PLM van Leer / MC limiters ........ closed forms + brute-force monotonicity (test_oracle_pointwise)
//   PPM getPPMStates .................. exact on cubics, on quadratics with an interior extremum,
//                                        fourth order at smooth extrema, no new extrema (pointwise, pins)
//   K_PrimToCons / K_ConsToPrim ........ round trip, pressure floor branch (test_oracle_pointwise)
//   K_Flux, HLLC, HLL, TVDLF, Roe HD ... consistency, upwind limits, jump conditions, Rusanov form
//   HLLD (double-star included) ........ exact moving contacts with B_n != 0 and rotational
//                                        discontinuities (test_oracle_pins)
//   UCT-HLL / UCT-HLLD stores .......... coefficient identity with the solver's flux (pins)
//   CalcCornerEMF ...................... uct_contact/uct0 reduce to 1-D, field-loop energy,
//                                        Londrillo-Del Zanna edge formula (pins); div B, conservation
//   HLLC + PLM + RK2 + outflow ......... Sod exact solution (Toro table 4.2 star state) (test_oracle_sod)
//   HLLD + CT (all EMF schemes) ........ div B at round-off, conservation, oblique linear fast
//                                        wave convergence (test_oracle_mhd)
//   Roe MHD + CT + RK2/RK3 in 3-D ...... the reference's four-mode linear-wave test (pins)
//   dt reduction ....................... brute-force min (test_oracle_pointwise)
#include "oracle.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int IDIR = 0, JDIR = 1, KDIR = 2;

// Planted mistake (negative controls of the tests, never set in normal use): the tests
// switch one of these on to show that a check fails when the restated branch is wrong.
//   ORC_PLANT_HLLD_SBX      sign of B_n flipped in the HLLD double-star state
//   ORC_PLANT_CONTACT_SV    upwind switch of uct_contact inverted (sv -> 1 - sv)
//   ORC_PLANT_UCT0_INDEX    uct0's cell-EMF average of the x faces taken along y
//   ORC_PLANT_RIEMANN_PHI   sign of the dissipative term of UCT-HLL/HLLD reversed
//   ORC_PLANT_PPM_EXTREMUM  PPM clips every extremum (smooth-extremum test removed)
//   ORC_PLANT_ROE_ALFVEN    Roe MHD: Alfven-wave strengths of both families swapped
int g_plant = 0;

// Variable indices of the reference (src/idefix.hpp:74-108), with -1 for the
// components that COMPONENTS removes.
struct Vars {
  int RHO, MX1, MX2, MX3, ENG, BX1, BX2, BX3, nvar;
  explicit Vars(const orc_cfg &c) {
    RHO = 0;
    MX1 = 1;
    MX2 = c.comps >= 2 ? 2 : -1;
    MX3 = c.comps == 3 ? 3 : -1;
    // ISOTHERMAL builds have no energy variable (idefix.hpp:63-92)
    const int he = c.isothermal ? 0 : 1;
    ENG = he ? c.comps + 1 : -1;
    if (c.mhd) {
      BX1 = c.comps + 1 + he;
      BX2 = c.comps >= 2 ? BX1 + 1 : -1;
      BX3 = c.comps == 3 ? BX1 + 2 : -1;
      nvar = 2 * c.comps + 1 + he;
    } else {
      BX1 = BX2 = BX3 = -1;
      nvar = c.comps + 1 + he;
    }
  }
};

inline double at(const double *a, int idx) { return idx >= 0 ? a[idx] : 0.0; }
// Cs2Iso of a face: the square of the average of the sound speeds of its two cells
// (hlldMHD.hpp:188-191); constant here (eos_isothermal.hpp, "csiso constant").
inline double c2iso(const orc_cfg &c) {
  const double cs = 0.5 * (c.csiso + c.csiso);
  return cs * cs;
}
inline void put(double *a, int idx, double v) { if (idx >= 0) a[idx] = v; }
// FMIN/FMAX/FABS are fmin/fmax/fabs in double precision (real_types.hpp:49-51)
inline double FMIN(double a, double b) { return std::fmin(a, b); }
inline double FMAX(double a, double b) { return std::fmax(a, b); }

// --------------------------------------------------------------------------------------
// Slope limiters: src/fluid/RiemannSolver/slopeLimiter.hpp
// --------------------------------------------------------------------------------------
// VanLeerLim, slopeLimiter.hpp:54-57
inline double vanleer(double dvp, double dvm) {
  return (dvp * dvm > 0.0 ? 2.0 * dvp * dvm / (dvp + dvm) : 0.0);
}
// LimO3Lim, slopeLimiter.hpp:25-51 (Cada & Torrilhon 2009 third-order limiter, r = 0.1):
// the limited slope of a cell is dvp * LimO3Lim(dvp, dvm, dx).
inline double limo3lim(double dvp, double dvm, double dx) {
  const double r = 0.1, eps = 1.e-12;
  const double th = dvm / (dvp + 1.e-16);
  const double q = (2.0 + th) / 3.0;
  double a = FMIN(1.5, 2.0 * th);
  a = FMIN(q, a);
  const double b = FMAX(-0.5 * th, a);
  const double cc = FMIN(q, b);
  double psi = FMAX(0.0, cc);
  double eta = r * dx;
  eta = (dvm * dvm + dvp * dvp) / (eta * eta);
  double lim;
  if (eta <= 1.0 - eps) {
    lim = q;
  } else if (eta >= 1.0 + eps) {
    lim = psi;
  } else {
    psi = (1.0 - (eta - 1.0) / eps) * q + (1.0 + (eta - 1.0) / eps) * psi;
    lim = 0.5 * psi;
  }
  return lim;
}
// MinModLim, slopeLimiter.hpp:16-23, as written: the limited value is assigned to a
// variable declared inside the if-block, so the function always returns 0 and the
// positivity fallback of LimO3 is first order.
inline double minmod_as_built(double, double) { return 0.0; }
// McLim, slopeLimiter.hpp:67-75
inline double mclim(double dvp, double dvm) {
  double dq = 0;
  if (dvp * dvm > 0.0) {
    double dqc = 0.5 * (dvp + dvm);
    double d2q = 2.0 * (std::fabs(dvp) < std::fabs(dvm) ? dvp : dvm);
    dq = std::fabs(d2q) < std::fabs(dqc) ? d2q : dqc;
  }
  return dq;
}
inline int sgn(double v) { return (0.0 < v) - (v < 0.0); }
// limitPPMFaceValues, slopeLimiter.hpp:121-140
inline void limit_ppm_face(double vm1, double v0, double vp1, double vp2, double &vph) {
  if ((vp1 - vph) * (vph - v0) < 0.0) {
    const double deltaL = (vm1 - 2 * v0 + vp1);
    const double deltaC = 3 * (v0 - 2 * vph + vp1);
    const double deltaR = (v0 - 2 * vp1 + vp2);
    double delta = 0.0;
    if (sgn(deltaL) == sgn(deltaC) && sgn(deltaR) == sgn(deltaC)) {
      const double C = 1.25;
      delta = C * FMIN(std::fabs(deltaL), std::fabs(deltaR));
      delta = sgn(deltaC) * FMIN(delta, std::fabs(deltaC));
    }
    vph = 0.5 * (v0 + vp1) - delta / 6.0;
  }
}
// getPPMStates, slopeLimiter.hpp:142-191
inline void ppm_states(double vm2, double vm1, double v0, double vp1, double vp2,
                       double &vl, double &vr) {
  const int n = 2;
  vr = 7.0 / 12.0 * (v0 + vp1) - 1.0 / 12.0 * (vm1 + vp2);
  vl = 7.0 / 12.0 * (vm1 + v0) - 1.0 / 12.0 * (vm2 + vp1);
  limit_ppm_face(vm2, vm1, v0, vp1, vl);
  limit_ppm_face(vm1, v0, vp1, vp2, vr);
  double d2qf = 6.0 * (vl + vr - 2.0 * v0);
  double d2qc0 = vm1 + vp1 - 2.0 * v0;
  double d2qcp1 = v0 + vp2 - 2.0 * vp1;
  double d2qcm1 = vm2 + v0 - 2.0 * vm1;
  double d2q = 0.0;
  if (sgn(d2qf) == sgn(d2qc0) && sgn(d2qf) == sgn(d2qcp1) && sgn(d2qf) == sgn(d2qcm1)) {
    const double C = 1.25;
    d2q = FMIN(std::fabs(d2qc0), std::fabs(d2qcp1));
    d2q = C * FMIN(std::fabs(d2qcm1), d2q);
    d2q = sgn(d2qf) * FMIN(std::fabs(d2qf), d2q);
  }
  double qmax = FMAX(FMAX(std::fabs(vm1), std::fabs(v0)), std::fabs(vp1));
  double rho = 0.0;
  if (std::fabs(d2qf) > 1e-12 * qmax) rho = d2q / d2qf;
  if (g_plant == ORC_PLANT_PPM_EXTREMUM) rho = 0.0;
  if (((vr - v0) * (v0 - vl) <= 0) || (vm1 - v0) * (v0 - vp1) <= 0) {
    if (rho <= (1.0 - 1e-12)) {
      vl = v0 - rho * (v0 - vl);
      vr = v0 + rho * (vr - v0);
    }
  } else {
    if (std::fabs(vr - v0) >= n * std::fabs(v0 - vl)) vr = v0 + n * (v0 - vl);
    if (std::fabs(vl - v0) >= n * std::fabs(v0 - vr)) vl = v0 + n * (v0 - vr);
  }
}

// --------------------------------------------------------------------------------------
// Pointwise conversions: src/fluid/convertConsToPrim.hpp, eos_adiabatic.hpp:39-45
// --------------------------------------------------------------------------------------
// K_PrimToCons, convertConsToPrim.hpp:69-100
inline void prim2cons(const orc_cfg &c, const Vars &V, const double *v, double *u) {
  u[V.RHO] = v[V.RHO];
  u[V.MX1] = v[V.MX1] * v[V.RHO];
  if (c.comps >= 2) u[V.MX2] = v[V.MX2] * v[V.RHO];
  if (c.comps >= 3) u[V.MX3] = v[V.MX3] * v[V.RHO];
  double v2 = v[V.MX1] * v[V.MX1];
  if (c.comps >= 2) v2 = v2 + v[V.MX2] * v[V.MX2];
  if (c.comps >= 3) v2 = v2 + v[V.MX3] * v[V.MX3];
  if (c.isothermal) {
    if (c.mhd) {
      u[V.BX1] = v[V.BX1];
      if (c.comps >= 2) u[V.BX2] = v[V.BX2];
      if (c.comps >= 3) u[V.BX3] = v[V.BX3];
    }
    return;
  }
  const double eint = v[V.ENG] / (c.gamma - 1.0);
  if (c.mhd) {
    u[V.BX1] = v[V.BX1];
    if (c.comps >= 2) u[V.BX2] = v[V.BX2];
    if (c.comps >= 3) u[V.BX3] = v[V.BX3];
    double b2 = u[V.BX1] * u[V.BX1];
    if (c.comps >= 2) b2 = b2 + u[V.BX2] * u[V.BX2];
    if (c.comps >= 3) b2 = b2 + u[V.BX3] * u[V.BX3];
    u[V.ENG] = eint + 0.5 * v[V.RHO] * v2 + 0.5 * b2;
  } else {
    u[V.ENG] = eint + 0.5 * v[V.RHO] * v2;
  }
}

// K_ConsToPrim, convertConsToPrim.hpp:15-67 (SMALL_PRESSURE_FIX = 1e-5,
// fluid_defs.hpp:23).  The re-synced energy stays local: the caller writes only V.
inline void cons2prim(const orc_cfg &c, const Vars &V, const double *u_in, double *v) {
  double u[8];
  for (int n = 0; n < V.nvar; n++) u[n] = u_in[n];
  v[V.RHO] = u[V.RHO];
  v[V.MX1] = u[V.MX1] / u[V.RHO];
  if (c.comps >= 2) v[V.MX2] = u[V.MX2] / u[V.RHO];
  if (c.comps >= 3) v[V.MX3] = u[V.MX3] / u[V.RHO];
  double m2 = u[V.MX1] * u[V.MX1];
  if (c.comps >= 2) m2 = m2 + u[V.MX2] * u[V.MX2];
  if (c.comps >= 3) m2 = m2 + u[V.MX3] * u[V.MX3];
  if (c.isothermal) {
    if (c.mhd) {
      v[V.BX1] = u[V.BX1];
      if (c.comps >= 2) v[V.BX2] = u[V.BX2];
      if (c.comps >= 3) v[V.BX3] = u[V.BX3];
    }
    return;
  }
  const double kin = 0.5 / u[V.RHO] * m2;
  if (c.mhd) {
    v[V.BX1] = u[V.BX1];
    if (c.comps >= 2) v[V.BX2] = u[V.BX2];
    if (c.comps >= 3) v[V.BX3] = u[V.BX3];
    double b2 = u[V.BX1] * u[V.BX1];
    if (c.comps >= 2) b2 = b2 + u[V.BX2] * u[V.BX2];
    if (c.comps >= 3) b2 = b2 + u[V.BX3] * u[V.BX3];
    const double mag = 0.5 * b2;
    v[V.ENG] = (u[V.ENG] - kin - mag) * (c.gamma - 1.0);
    if (v[V.ENG] <= 0.0) v[V.ENG] = 1.0e-5;
  } else {
    v[V.ENG] = (u[V.ENG] - kin) * (c.gamma - 1.0);
    if (v[V.ENG] <= 0.0) v[V.ENG] = 1.0e-5;
  }
}

// --------------------------------------------------------------------------------------
// K_Flux<Phys,DIR>: src/fluid/RiemannSolver/flux.hpp:28-95
// --------------------------------------------------------------------------------------
struct Dir {
  int Xn, Xt, Xb, BXn, BXt, BXb;
  Dir(const Vars &V, int dir) {
    Xn = V.MX1 + dir;
    Xt = (dir == IDIR ? V.MX2 : V.MX1);
    Xb = (dir == KDIR ? V.MX2 : V.MX3);
    if (V.BX1 >= 0) {
      BXn = V.BX1 + dir;
      BXt = (dir == IDIR ? V.BX2 : V.BX1);
      BXb = (dir == KDIR ? V.BX2 : V.BX3);
    } else {
      BXn = BXt = BXb = -1;
    }
  }
};

// c2 is Cs2Iso, the squared isothermal sound speed (used only when isothermal)
inline void kflux(const orc_cfg &c, const Vars &V, const Dir &D, double *F, const double *v,
                  const double *u, double c2 = 0.0) {
  F[V.RHO] = u[D.Xn];
  if (c.mhd) {
    F[V.MX1] = u[V.MX1] * v[D.Xn] - v[D.BXn] * v[V.BX1];
    if (c.comps >= 2) F[V.MX2] = u[V.MX2] * v[D.Xn] - v[D.BXn] * v[V.BX2];
    if (c.comps >= 3) F[V.MX3] = u[V.MX3] * v[D.Xn] - v[D.BXn] * v[V.BX3];
    F[D.BXn] = 0.0;
    if (c.comps >= 2) F[D.BXt] = v[D.Xn] * v[D.BXt] - v[D.BXn] * v[D.Xt];
    if (c.comps >= 3) F[D.BXb] = v[D.Xn] * v[D.BXb] - v[D.BXn] * v[D.Xb];
    double b2 = v[V.BX1] * v[V.BX1];
    if (c.comps >= 2) b2 = b2 + v[V.BX2] * v[V.BX2];
    if (c.comps >= 3) b2 = b2 + v[V.BX3] * v[V.BX3];
    if (c.isothermal) {
      F[D.Xn] += c2 * v[V.RHO] + 0.5 * b2;
      return;
    }
    const double ptot = v[V.ENG] + 0.5 * b2;
    double vb = v[V.MX1] * v[V.BX1];
    if (c.comps >= 2) vb = vb + v[V.MX2] * v[V.BX2];
    if (c.comps >= 3) vb = vb + v[V.MX3] * v[V.BX3];
    F[V.ENG] = (u[V.ENG] + ptot) * v[D.Xn] - v[D.BXn] * vb;
    F[D.Xn] += ptot;
  } else {
    F[V.MX1] = u[V.MX1] * v[D.Xn];
    if (c.comps >= 2) F[V.MX2] = u[V.MX2] * v[D.Xn];
    if (c.comps >= 3) F[V.MX3] = u[V.MX3] * v[D.Xn];
    if (c.isothermal) {
      F[D.Xn] += c2 * v[V.RHO];
      return;
    }
    const double ptot = v[V.ENG];
    F[V.ENG] = (u[V.ENG] + ptot) * v[D.Xn];
    F[D.Xn] += ptot;
  }
}

// Face EMF data written by the K_Store* functions (storeFlux.hpp).
struct FaceEmf {
  double Et, Eb, sv, aL, aR, dL, dR;
  bool hasEt, hasEb, hasSv, hasAD;
};

// --------------------------------------------------------------------------------------
// HllcHD: src/fluid/RiemannSolver/HDsolvers/hllcHD.hpp:35-163 (HAVE_ENERGY branch)
// --------------------------------------------------------------------------------------
void hllc_hd(const orc_cfg &c, const Vars &V, const Dir &D, const double *vL, const double *vR,
             double *flux, double &cmax) {
  const int nv_ = V.nvar;
  double uL[8], uR[8], fL[8], fR[8];
  // isothermal: the face average of the sound speed at the two cells (hllcHD.hpp:67-70)
  const double cL = c.isothermal ? 0.5 * (c.csiso + c.csiso)
                                 : std::sqrt(c.gamma * (vL[V.ENG] / vL[V.RHO]));
  const double cR = c.isothermal ? cL : std::sqrt(c.gamma * (vR[V.ENG] / vR[V.RHO]));
  const double cminL = vL[D.Xn] - cL, cmaxL = vL[D.Xn] + cL;
  const double cminR = vR[D.Xn] - cR, cmaxR = vR[D.Xn] + cR;
  const double SL = FMIN(cminL, cminR);
  const double SR = FMAX(cmaxL, cmaxR);
  cmax = FMAX(std::fabs(SL), std::fabs(SR));
  prim2cons(c, V, vL, uL);
  prim2cons(c, V, vR, uR);
  kflux(c, V, D, fL, vL, uL, cL * cL);
  kflux(c, V, D, fR, vR, uR, cR * cR);
  if (SL > 0) {
    for (int n = 0; n < nv_; n++) flux[n] = fL[n];
  } else if (SR < 0) {
    for (int n = 0; n < nv_; n++) flux[n] = fR[n];
  } else if (c.isothermal) {
    // hllcHD.hpp:129-143
    double usL[8], usR[8];
    const double scrh = 1.0 / (SR - SL);
    const double rho = (SR * uR[V.RHO] - SL * uL[V.RHO] - fR[V.RHO] + fL[V.RHO]) * scrh;
    const double mx = (SR * uR[D.Xn] - SL * uL[D.Xn] - fR[D.Xn] + fL[D.Xn]) * scrh;
    usL[V.RHO] = usR[V.RHO] = rho;
    usL[D.Xn] = usR[D.Xn] = mx;
    double vs = (SR * fL[V.RHO] - SL * fR[V.RHO] + SR * SL * (uR[V.RHO] - uL[V.RHO]));
    vs *= scrh;
    vs /= rho;
    if (c.comps >= 2) {
      usL[D.Xt] = rho * vL[D.Xt];
      usR[D.Xt] = rho * vR[D.Xt];
    }
    if (c.comps >= 3) {
      usL[D.Xb] = rho * vL[D.Xb];
      usR[D.Xb] = rho * vR[D.Xb];
    }
    if (vs >= 0.0) {
      for (int n = 0; n < nv_; n++) flux[n] = fL[n] + SL * (usL[n] - uL[n]);
    } else {
      for (int n = 0; n < nv_; n++) flux[n] = fR[n] + SR * (usR[n] - uR[n]);
    }
  } else {
    double usL[8], usR[8];
    const double qL = vL[V.ENG] + uL[D.Xn] * (vL[D.Xn] - SL);
    const double qR = vR[V.ENG] + uR[D.Xn] * (vR[D.Xn] - SR);
    const double wL = vL[V.RHO] * (vL[D.Xn] - SL);
    const double wR = vR[V.RHO] * (vR[D.Xn] - SR);
    const double vs = (qR - qL) / (wR - wL);
    usL[V.RHO] = uL[V.RHO] * (SL - vL[D.Xn]) / (SL - vs);
    usR[V.RHO] = uR[V.RHO] * (SR - vR[D.Xn]) / (SR - vs);
    usL[D.Xn] = usL[V.RHO] * vs;
    usR[D.Xn] = usR[V.RHO] * vs;
    if (c.comps >= 2) {
      usL[D.Xt] = usL[V.RHO] * vL[D.Xt];
      usR[D.Xt] = usR[V.RHO] * vR[D.Xt];
    }
    if (c.comps >= 3) {
      usL[D.Xb] = usL[V.RHO] * vL[D.Xb];
      usR[D.Xb] = usR[V.RHO] * vR[D.Xb];
    }
    usL[V.ENG] = uL[V.ENG] / vL[V.RHO] +
                 (vs - vL[D.Xn]) * (vs + vL[V.ENG] / (vL[V.RHO] * (SL - vL[D.Xn])));
    usR[V.ENG] = uR[V.ENG] / vR[V.RHO] +
                 (vs - vR[D.Xn]) * (vs + vR[V.ENG] / (vR[V.RHO] * (SR - vR[D.Xn])));
    usL[V.ENG] *= usL[V.RHO];
    usR[V.ENG] *= usR[V.RHO];
    if (vs >= 0.0) {
      for (int n = 0; n < nv_; n++) flux[n] = fL[n] + SL * (usL[n] - uL[n]);
    } else {
      for (int n = 0; n < nv_; n++) flux[n] = fR[n] + SR * (usR[n] - uR[n]);
    }
  }
}

// --------------------------------------------------------------------------------------
// HllHD: src/fluid/RiemannSolver/HDsolvers/hllHD.hpp:35-120 (HAVE_ENERGY branch)
// --------------------------------------------------------------------------------------
void hll_hd(const orc_cfg &c, const Vars &V, const Dir &D, const double *vL, const double *vR,
            double *flux, double &cmax) {
  const int nv_ = V.nvar;
  double uL[8], uR[8], fL[8], fR[8];
  const double cL = c.isothermal ? 0.5 * (c.csiso + c.csiso)
                                 : std::sqrt(c.gamma * (vL[V.ENG] / vL[V.RHO]));
  const double cR = c.isothermal ? cL : std::sqrt(c.gamma * (vR[V.ENG] / vR[V.RHO]));
  const double cminL = vL[D.Xn] - cL, cmaxL = vL[D.Xn] + cL;
  const double cminR = vR[D.Xn] - cR, cmaxR = vR[D.Xn] + cR;
  const double SL = FMIN(cminL, cminR);
  const double SR = FMAX(cmaxL, cmaxR);
  cmax = FMAX(std::fabs(SL), std::fabs(SR));
  prim2cons(c, V, vL, uL);
  prim2cons(c, V, vR, uR);
  kflux(c, V, D, fL, vL, uL, cL * cL);
  kflux(c, V, D, fR, vR, uR, cR * cR);
  if (SL > 0) {
    for (int n = 0; n < nv_; n++) flux[n] = fL[n];
  } else if (SR < 0) {
    for (int n = 0; n < nv_; n++) flux[n] = fR[n];
  } else {
    for (int n = 0; n < nv_; n++) {
      flux[n] = SL * SR * uR[n] - SL * SR * uL[n] + SR * fL[n] - SL * fR[n];
      flux[n] /= (SR - SL);
    }
  }
}

// --------------------------------------------------------------------------------------
// TvdlfHD: src/fluid/RiemannSolver/HDsolvers/tvdlfHD.hpp:37-96 (HAVE_ENERGY branch):
// Rusanov flux with the sound speed of the arithmetic mean state.
// --------------------------------------------------------------------------------------
void tvdlf_hd(const orc_cfg &c, const Vars &V, const Dir &D, const double *vL,
              const double *vR, double *flux, double &cmax) {
  const int nv_ = V.nvar;
  double v[8], uL[8], uR[8], fL[8], fR[8];
  for (int n = 0; n < nv_; n++) v[n] = 0.5 * (vL[n] + vR[n]);
  const double cRL = c.isothermal ? 0.5 * (c.csiso + c.csiso)
                                  : std::sqrt(c.gamma * (v[V.ENG] / v[V.RHO]));
  cmax = FMAX(std::fabs(v[D.Xn] + cRL), std::fabs(v[D.Xn] - cRL));
  prim2cons(c, V, vL, uL);
  prim2cons(c, V, vR, uR);
  kflux(c, V, D, fL, vL, uL, cRL * cRL);
  kflux(c, V, D, fR, vR, uR, cRL * cRL);
  for (int n = 0; n < nv_; n++) flux[n] = 0.5 * (fL[n] + fR[n] - cmax * (uR[n] - uL[n]));
}

// --------------------------------------------------------------------------------------
// RoeHD: src/fluid/RiemannSolver/HDsolvers/roeHD.hpp:55-364 (HAVE_ENERGY,
// ROE_AVERAGE == 0).  Returns false when the reference leaves the flux unwritten
// (1-D strong shock, roeHD.hpp:315-328 under DIMENSIONS > 1 only).
// --------------------------------------------------------------------------------------
// RoeHD isothermal branches (roeHD.hpp:29-36, :95-99, :164-167, :186-194, :254-264):
// modes u -+ c and the transverse velocities, a^2 the squared sound speed.
bool roe_hd_iso(const orc_cfg &c, const Vars &V, const Dir &D, const double *vL,
                const double *vR, double *flux, double &cmax) {
  const int nv_ = V.nvar;
  double dv[8], uL[8], uR[8], fL[8], fR[8], Rc[8][8], um[8];
  for (int n = 0; n < nv_; n++) dv[n] = vR[n] - vL[n];
  double a2L = 0.5 * (c.csiso + c.csiso);
  double a2R = a2L;
  a2L = a2L * a2L;
  a2R = a2R * a2R;
  prim2cons(c, V, vL, uL);
  prim2cons(c, V, vR, uR);
  kflux(c, V, D, fL, vL, uL, a2L);
  kflux(c, V, D, fR, vR, uR, a2R);
  for (int n = 0; n < nv_; n++) um[n] = 0.5 * (vR[n] + vL[n]);
  const double a2 = 0.5 * (a2L + a2R);
  const double a = std::sqrt(a2);
  const int I0 = 0, I1 = 1, I2 = 2, I3 = 3;
  double lambda[8], alambda[8], eta[8];
  for (int a1 = 0; a1 < nv_; a1++)
    for (int a3 = 0; a3 < nv_; a3++) Rc[a1][a3] = 0;
  for (int n = 0; n < 8; n++) { lambda[n] = 0; eta[n] = 0; }
  lambda[I0] = um[D.Xn] - a;
  eta[I0] = 0.5 * (dv[V.RHO] - um[V.RHO] * dv[D.Xn] / a);
  Rc[V.RHO][I0] = 1.0;
  Rc[D.Xn][I0] = um[D.Xn] - a;
  if (c.comps >= 2) Rc[D.Xt][I0] = um[D.Xt];
  if (c.comps >= 3) Rc[D.Xb][I0] = um[D.Xb];
  lambda[I1] = um[D.Xn] + a;
  eta[I1] = 0.5 * (dv[V.RHO] + um[V.RHO] * dv[D.Xn] / a);
  Rc[V.RHO][I1] = 1.0;
  Rc[D.Xn][I1] = um[D.Xn] + a;
  if (c.comps >= 2) Rc[D.Xt][I1] = um[D.Xt];
  if (c.comps >= 3) Rc[D.Xb][I1] = um[D.Xb];
  if (c.comps > 1) {
    lambda[I2] = um[D.Xn];
    eta[I2] = um[V.RHO] * dv[D.Xt];
    Rc[D.Xt][I2] = 1.0;
  }
  if (c.comps > 2) {
    lambda[I3] = um[D.Xn];
    eta[I3] = um[V.RHO] * dv[D.Xb];
    Rc[D.Xb][I3] = 1.0;
  }
  cmax = std::fabs(um[D.Xn]) + a;
  double scrh = std::fabs(vL[V.RHO] - vR[V.RHO]);
  scrh /= FMIN(vL[V.RHO], vR[V.RHO]);
  scrh *= a * a;
  if (scrh > 0.5 && (vR[D.Xn] < vL[D.Xn])) {
    if (c.dim > 1) {
      const double bmin = FMIN(0.0, lambda[0]);
      const double bmax = FMAX(0.0, lambda[1]);
      const double scrh1 = 1.0 / (bmax - bmin);
      for (int n = 0; n < nv_; n++) {
        flux[n] = bmin * bmax * (uR[n] - uL[n]) + bmax * fL[n] - bmin * fR[n];
        flux[n] *= scrh1;
      }
      return true;
    }
    return false;
  }
  for (int n = 0; n < nv_; n++) alambda[n] = std::fabs(lambda[n]);
  const double delta = 1.e-7;
  if (alambda[0] <= delta) alambda[0] = 0.5 * lambda[0] * lambda[0] / delta + 0.5 * delta;
  if (alambda[1] <= delta) alambda[1] = 0.5 * lambda[1] * lambda[1] / delta + 0.5 * delta;
  for (int n = 0; n < nv_; n++) {
    double f = fL[n] + fR[n];
    for (int n2 = 0; n2 < nv_; n2++) f -= alambda[n2] * eta[n2] * Rc[n][n2];
    flux[n] = f * 0.5;
  }
  return true;
}

bool roe_hd(const orc_cfg &c, const Vars &V, const Dir &D, const double *vL, const double *vR,
            double *flux, double &cmax) {
  if (c.isothermal) return roe_hd_iso(c, V, D, vL, vR, flux, cmax);
  const int nv_ = V.nvar;
  double dv[8], uL[8], uR[8], fL[8], fR[8], Rc[8][8], um[8];
  for (int n = 0; n < nv_; n++) dv[n] = vR[n] - vL[n];
  double a2L = std::sqrt(c.gamma * (vL[V.ENG] / vL[V.RHO]));
  double a2R = std::sqrt(c.gamma * (vR[V.ENG] / vR[V.RHO]));
  a2L = a2L * a2L;
  a2R = a2R * a2R;
  prim2cons(c, V, vL, uL);
  prim2cons(c, V, vR, uR);
  kflux(c, V, D, fL, vL, uL);
  kflux(c, V, D, fR, vR, uR);
  const double gamma = c.gamma;
  const double gamma_m1 = gamma - 1;
  for (int n = 0; n < nv_; n++) um[n] = 0.5 * (vR[n] + vL[n]);
  const double a2 = gamma * um[V.ENG] / um[V.RHO];
  const double a = std::sqrt(a2);
  double vel2 = um[V.MX1] * um[V.MX1];
  if (c.comps >= 2) vel2 = vel2 + um[V.MX2] * um[V.MX2];
  if (c.comps >= 3) vel2 = vel2 + um[V.MX3] * um[V.MX3];
  const double h = 0.5 * vel2 + a2 / gamma_m1;

  // mode numbering I0, I1, IE, I2, I3 = 0, 1, 2, 3, 4 (roeHD.hpp:20-27)
  const int I0 = 0, I1 = 1, IE = 2, I2 = 3, I3 = 4;
  double lambda[8], alambda[8], eta[8];
  for (int a1 = 0; a1 < nv_; a1++)
    for (int a3 = 0; a3 < nv_; a3++) Rc[a1][a3] = 0;
  for (int n = 0; n < 8; n++) { lambda[n] = 0; eta[n] = 0; }

  lambda[I0] = um[D.Xn] - a;
  eta[I0] = 0.5 / a2 * (dv[V.ENG] - dv[D.Xn] * um[V.RHO] * a);
  Rc[V.RHO][I0] = 1.0;
  Rc[D.Xn][I0] = um[D.Xn] - a;
  if (c.comps >= 2) Rc[D.Xt][I0] = um[D.Xt];
  if (c.comps >= 3) Rc[D.Xb][I0] = um[D.Xb];
  Rc[V.ENG][I0] = h - um[D.Xn] * a;

  lambda[I1] = um[D.Xn] + a;
  eta[I1] = 0.5 / a2 * (dv[V.ENG] + dv[D.Xn] * um[V.RHO] * a);
  Rc[V.RHO][I1] = 1.0;
  Rc[D.Xn][I1] = um[D.Xn] + a;
  if (c.comps >= 2) Rc[D.Xt][I1] = um[D.Xt];
  if (c.comps >= 3) Rc[D.Xb][I1] = um[D.Xb];
  Rc[V.ENG][I1] = h + um[D.Xn] * a;

  lambda[IE] = um[D.Xn];
  eta[IE] = dv[V.RHO] - dv[V.ENG] / a2;
  Rc[V.RHO][IE] = 1.0;
  Rc[V.MX1][IE] = um[V.MX1];
  if (c.comps >= 2) Rc[V.MX2][IE] = um[V.MX2];
  if (c.comps >= 3) Rc[V.MX3][IE] = um[V.MX3];
  Rc[V.ENG][IE] = 0.5 * vel2;

  if (c.comps > 1) {
    lambda[I2] = um[D.Xn];
    eta[I2] = um[V.RHO] * dv[D.Xt];
    Rc[D.Xt][I2] = 1.0;
    Rc[V.ENG][I2] = um[D.Xt];
  }
  if (c.comps > 2) {
    lambda[I3] = um[D.Xn];
    eta[I3] = um[V.RHO] * dv[D.Xb];
    Rc[D.Xb][I3] = 1.0;
    Rc[V.ENG][I3] = um[D.Xb];
  }
  cmax = std::fabs(um[D.Xn]) + a;

  double scrh = std::fabs(vL[V.ENG] - vR[V.ENG]);
  scrh /= FMIN(vL[V.ENG], vR[V.ENG]);
  if (scrh > 0.5 && (vR[D.Xn] < vL[D.Xn])) {
    if (c.dim > 1) {
      const double bmin = FMIN(0.0, lambda[0]);
      const double bmax = FMAX(0.0, lambda[1]);
      const double scrh1 = 1.0 / (bmax - bmin);
      for (int n = 0; n < nv_; n++) {
        flux[n] = bmin * bmax * (uR[n] - uL[n]) + bmax * fL[n] - bmin * fR[n];
        flux[n] *= scrh1;
      }
      return true;
    }
    return false;
  }
  for (int n = 0; n < nv_; n++) alambda[n] = std::fabs(lambda[n]);
  const double delta = 1.e-7;
  if (alambda[0] <= delta) alambda[0] = 0.5 * lambda[0] * lambda[0] / delta + 0.5 * delta;
  if (alambda[1] <= delta) alambda[1] = 0.5 * lambda[1] * lambda[1] / delta + 0.5 * delta;
  for (int n = 0; n < nv_; n++) {
    double f = fL[n] + fR[n];
    for (int n2 = 0; n2 < nv_; n2++) f -= alambda[n2] * eta[n2] * Rc[n][n2];
    flux[n] = f * 0.5;
  }
  return true;
}

void store_face_emf(const orc_cfg &c, const Vars &V, const Dir &D, int dir, const double *vL,
                    const double *vR, const double *uL, const double *uR, const double *flux,
                    double sl, double sr, double ptL, double ptR, FaceEmf &E);

// --------------------------------------------------------------------------------------
// HlldMHD: src/fluid/RiemannSolver/MHDsolvers/hlldMHD.hpp:147-589 (HAVE_ENERGY branch),
// with the face-EMF stores of storeFlux.hpp:14-208.
// --------------------------------------------------------------------------------------
// HlldMHD isothermal branch (hlldMHD.hpp:188-213, :459-571; Mignone 2007): the outer
// waves of the fast speeds with gpr = Cs2Iso rho, one averaged density, the rotational
// discontinuities S1L, S1R, and a reversion to HLL near degeneracy.
void hlld_mhd_iso(const orc_cfg &c, const Vars &V, const Dir &D, int dir, const double *vL,
                  const double *vR, double *flux, double &cmax, FaceEmf &E) {
  const int nv_ = V.nvar;
  const int comps = c.comps;
  double uL[8] = {}, uR[8] = {}, fluxL[8] = {}, fluxR[8] = {};
  const double c2 = c2iso(c);
  double gpr = c2 * vL[V.RHO];
  double b1 = vL[D.BXn], b2 = at(vL, D.BXt), b3 = at(vL, D.BXb);
  double Btmag2 = b2 * b2 + b3 * b3;
  double Bmag2 = b1 * b1 + Btmag2;
  double cL = gpr - Bmag2;
  cL = gpr + Bmag2 + std::sqrt(cL * cL + 4.0 * gpr * Btmag2);
  cL = std::sqrt(0.5 * cL / vL[V.RHO]);
  gpr = c2 * vR[V.RHO];
  b1 = vR[D.BXn];
  b2 = at(vR, D.BXt);
  b3 = at(vR, D.BXb);
  Btmag2 = b2 * b2 + b3 * b3;
  Bmag2 = b1 * b1 + Btmag2;
  double cR = gpr - Bmag2;
  cR = gpr + Bmag2 + std::sqrt(cR * cR + 4.0 * gpr * Btmag2);
  cR = std::sqrt(0.5 * cR / vR[V.RHO]);
  const double cminL = vL[D.Xn] - cL, cmaxL = vL[D.Xn] + cL;
  const double cminR = vR[D.Xn] - cR, cmaxR = vR[D.Xn] + cR;
  const double sl = FMIN(cminL, cminR);
  const double sr = FMAX(cmaxL, cmaxR);
  cmax = std::fmax(std::fabs(sl), std::fabs(sr));
  prim2cons(c, V, vL, uL);
  prim2cons(c, V, vR, uR);
  for (int n = 0; n < nv_; n++) {
    fluxL[n] = uL[n];
    fluxR[n] = uR[n];
  }
  kflux(c, V, D, fluxL, vL, fluxL, c2);
  kflux(c, V, D, fluxR, vR, fluxR, c2);
  if (sl > 0) {
    for (int n = 0; n < nv_; n++) flux[n] = fluxL[n];
  } else if (sr < 0) {
    for (int n = 0; n < nv_; n++) flux[n] = fluxR[n];
  } else {
    double usL[8], usR[8], usc[8];
    double scrh = 1.0 / (sr - sl);
    const double duL = sl - vL[D.Xn];
    const double duR = sr - vR[D.Xn];
    const double Bx = (sr * vR[D.BXn] - sl * vL[D.BXn]) * scrh;
    const double rho = (uR[V.RHO] * duR - uL[V.RHO] * duL) * scrh;
    flux[V.RHO] = (sl * uR[V.RHO] * duR - sr * uL[V.RHO] * duL) * scrh;
    const double sqrho = std::sqrt(rho);
    const double SM = flux[V.RHO] / rho;
    const double S1L = SM - std::fabs(Bx) / sqrho;
    const double S1R = SM + std::fabs(Bx) / sqrho;
    int revert_to_hll = 0;
    if ((S1L - sl) < 1.e-4 * (sr - sl)) revert_to_hll = 1;
    if ((S1R - sr) > -1.e-4 * (sr - sl)) revert_to_hll = 1;
    if (revert_to_hll) {
      scrh = 1.0 / (sr - sl);
      for (int n = 0; n < nv_; n++) {
        flux[n] = sl * sr * (uR[n] - uL[n]) + sr * fluxL[n] - sl * fluxR[n];
        flux[n] *= scrh;
      }
    } else {
      flux[D.Xn] = (sr * fluxL[D.Xn] - sl * fluxR[D.Xn] + sr * sl * (uR[D.Xn] - uL[D.Xn])) * scrh;
      flux[D.BXn] = sr * sl * (uR[D.BXn] - uL[D.BXn]) * scrh;
      const double scrhL = 1.0 / ((sl - S1L) * (sl - S1R));
      const double scrhR = 1.0 / ((sr - S1L) * (sr - S1R));
      if (comps >= 2) {
        usL[D.Xt] = rho * vL[D.Xt] - Bx * uL[D.BXt] * (SM - vL[D.Xn]) * scrhL;
        usR[D.Xt] = rho * vR[D.Xt] - Bx * uR[D.BXt] * (SM - vR[D.Xn]) * scrhR;
        usL[D.BXt] = uL[D.BXt] / rho * (uL[V.RHO] * duL * duL - Bx * Bx) * scrhL;
        usR[D.BXt] = uR[D.BXt] / rho * (uR[V.RHO] * duR * duR - Bx * Bx) * scrhR;
      }
      if (comps >= 3) {
        usL[D.Xb] = rho * vL[D.Xb] - Bx * uL[D.BXb] * (SM - vL[D.Xn]) * scrhL;
        usR[D.Xb] = rho * vR[D.Xb] - Bx * uR[D.BXb] * (SM - vR[D.Xn]) * scrhR;
        usL[D.BXb] = uL[D.BXb] / rho * (uL[V.RHO] * duL * duL - Bx * Bx) * scrhL;
        usR[D.BXb] = uR[D.BXb] / rho * (uR[V.RHO] * duR * duR - Bx * Bx) * scrhR;
      }
      if (S1L >= 0.0) {
        if (comps >= 2) {
          flux[D.Xt] = fluxL[D.Xt] + sl * (usL[D.Xt] - uL[D.Xt]);
          flux[D.BXt] = fluxL[D.BXt] + sl * (usL[D.BXt] - uL[D.BXt]);
        }
        if (comps >= 3) {
          flux[D.Xb] = fluxL[D.Xb] + sl * (usL[D.Xb] - uL[D.Xb]);
          flux[D.BXb] = fluxL[D.BXb] + sl * (usL[D.BXb] - uL[D.BXb]);
        }
      } else if (S1R <= 0.0) {
        if (comps >= 2) {
          flux[D.Xt] = fluxR[D.Xt] + sr * (usR[D.Xt] - uR[D.Xt]);
          flux[D.BXt] = fluxR[D.BXt] + sr * (usR[D.BXt] - uR[D.BXt]);
        }
        if (comps >= 3) {
          flux[D.Xb] = fluxR[D.Xb] + sr * (usR[D.Xb] - uR[D.Xb]);
          flux[D.BXb] = fluxR[D.BXb] + sr * (usR[D.BXb] - uR[D.BXb]);
        }
      } else {
        const double sBx = (Bx > 0.0 ? 1.0 : -1.0);
        if (comps >= 2) {
          usc[D.Xt] = 0.5 * (usR[D.Xt] + usL[D.Xt] + (usR[D.BXt] - usL[D.BXt]) * sBx * sqrho);
          usc[D.BXt] = 0.5 * (usR[D.BXt] + usL[D.BXt] + (usR[D.Xt] - usL[D.Xt]) * sBx / sqrho);
          flux[D.Xt] = usc[D.Xt] * SM - Bx * usc[D.BXt];
          flux[D.BXt] = usc[D.BXt] * SM - Bx * usc[D.Xt] / rho;
        }
        if (comps >= 3) {
          usc[D.Xb] = 0.5 * (usR[D.Xb] + usL[D.Xb] + (usR[D.BXb] - usL[D.BXb]) * sBx * sqrho);
          usc[D.BXb] = 0.5 * (usR[D.BXb] + usL[D.BXb] + (usR[D.Xb] - usL[D.Xb]) * sBx / sqrho);
          flux[D.Xb] = usc[D.Xb] * SM - Bx * usc[D.BXb];
          flux[D.BXb] = usc[D.BXb] * SM - Bx * usc[D.Xb] / rho;
        }
      }
    }
  }
  double b2L = vL[V.BX1] * vL[V.BX1];
  if (comps >= 2) b2L = b2L + vL[V.BX2] * vL[V.BX2];
  if (comps >= 3) b2L = b2L + vL[V.BX3] * vL[V.BX3];
  double b2R = vR[V.BX1] * vR[V.BX1];
  if (comps >= 2) b2R = b2R + vR[V.BX2] * vR[V.BX2];
  if (comps >= 3) b2R = b2R + vR[V.BX3] * vR[V.BX3];
  store_face_emf(c, V, D, dir, vL, vR, uL, uR, flux, sl, sr, c2 * vL[V.RHO] + 0.5 * b2L,
                 c2 * vR[V.RHO] + 0.5 * b2R, E);
}

void hlld_mhd(const orc_cfg &c, const Vars &V, const Dir &D, int dir, const double *vL,
              const double *vR, double *flux, double &cmax, FaceEmf &E) {
  if (c.isothermal) return hlld_mhd_iso(c, V, D, dir, vL, vR, flux, cmax, E);
  const int nv_ = V.nvar;
  const int comps = c.comps;
  double uL[8], uR[8], fluxL[8], fluxR[8];
  const double gamma = c.gamma;

  // fast magnetosonic speeds, hlldMHD.hpp:183-226
  double gpr = gamma * vL[V.ENG];
  double b1 = vL[D.BXn], b2 = at(vL, D.BXt), b3 = at(vL, D.BXb);
  double Btmag2 = b2 * b2 + b3 * b3;
  double Bmag2 = b1 * b1 + Btmag2;
  double cL = gpr - Bmag2;
  cL = gpr + Bmag2 + std::sqrt(cL * cL + 4.0 * gpr * Btmag2);
  cL = std::sqrt(0.5 * cL / vL[V.RHO]);

  gpr = gamma * vR[V.ENG];
  b1 = vR[D.BXn];
  b2 = at(vR, D.BXt);
  b3 = at(vR, D.BXb);
  Btmag2 = b2 * b2 + b3 * b3;
  Bmag2 = b1 * b1 + Btmag2;
  double cR = gpr - Bmag2;
  cR = gpr + Bmag2 + std::sqrt(cR * cR + 4.0 * gpr * Btmag2);
  cR = std::sqrt(0.5 * cR / vR[V.RHO]);

  const double cminL = vL[D.Xn] - cL, cmaxL = vL[D.Xn] + cL;
  const double cminR = vR[D.Xn] - cR, cmaxR = vR[D.Xn] + cR;
  const double sl = FMIN(cminL, cminR);
  const double sr = FMAX(cmaxL, cmaxR);
  cmax = std::fmax(std::fabs(sl), std::fabs(sr));

  prim2cons(c, V, vL, uL);
  prim2cons(c, V, vR, uR);
  for (int n = 0; n < nv_; n++) {
    fluxL[n] = uL[n];
    fluxR[n] = uR[n];
  }
  kflux(c, V, D, fluxL, vL, fluxL);
  kflux(c, V, D, fluxR, vR, fluxR);

  double b2L = vL[V.BX1] * vL[V.BX1];
  if (comps >= 2) b2L = b2L + vL[V.BX2] * vL[V.BX2];
  if (comps >= 3) b2L = b2L + vL[V.BX3] * vL[V.BX3];
  double b2R = vR[V.BX1] * vR[V.BX1];
  if (comps >= 2) b2R = b2R + vR[V.BX2] * vR[V.BX2];
  if (comps >= 3) b2R = b2R + vR[V.BX3] * vR[V.BX3];
  const double ptL = vL[V.ENG] + 0.5 * b2L;
  const double ptR = vR[V.ENG] + 0.5 * b2R;

  if (sl > 0) {
    for (int n = 0; n < nv_; n++) flux[n] = fluxL[n];
  } else if (sr < 0) {
    for (int n = 0; n < nv_; n++) flux[n] = fluxR[n];
  } else {
    double usL[8], usR[8], Uhll[8];
    double scrh, scrhL, scrhR, duL, duR, sBx, Bx, SM, S1L, S1R;
    double pts, sqrL, sqrR, vsL = 0, vsR = 0, wsL = 0, wsR = 0;
    for (int n = 0; n < 8; n++) { usL[n] = 0; usR[n] = 0; }
    int revert_to_hllc = 0;

    scrh = 1.0 / (sr - sl);
    Bx = (sr * vR[D.BXn] - sl * vL[D.BXn]) * scrh;
    sBx = (Bx > 0.0 ? 1.0 : -1.0);
    duL = sl - vL[D.Xn];
    duR = sr - vR[D.Xn];
    scrh = 1.0 / (duR * uR[V.RHO] - duL * uL[V.RHO]);
    SM = (duR * uR[D.Xn] - duL * uL[D.Xn] - ptR + ptL) * scrh;
    pts = duR * uR[V.RHO] * ptL - duL * uL[V.RHO] * ptR +
          vL[V.RHO] * vR[V.RHO] * duR * duL * (vR[D.Xn] - vL[D.Xn]);
    pts *= scrh;
    usL[V.RHO] = uL[V.RHO] * duL / (sl - SM);
    usR[V.RHO] = uR[V.RHO] * duR / (sr - SM);
    sqrL = std::sqrt(usL[V.RHO]);
    sqrR = std::sqrt(usR[V.RHO]);
    S1L = SM - std::fabs(Bx) / sqrL;
    S1R = SM + std::fabs(Bx) / sqrR;

    if ((S1L - sl) < 1.e-4 * (SM - sl)) revert_to_hllc = 1;
    if ((S1R - sr) > -1.e-4 * (sr - SM)) revert_to_hllc = 1;

    if (revert_to_hllc) {
      scrh = 1.0 / (sr - sl);
      for (int n = 0; n < nv_; n++) {
        Uhll[n] = sr * uR[n] - sl * uL[n] + fluxL[n] - fluxR[n];
        Uhll[n] *= scrh;
      }
      usL[D.BXn] = usR[D.BXn] = Uhll[D.BXn];
      if (comps >= 2) usL[D.BXt] = usR[D.BXt] = Uhll[D.BXt];
      if (comps >= 3) usL[D.BXb] = usR[D.BXb] = Uhll[D.BXb];
      S1L = S1R = SM;
    } else {
      scrhL = (uL[V.RHO] * duL * duL - Bx * Bx) / (uL[V.RHO] * duL * (sl - SM) - Bx * Bx);
      scrhR = (uR[V.RHO] * duR * duR - Bx * Bx) / (uR[V.RHO] * duR * (sr - SM) - Bx * Bx);
      usL[D.BXn] = Bx;
      if (comps >= 2) usL[D.BXt] = uL[D.BXt] * scrhL;
      if (comps >= 3) usL[D.BXb] = uL[D.BXb] * scrhL;
      usR[D.BXn] = Bx;
      if (comps >= 2) usR[D.BXt] = uR[D.BXt] * scrhR;
      if (comps >= 3) usR[D.BXb] = uR[D.BXb] * scrhR;
    }

    scrhL = Bx / (uL[V.RHO] * duL);
    scrhR = Bx / (uR[V.RHO] * duR);
    if (comps >= 2) {
      vsL = vL[D.Xt] - scrhL * (usL[D.BXt] - uL[D.BXt]);
      vsR = vR[D.Xt] - scrhR * (usR[D.BXt] - uR[D.BXt]);
    }
    if (comps >= 3) {
      wsL = vL[D.Xb] - scrhL * (usL[D.BXb] - uL[D.BXb]);
      wsR = vR[D.Xb] - scrhR * (usR[D.BXb] - uR[D.BXb]);
    }
    usL[D.Xn] = usL[V.RHO] * SM;
    usR[D.Xn] = usR[V.RHO] * SM;
    if (comps >= 2) {
      usL[D.Xt] = usL[V.RHO] * vsL;
      usR[D.Xt] = usR[V.RHO] * vsR;
    }
    if (comps >= 3) {
      usL[D.Xb] = usL[V.RHO] * wsL;
      usR[D.Xb] = usR[V.RHO] * wsR;
    }

    scrhL = vL[D.Xn] * Bx;
    if (comps >= 2) scrhL = scrhL + vL[D.Xt] * uL[D.BXt];
    if (comps >= 3) scrhL = scrhL + vL[D.Xb] * uL[D.BXb];
    {
      double s = SM * Bx;
      if (comps >= 2) s = s + vsL * usL[D.BXt];
      if (comps >= 3) s = s + wsL * usL[D.BXb];
      scrhL -= s;
    }
    usL[V.ENG] = duL * uL[V.ENG] - ptL * vL[D.Xn] + pts * SM + Bx * scrhL;
    usL[V.ENG] /= sl - SM;

    scrhR = vR[D.Xn] * Bx;
    if (comps >= 2) scrhR = scrhR + vR[D.Xt] * uR[D.BXt];
    if (comps >= 3) scrhR = scrhR + vR[D.Xb] * uR[D.BXb];
    {
      double s = SM * Bx;
      if (comps >= 2) s = s + vsR * usR[D.BXt];
      if (comps >= 3) s = s + wsR * usR[D.BXb];
      scrhR -= s;
    }
    usR[V.ENG] = duR * uR[V.ENG] - ptR * vR[D.Xn] + pts * SM + Bx * scrhR;
    usR[V.ENG] /= sr - SM;

    if (S1L >= 0.0) {
      for (int n = 0; n < nv_; n++) flux[n] = fluxL[n] + sl * (usL[n] - uL[n]);
    } else if (S1R <= 0.0) {
      for (int n = 0; n < nv_; n++) flux[n] = fluxR[n] + sr * (usR[n] - uR[n]);
    } else {
      double vss = 0, wss = 0, ussl[8], ussr[8];
      for (int n = 0; n < 8; n++) { ussl[n] = 0; ussr[n] = 0; }
      if (g_plant == ORC_PLANT_HLLD_SBX) sBx = -sBx;
      ussl[V.RHO] = usL[V.RHO];
      ussr[V.RHO] = usR[V.RHO];
      if (comps >= 2) {
        vss = sqrL * vsL + sqrR * vsR + (usR[D.BXt] - usL[D.BXt]) * sBx;
        vss /= sqrL + sqrR;
      }
      if (comps >= 3) {
        wss = sqrL * wsL + sqrR * wsR + (usR[D.BXb] - usL[D.BXb]) * sBx;
        wss /= sqrL + sqrR;
      }
      ussl[D.Xn] = ussl[V.RHO] * SM;
      ussr[D.Xn] = ussr[V.RHO] * SM;
      if (comps >= 2) {
        ussl[D.Xt] = ussl[V.RHO] * vss;
        ussr[D.Xt] = ussr[V.RHO] * vss;
      }
      if (comps >= 3) {
        ussl[D.Xb] = ussl[V.RHO] * wss;
        ussr[D.Xb] = ussr[V.RHO] * wss;
      }
      ussl[D.BXn] = ussr[D.BXn] = Bx;
      if (comps >= 2) {
        ussl[D.BXt] = sqrL * usR[D.BXt] + sqrR * usL[D.BXt] + sqrL * sqrR * (vsR - vsL) * sBx;
        ussl[D.BXt] /= sqrL + sqrR;
        ussr[D.BXt] = ussl[D.BXt];
      }
      if (comps >= 3) {
        ussl[D.BXb] = sqrL * usR[D.BXb] + sqrR * usL[D.BXb] + sqrL * sqrR * (wsR - wsL) * sBx;
        ussl[D.BXb] /= sqrL + sqrR;
        ussr[D.BXb] = ussl[D.BXb];
      }
      scrhL = SM * Bx;
      if (comps >= 2) scrhL = scrhL + vsL * usL[D.BXt];
      if (comps >= 3) scrhL = scrhL + wsL * usL[D.BXb];
      {
        double s = SM * Bx;
        if (comps >= 2) s = s + vss * ussl[D.BXt];
        if (comps >= 3) s = s + wss * ussl[D.BXb];
        scrhL -= s;
      }
      scrhR = SM * Bx;
      if (comps >= 2) scrhR = scrhR + vsR * usR[D.BXt];
      if (comps >= 3) scrhR = scrhR + wsR * usR[D.BXb];
      {
        double s = SM * Bx;
        if (comps >= 2) s = s + vss * ussr[D.BXt];
        if (comps >= 3) s = s + wss * ussr[D.BXb];
        scrhR -= s;
      }
      ussl[V.ENG] = usL[V.ENG] - sqrL * scrhL * sBx;
      ussr[V.ENG] = usR[V.ENG] + sqrR * scrhR * sBx;
      if (SM >= 0.0) {
        for (int n = 0; n < nv_; n++)
          flux[n] = fluxL[n] + S1L * (ussl[n] - usL[n]) + sl * (usL[n] - uL[n]);
      } else {
        for (int n = 0; n < nv_; n++)
          flux[n] = fluxR[n] + S1R * (ussr[n] - usR[n]) + sr * (usR[n] - uR[n]);
      }
    }
  }

  store_face_emf(c, V, D, dir, vL, vR, uL, uR, flux, sl, sr, ptL, ptR, E);

}

// Face EMF stores (storeFlux.hpp:14-208) with the signs of hlldMHD.hpp:79-142 /
// roeMHD.hpp:106-170; shared by HLLD and Roe MHD.
void store_face_emf(const orc_cfg &c, const Vars &V, const Dir &D, int dir, const double *vL,
                    const double *vR, const double *uL, const double *uR, const double *flux,
                    double sl, double sr, double ptL, double ptR, FaceEmf &E) {
  const int comps = c.comps;
  double st = 1.0, sb = 1.0;  // hlldMHD.hpp:79-142
  if (dir == IDIR) { st = -1.0; if (c.dim == 3) sb = +1.0; }
  if (dir == JDIR) { st = +1.0; if (c.dim == 3) sb = -1.0; }
  if (dir == KDIR) { st = -1.0; sb = +1.0; }
  E.hasEt = E.hasEb = E.hasSv = E.hasAD = false;
  if (c.emf == ORC_EMF_ARITHMETIC || c.emf == ORC_EMF_UCT0 || c.emf == ORC_EMF_UCT_CONTACT) {
    // K_StoreEMF, storeFlux.hpp:14-29
    if (comps > 1) {
      E.Et = st * flux[D.BXt];
      E.hasEt = true;
      if (c.dim == 3) {
        E.Eb = sb * flux[D.BXb];
        E.hasEb = true;
      }
    }
    if (c.emf == ORC_EMF_UCT_CONTACT) {
      // K_StoreContact, storeFlux.hpp:31-44 (eps_UCT_CONTACT = 1e-6)
      double s = 0.5;
      if (flux[V.RHO] > 1.0e-6) s = 1.0;
      if (flux[V.RHO] < -1.0e-6) s = 0.0;
      E.sv = s;
      E.hasSv = true;
    }
  } else if (c.emf == ORC_EMF_UCT_HLL) {
    // K_StoreHLL, storeFlux.hpp:46-75
    const double ar = std::fmax(0.0, sr);
    const double al = std::fmin(0.0, sl);
    const double scrh = 1.0 / (ar - al);
    if (comps > 1) {
      E.Et = -st * (ar * vL[D.Xt] - al * vR[D.Xt]) * scrh;
      E.hasEt = true;
      if (comps >= 3) {
        E.Eb = -sb * (ar * vL[D.Xb] - al * vR[D.Xb]) * scrh;
        E.hasEb = true;
      }
    }
    E.aL = ar * scrh;
    E.aR = -al * scrh;
    E.dR = -al * ar * scrh;
    E.dL = E.dR;
    E.hasAD = true;
  } else if (c.emf == ORC_EMF_UCT_HLLD) {
    // K_StoreHLLD, storeFlux.hpp:77-208 (HAVE_ENERGY branch)
    const double ptL2 = ptL, ptR2 = ptR;
    int revert_to_hll = 0;
    const double Bn = (sr * vR[D.BXn] - sl * vL[D.BXn]) / (sr - sl);
    const double duL = sl - vL[D.Xn];
    const double duR = sr - vR[D.Xn];
    double scrh, Sc, SaL, SaR;
    if (c.isothermal) {
      // storeFlux.hpp:139-147
      scrh = 1.0 / (sr - sl);
      const double rho_h = (uR[V.RHO] * duR - uL[V.RHO] * duL) * scrh;
      Sc = (sl * uR[V.RHO] * duR - sr * uL[V.RHO] * duL) * scrh / rho_h;
      const double sqrho_h = std::sqrt(rho_h);
      SaL = Sc - std::fabs(Bn) / sqrho_h;
      SaR = Sc + std::fabs(Bn) / sqrho_h;
    } else {
      scrh = 1.0 / (duR * uR[V.RHO] - duL * uL[V.RHO]);
      Sc = (duR * uR[D.Xn] - duL * uL[D.Xn] - ptR2 + ptL2) * scrh;
      const double usLRHO = uL[V.RHO] * duL / (sl - Sc);
      const double usRRHO = uR[V.RHO] * duR / (sr - Sc);
      const double sqrL = std::sqrt(usLRHO);
      const double sqrR = std::sqrt(usRRHO);
      SaL = Sc - std::fabs(Bn) / sqrL;
      SaR = Sc + std::fabs(Bn) / sqrR;
    }
    if ((SaL - sl) < 1e-4 * (Sc - sl)) revert_to_hll = 1;
    if ((SaR - sr) > -1e-4 * (sr - Sc)) revert_to_hll = 1;
    const double chiL = (vL[D.Xn] - Sc) * (sl - Sc) / (SaL + sl - 2.0 * Sc);
    const double chiR = (vR[D.Xn] - Sc) * (sr - Sc) / (SaR + sr - 2.0 * Sc);
    const double nuL = (SaL + sl) / (std::fabs(SaL) + std::fabs(sl));
    const double nuR = (SaR + sr) / (std::fabs(SaR) + std::fabs(sr));
    const double nuLR = (SaL + SaR) / (std::fabs(SaL) + std::fabs(SaR));
    if (std::fabs(SaR - SaL) > 1.e-9 * std::fabs(sr - sl)) {
      E.dL = 0.5 * (chiL * nuL - chiL * nuLR) + 0.5 * (std::fabs(SaL) - nuLR * SaL);
      E.dR = 0.5 * (chiR * nuR - chiR * nuLR) + 0.5 * (std::fabs(SaR) - nuLR * SaR);
      E.aL = 0.5 * (1.0 + nuLR);
      E.aR = 0.5 * (1.0 - nuLR);
    } else {
      E.dL = 0.5 * chiL * nuL + 0.5 * std::fabs(SaL);
      E.dR = 0.5 * chiR * nuR + 0.5 * std::fabs(SaR);
      E.aL = 0.5;
      E.aR = 0.5;
    }
    const double ar = std::fmax(0.0, sr);
    const double al = std::fmin(0.0, sl);
    scrh = 1.0 / (ar - al);
    if (revert_to_hll) {
      E.aL = ar * scrh;
      E.aR = -al * scrh;
      E.dR = -al * ar * scrh;
      E.dL = E.dR;
    }
    E.hasAD = true;
    if (comps > 1) {
      E.Et = -st * (ar * vL[D.Xt] - al * vR[D.Xt]) * scrh;
      E.hasEt = true;
      if (comps >= 3) {
        E.Eb = -sb * (ar * vL[D.Xb] - al * vR[D.Xb]) * scrh;
        E.hasEb = true;
      }
    }
  }}

// --------------------------------------------------------------------------------------
// HllMHD: src/fluid/RiemannSolver/MHDsolvers/hllMHD.hpp:170-453 (HAVE_ENERGY, no Hall):
// density and momentum fluxes divide by (sr - sl), field and energy fluxes multiply by
// 1/(sr - sl); the face stores cover arithmetic, uct0, uct_contact and uct_hll (the
// reference has no uct_hlld store for this solver).
// --------------------------------------------------------------------------------------
void hll_mhd(const orc_cfg &c, const Vars &V, const Dir &D, int dir, const double *vL,
             const double *vR, double *flux, double &cmax, FaceEmf &E) {
  const int comps = c.comps;
  double uL[8], uR[8], fluxL[8], fluxR[8];
  const double gamma = c.gamma;
  const double c2 = c.isothermal ? c2iso(c) : 0.0;
  double gpr = c.isothermal ? c2 * vL[V.RHO] : gamma * vL[V.ENG];
  double b1 = vL[D.BXn], b2 = at(vL, D.BXt), b3 = at(vL, D.BXb);
  double Btmag2 = b2 * b2 + b3 * b3;
  double Bmag2 = b1 * b1 + Btmag2;
  double cL = gpr - Bmag2;
  cL = gpr + Bmag2 + std::sqrt(cL * cL + 4.0 * gpr * Btmag2);
  cL = std::sqrt(0.5 * cL / vL[V.RHO]);
  gpr = c.isothermal ? c2 * vR[V.RHO] : gamma * vR[V.ENG];
  b1 = vR[D.BXn];
  b2 = at(vR, D.BXt);
  b3 = at(vR, D.BXb);
  Btmag2 = b2 * b2 + b3 * b3;
  Bmag2 = b1 * b1 + Btmag2;
  double cR = gpr - Bmag2;
  cR = gpr + Bmag2 + std::sqrt(cR * cR + 4.0 * gpr * Btmag2);
  cR = std::sqrt(0.5 * cR / vR[V.RHO]);
  const double cminL = vL[D.Xn] - cL, cmaxL = vL[D.Xn] + cL;
  const double cminR = vR[D.Xn] - cR, cmaxR = vR[D.Xn] + cR;
  const double sl = FMIN(cminL, cminR);
  const double sr = FMAX(cmaxL, cmaxR);
  cmax = FMAX(std::fabs(sl), std::fabs(sr));
  prim2cons(c, V, vL, uL);
  prim2cons(c, V, vR, uR);
  kflux(c, V, D, fluxL, vL, uL, c2);
  kflux(c, V, D, fluxR, vR, uR, c2);
  // density and momentum
  for (int n = V.RHO; n <= V.RHO + comps; n++) {
    if (sl > 0) flux[n] = fluxL[n];
    else if (sr < 0) flux[n] = fluxR[n];
    else flux[n] = (sl * sr * uR[n] - sl * sr * uL[n] + sr * fluxL[n] - sl * fluxR[n]) / (sr - sl);
  }
  // magnetic field and energy
  for (int n = c.isothermal ? V.BX1 : V.ENG; n < V.nvar; n++) {
    if (sl > 0) {
      flux[n] = fluxL[n];
    } else if (sr < 0) {
      flux[n] = fluxR[n];
    } else {
      flux[n] = sl * sr * uR[n] - sl * sr * uL[n] + sr * fluxL[n] - sl * fluxR[n];
      flux[n] *= (1.0 / (sr - sl));
    }
  }
  double b2L = vL[V.BX1] * vL[V.BX1];
  if (comps >= 2) b2L = b2L + vL[V.BX2] * vL[V.BX2];
  if (comps >= 3) b2L = b2L + vL[V.BX3] * vL[V.BX3];
  double b2R = vR[V.BX1] * vR[V.BX1];
  if (comps >= 2) b2R = b2R + vR[V.BX2] * vR[V.BX2];
  if (comps >= 3) b2R = b2R + vR[V.BX3] * vR[V.BX3];
  store_face_emf(c, V, D, dir, vL, vR, uL, uR, flux, sl, sr,
                 (c.isothermal ? c2 * vL[V.RHO] : vL[V.ENG]) + 0.5 * b2L,
                 (c.isothermal ? c2 * vR[V.RHO] : vR[V.ENG]) + 0.5 * b2R, E);
}

// --------------------------------------------------------------------------------------
// TvdlfMHD: src/fluid/RiemannSolver/MHDsolvers/tvdlfMHD.hpp:138-234 (HAVE_ENERGY, no
// Hall): Rusanov flux with the fast speed of the arithmetic mean state; the face range
// always extends by one transverse cell and the uct_hll store sees sl = -cmax, sr = cmax.
// --------------------------------------------------------------------------------------
void tvdlf_mhd(const orc_cfg &c, const Vars &V, const Dir &D, int dir, const double *vL,
               const double *vR, double *flux, double &cmax, FaceEmf &E) {
  const int nv_ = V.nvar;
  const int comps = c.comps;
  double v[8], uL[8], uR[8], fluxL[8], fluxR[8];
  for (int n = 0; n < nv_; n++) v[n] = 0.5 * (vL[n] + vR[n]);
  const double c2 = c.isothermal ? c2iso(c) : 0.0;
  const double gpr = c.isothermal ? c2 * v[V.RHO] : c.gamma * v[V.ENG];
  double Bt2 = 0.0;
  if (comps >= 2) Bt2 = Bt2 + v[D.BXt] * v[D.BXt];
  if (comps >= 3) Bt2 = Bt2 + v[D.BXb] * v[D.BXb];
  const double B2 = Bt2 + v[D.BXn] * v[D.BXn];
  double cRL = gpr - B2;
  cRL = cRL + B2 + std::sqrt(cRL * cRL + 4.0 * gpr * Bt2);
  cRL = std::sqrt(0.5 * cRL / v[V.RHO]);
  cmax = std::fmax(std::fabs(v[D.Xn] + cRL), std::fabs(v[D.Xn] - cRL));
  const double sl = -cmax, sr = cmax;
  prim2cons(c, V, vL, uL);
  prim2cons(c, V, vR, uR);
  kflux(c, V, D, fluxL, vL, uL, c2);
  kflux(c, V, D, fluxR, vR, uR, c2);
  for (int n = 0; n < nv_; n++) flux[n] = 0.5 * (fluxL[n] + fluxR[n] + cmax * (uL[n] - uR[n]));
  store_face_emf(c, V, D, dir, vL, vR, uL, uR, flux, sl, sr, 0.0, 0.0, E);
}

// --------------------------------------------------------------------------------------
// RoeMHD: src/fluid/RiemannSolver/MHDsolvers/roeMHD.hpp:175-651 (ROE_AVERAGE 0).  With
// an energy equation 8 waves: fast-, fast+, div B, entropy, slow-, slow+, Alfven-,
// Alfven+; isothermal (the !HAVE_ENERGY branches, :21-41, :225-232, :296-302, :325-328,
// :427-431, :454-461, :515-520, :541-546) 7 waves without the entropy wave, the sound
// speed squared a2 = Cs2Iso + X, and the fast and slow wave strengths built on a2 dV[RHO].
// --------------------------------------------------------------------------------------
void roe_mhd(const orc_cfg &c, const Vars &V, const Dir &D, int dir, const double *vL,
             const double *vR, double *flux, double &cmax, FaceEmf &E) {
  const int nv_ = V.nvar;
  const int comps = c.comps;
  const bool iso = c.isothermal != 0;
  const int KFASTM = 0, KFASTP = 1, KDIVB = 2, KENTRP = 3, KSLOWM = iso ? 3 : 4,
            KSLOWP = iso ? 4 : 5, KALFVM = iso ? 5 : 6, KALFVP = iso ? 6 : 7;
  const double delta = 1.e-6;
  auto DSIGN = [](double x) { return x >= 0.0 ? 1.0 : -1.0; };
  double dV[8], uL[8], uR[8], dU[8], fluxL[8], fluxR[8], Rc[8][8];
  for (int n = 0; n < nv_; n++) dV[n] = vR[n] - vL[n];
  prim2cons(c, V, vL, uL);
  prim2cons(c, V, vR, uR);
  // with an energy equation a2L, a2R are placeholders (roeMHD.hpp:223-227)
  const double a2L = iso ? c2iso(c) : 1.0, a2R = iso ? a2L : 1.0;
  const double gamma = c.gamma;
  for (int n = 0; n < nv_; n++) {
    fluxL[n] = uL[n];
    fluxR[n] = uR[n];
    dU[n] = uR[n] - uL[n];
  }
  kflux(c, V, D, fluxL, vL, fluxL, a2L);
  kflux(c, V, D, fluxR, vR, fluxR, a2R);
  for (int p = 0; p < 8; p++)
    for (int q = 0; q < 8; q++) Rc[p][q] = 0;

  const double sqr_rho_L = std::sqrt(vL[V.RHO]);
  const double sqr_rho_R = std::sqrt(vR[V.RHO]);
  double sl = sqr_rho_L / (sqr_rho_L + sqr_rho_R);
  double sr = sqr_rho_R / (sqr_rho_L + sqr_rho_R);
  const double rho = sr * vL[V.RHO] + sl * vR[V.RHO];
  const double sqrt_rho = std::sqrt(rho);
  double u = 0, v = 0, w = 0, Bx = 0, By = 0, Bz = 0, bx = 0, by = 0, bz = 0;
  u = sl * vL[D.Xn] + sr * vR[D.Xn];
  if (comps >= 2) v = sl * vL[D.Xt] + sr * vR[D.Xt];
  if (comps >= 3) w = sl * vL[D.Xb] + sr * vR[D.Xb];
  Bx = sr * vL[D.BXn] + sl * vR[D.BXn];
  if (comps >= 2) By = sr * vL[D.BXt] + sl * vR[D.BXt];
  if (comps >= 3) Bz = sr * vL[D.BXb] + sl * vR[D.BXb];
  const double sBx = (Bx >= 0.0 ? 1.0 : -1.0);
  bx = Bx / sqrt_rho;
  if (comps >= 2) by = By / sqrt_rho;
  if (comps >= 3) bz = Bz / sqrt_rho;
  double bt2 = 0.0;
  if (comps >= 2) bt2 = bt2 + by * by;
  if (comps >= 3) bt2 = bt2 + bz * bz;
  const double b2 = bx * bx + bt2;
  const double Btmag = std::sqrt(bt2 * rho);
  double X = dV[D.BXn] * dV[D.BXn];
  if (comps >= 2) X = X + dV[D.BXt] * dV[D.BXt];
  if (comps >= 3) X = X + dV[D.BXb] * dV[D.BXb];
  X /= (sqr_rho_L + sqr_rho_R) * (sqr_rho_L + sqr_rho_R) * 2.0;
  double Bmag2L = vL[V.BX1] * vL[V.BX1];
  if (comps >= 2) Bmag2L = Bmag2L + vL[V.BX2] * vL[V.BX2];
  if (comps >= 3) Bmag2L = Bmag2L + vL[V.BX3] * vL[V.BX3];
  double Bmag2R = vR[V.BX1] * vR[V.BX1];
  if (comps >= 2) Bmag2R = Bmag2R + vR[V.BX2] * vR[V.BX2];
  if (comps >= 3) Bmag2R = Bmag2R + vR[V.BX3] * vR[V.BX3];
  const double pL = (iso ? a2L * vL[V.RHO] : vL[V.ENG]) + 0.5 * Bmag2L;
  const double pR = (iso ? a2R * vR[V.RHO] : vR[V.ENG]) + 0.5 * Bmag2R;

  double vel2 = 0.0, Hgas = 0.0, a2;
  if (!iso) {
    double vdm = u * dU[D.Xn];
    if (comps >= 2) vdm = vdm + v * dU[D.Xt];
    if (comps >= 3) vdm = vdm + w * dU[D.Xb];
    double BdB = Bx * dU[D.BXn];
    if (comps >= 2) BdB = BdB + By * dU[D.BXt];
    if (comps >= 3) BdB = BdB + Bz * dU[D.BXb];
    vel2 = u * u;
    if (comps >= 2) vel2 = vel2 + v * v;
    if (comps >= 3) vel2 = vel2 + w * w;
    dV[V.ENG] = (gamma - 1.0) * ((0.5 * vel2 - X) * dV[V.RHO] - vdm + dU[V.ENG] - BdB);
    const double HL = (uL[V.ENG] + pL) / vL[V.RHO];
    const double HR = (uR[V.ENG] + pR) / vR[V.RHO];
    const double H = sl * HL + sr * HR;
    Hgas = H - b2;
    a2 = (2.0 - gamma) * X + (gamma - 1.0) * (Hgas - 0.5 * vel2);
  } else {
    a2 = 0.5 * (a2L + a2R) + X;
  }

  double scrh = a2 - b2;
  const double ca2 = bx * bx;
  scrh = scrh * scrh + 4.0 * bt2 * a2;
  scrh = std::sqrt(scrh);
  const double cf2 = 0.5 * (a2 + b2 + scrh);
  const double cs2 = a2 * ca2 / cf2;
  const double cf = std::sqrt(cf2);
  const double cs = std::sqrt(cs2);
  const double ca = std::sqrt(ca2);
  const double a = std::sqrt(a2);
  double alpha_f, alpha_s;
  if (cf == cs) {
    alpha_f = 1.0;
    alpha_s = 0.0;
  } else if (a <= cs) {
    alpha_f = 0.0;
    alpha_s = 1.0;
  } else if (cf <= a) {
    alpha_f = 1.0;
    alpha_s = 0.0;
  } else {
    scrh = 1.0 / (cf2 - cs2);
    alpha_f = (a2 - cs2) * scrh;
    alpha_s = (cf2 - a2) * scrh;
    alpha_f = FMAX(0.0, alpha_f);
    alpha_s = FMAX(0.0, alpha_s);
    alpha_f = std::sqrt(alpha_f);
    alpha_s = std::sqrt(alpha_s);
  }
  double beta_y = 0.0, beta_z = 0.0;
  if (Btmag > 1.e-9) {
    if (comps == 2) beta_y = DSIGN(By);
    if (comps == 3) {
      beta_y = By / Btmag;
      beta_z = Bz / Btmag;
    }
  } else {
    if (comps == 2) beta_y = 1.0;
    if (comps == 3) beta_z = beta_y = 1.0;
  }

  double lambda[8], alambda[8], eta[8];
  for (int n = 0; n < 8; n++) { lambda[n] = 0; eta[n] = 0; }
  // fast wave u - c_f
  int kk = KFASTM;
  lambda[kk] = u - cf;
  scrh = alpha_s * cs * sBx;
  double beta_dv = 0.0, beta_dB = 0.0, beta_v = 0.0;
  if (comps >= 2) {
    beta_dv = beta_dv + beta_y * dV[D.Xt];
    beta_dB = beta_dB + beta_y * dV[D.BXt];
    beta_v = beta_v + beta_y * v;
  }
  if (comps >= 3) {
    beta_dv = beta_dv + beta_z * dV[D.Xb];
    beta_dB = beta_dB + beta_z * dV[D.BXb];
    beta_v = beta_v + beta_z * w;
  }
  Rc[V.RHO][kk] = alpha_f;
  Rc[D.Xn][kk] = alpha_f * lambda[kk];
  if (comps >= 2) Rc[D.Xt][kk] = alpha_f * v + scrh * beta_y;
  if (comps >= 3) Rc[D.Xb][kk] = alpha_f * w + scrh * beta_z;
  if (comps >= 2) Rc[D.BXt][kk] = alpha_s * a * beta_y / sqrt_rho;
  if (comps >= 3) Rc[D.BXb][kk] = alpha_s * a * beta_z / sqrt_rho;
  if (!iso) {
    Rc[V.ENG][kk] = alpha_f * (Hgas - u * cf) + scrh * beta_v + alpha_s * a * Btmag / sqrt_rho;
    eta[kk] = alpha_f * (X * dV[V.RHO] + dV[V.ENG]);
  } else {
    eta[kk] = alpha_f * a2 * dV[V.RHO];
  }
  eta[kk] += rho * scrh * beta_dv - rho * alpha_f * cf * dV[D.Xn] + sqrt_rho * alpha_s * a * beta_dB;
  eta[kk] *= 0.5 / a2;
  // fast wave u + c_f
  kk = KFASTP;
  lambda[kk] = u + cf;
  Rc[V.RHO][kk] = alpha_f;
  Rc[D.Xn][kk] = alpha_f * lambda[kk];
  if (comps >= 2) Rc[D.Xt][kk] = alpha_f * v - scrh * beta_y;
  if (comps >= 3) Rc[D.Xb][kk] = alpha_f * w - scrh * beta_z;
  if (comps >= 2) Rc[D.BXt][kk] = Rc[D.BXt][KFASTM];
  if (comps >= 3) Rc[D.BXb][kk] = Rc[D.BXb][KFASTM];
  if (!iso) {
    Rc[V.ENG][kk] = alpha_f * (Hgas + u * cf) - scrh * beta_v + alpha_s * a * Btmag / sqrt_rho;
    eta[kk] = alpha_f * (X * dV[V.RHO] + dV[V.ENG]) - rho * scrh * beta_dv +
              rho * alpha_f * cf * dV[D.Xn] + sqrt_rho * alpha_s * a * beta_dB;
  } else {
    eta[kk] = alpha_f * (0. * X + a2) * dV[V.RHO] - rho * scrh * beta_dv +
              rho * alpha_f * cf * dV[D.Xn] + sqrt_rho * alpha_s * a * beta_dB;
  }
  eta[kk] *= 0.5 / a2;
  // entropy wave u (with an energy equation only)
  if (!iso) {
    kk = KENTRP;
    lambda[kk] = u;
    Rc[V.RHO][kk] = 1.0;
    Rc[D.Xn][kk] = u;
    if (comps >= 2) Rc[D.Xt][kk] = v;
    if (comps >= 3) Rc[D.Xb][kk] = w;
    Rc[V.ENG][kk] = 0.5 * vel2 + (gamma - 2.0) / (gamma - 1.0) * X;
    eta[kk] = ((a2 - X) * dV[V.RHO] - dV[V.ENG]) / a2;
  }
  // div B wave u (zero jump)
  kk = KDIVB;
  lambda[kk] = u;
  eta[kk] = 0.0;
  if (comps > 1) {
    scrh = alpha_f * cf * sBx;
    kk = KSLOWM;
    lambda[kk] = u - cs;
    Rc[V.RHO][kk] = alpha_s;
    Rc[D.Xn][kk] = alpha_s * lambda[kk];
    Rc[D.Xt][kk] = alpha_s * v - scrh * beta_y;
    if (comps >= 3) Rc[D.Xb][kk] = alpha_s * w - scrh * beta_z;
    Rc[D.BXt][kk] = -alpha_f * a * beta_y / sqrt_rho;
    if (comps >= 3) Rc[D.BXb][kk] = -alpha_f * a * beta_z / sqrt_rho;
    if (!iso) {
      Rc[V.ENG][kk] = alpha_s * (Hgas - u * cs) - scrh * beta_v - alpha_f * a * Btmag / sqrt_rho;
      eta[kk] = alpha_s * (X * dV[V.RHO] + dV[V.ENG]) - rho * scrh * beta_dv -
                rho * alpha_s * cs * dV[D.Xn] - sqrt_rho * alpha_f * a * beta_dB;
    } else {
      eta[kk] = alpha_s * (0. * X + a2) * dV[V.RHO] - rho * scrh * beta_dv -
                rho * alpha_s * cs * dV[D.Xn] - sqrt_rho * alpha_f * a * beta_dB;
    }
    eta[kk] *= 0.5 / a2;
    kk = KSLOWP;
    lambda[kk] = u + cs;
    Rc[V.RHO][kk] = alpha_s;
    Rc[D.Xn][kk] = alpha_s * lambda[kk];
    Rc[D.Xt][kk] = alpha_s * v + scrh * beta_y;
    if (comps >= 3) Rc[D.Xb][kk] = alpha_s * w + scrh * beta_z;
    Rc[D.BXt][kk] = Rc[D.BXt][KSLOWM];
    if (comps >= 3) Rc[D.BXb][kk] = Rc[D.BXb][KSLOWM];
    if (!iso) {
      Rc[V.ENG][kk] = alpha_s * (Hgas + u * cs) + scrh * beta_v - alpha_f * a * Btmag / sqrt_rho;
      eta[kk] = alpha_s * (X * dV[V.RHO] + dV[V.ENG]) + rho * scrh * beta_dv +
                rho * alpha_s * cs * dV[D.Xn] - sqrt_rho * alpha_f * a * beta_dB;
    } else {
      eta[kk] = alpha_s * (0. * X + a2) * dV[V.RHO] + rho * scrh * beta_dv +
                rho * alpha_s * cs * dV[D.Xn] - sqrt_rho * alpha_f * a * beta_dB;
    }
    eta[kk] *= 0.5 / a2;
  }
  if (comps == 3) {
    kk = KALFVM;
    lambda[kk] = u - ca;
    Rc[D.Xt][kk] = -rho * beta_z;
    Rc[D.Xb][kk] = +rho * beta_y;
    Rc[D.BXt][kk] = -sBx * sqrt_rho * beta_z;
    Rc[D.BXb][kk] = sBx * sqrt_rho * beta_y;
    if (!iso) Rc[V.ENG][kk] = -rho * (v * beta_z - w * beta_y);
    eta[kk] = +beta_y * dV[D.Xb] - beta_z * dV[D.Xt] +
              sBx / sqrt_rho * (beta_y * dV[D.BXb] - beta_z * dV[D.BXt]);
    eta[kk] *= 0.5;
    kk = KALFVP;
    lambda[kk] = u + ca;
    Rc[D.Xt][kk] = -Rc[D.Xt][KALFVM];
    Rc[D.Xb][kk] = -Rc[D.Xb][KALFVM];
    Rc[D.BXt][kk] = Rc[D.BXt][KALFVM];
    Rc[D.BXb][kk] = Rc[D.BXb][KALFVM];
    if (!iso) Rc[V.ENG][kk] = -Rc[V.ENG][KALFVM];
    eta[kk] = -beta_y * dV[D.Xb] + beta_z * dV[D.Xt] +
              sBx / sqrt_rho * (beta_y * dV[D.BXb] - beta_z * dV[D.BXt]);
    eta[kk] *= 0.5;
  }
  if (comps == 3 && g_plant == ORC_PLANT_ROE_ALFVEN) {
    const double e = eta[KALFVM];
    eta[KALFVM] = eta[KALFVP];
    eta[KALFVP] = e;
  }
  cmax = std::fabs(u) + cf;
  sl = lambda[KFASTM];
  sr = lambda[KFASTP];
  for (int n = 0; n < nv_; n++) alambda[n] = std::fabs(lambda[n]);
  if (alambda[KFASTM] < 0.5 * delta) alambda[KFASTM] = lambda[KFASTM] * lambda[KFASTM] / delta + 0.25 * delta;
  if (alambda[KFASTP] < 0.5 * delta) alambda[KFASTP] = lambda[KFASTP] * lambda[KFASTP] / delta + 0.25 * delta;
  if (comps > 1) {
    if (alambda[KSLOWM] < 0.5 * delta) alambda[KSLOWM] = lambda[KSLOWM] * lambda[KSLOWM] / delta + 0.25 * delta;
    if (alambda[KSLOWP] < 0.5 * delta) alambda[KSLOWP] = lambda[KSLOWP] * lambda[KSLOWP] / delta + 0.25 * delta;
  }
  for (int n1 = 0; n1 < nv_; n1++) {
    double sc = 0.0;
    for (int n2 = 0; n2 < nv_; n2++) sc += alambda[n2] * eta[n2] * Rc[n1][n2];
    flux[n1] = 0.5 * (fluxL[n1] + fluxR[n1] - sc);
  }
  store_face_emf(c, V, D, dir, vL, vR, uL, uR, flux, sl, sr, pL, pR, E);
}

// --------------------------------------------------------------------------------------
// The solver state: arrays of src/fluid/fluid.hpp:529-560 and constrainedTransport.hpp
// --------------------------------------------------------------------------------------
struct Oracle {
  orc_cfg c;
  Vars V;
  int np_tot[3], np_int[3], nghost[3], beg[3], end[3];
  int ntot;  // fluid variables + passive tracers
  std::vector<unsigned char> shock;  // FlagShock per cell (shockFlattening.hpp:14)
  // isothermal sound speed per cell ("csiso userdef", eos_isothermal.hpp:29-33, :70-75);
  // empty: the constant c.csiso
  std::vector<double> cs_cell;
  int offs[3];  // IOFFSET, JOFFSET, KOFFSET (macros.hpp:31-53)
  int vsn[3];   // Vs extents (i, j, k)
  std::vector<double> x[3], xl[3], xr[3], dx[3];
  std::vector<double> Vc, Uc, Vs, Flux, InvDt, cMax;
  std::vector<double> Uc0, Vs0;  // state "begin" (timeIntegrator.cpp:77-80)
  std::vector<double> ex, ey, ez, ezi, ezj, exj, exk, eyi, eyk, svx, svy, svz;
  std::vector<double> axL, axR, ayL, ayR, azL, azR, dxL, dxR, dyL, dyR, dzL, dzR;
  std::vector<double> Ex1, Ex2, Ex3;
  double t = 0, dt = 0;
  long ncycles = 0;
  double wc[2] = {0, 0}, w0[2] = {0, 0};
  bool last_flux_unwritten = false;
  // threads only pay off on blocks of some size: small (1-D) blocks run serially, which
  // also keeps them fast when other processes compete for the cores
  bool par_ = true;
  orc_userdef_fn user_fn = nullptr;
  void *user_arg = nullptr;

  explicit Oracle(const orc_cfg &cfg) : c(cfg), V(cfg) {
    for (int d = 0; d < 3; d++) {
      if (d < c.dim) {
        nghost[d] = (c.order < 4) ? 2 : 3;  // grid.cpp:60-65
        np_int[d] = c.nx[d];
      } else {
        nghost[d] = 0;
        np_int[d] = 1;
      }
      np_tot[d] = np_int[d] + 2 * nghost[d];
      beg[d] = nghost[d];
      end[d] = nghost[d] + np_int[d];
      offs[d] = (d < c.dim) ? 1 : 0;
      vsn[d] = np_tot[d] + offs[d];
      x[d].resize(np_tot[d]);
      xl[d].resize(np_tot[d]);
      xr[d].resize(np_tot[d]);
      dx[d].resize(np_tot[d]);
      // GridHost::MakeGrid uniform patch, gridHost.cpp:79-86; unused dims :200-221
      for (int i = 0; i < np_tot[d]; i++) {
        if (d < c.dim) {
          const double ps = c.xbeg[d], pe = c.xend[d];
          const int ii = i - nghost[d];
          dx[d][i] = (pe - ps) / (np_int[d]);
          x[d][i] = ps + (ii + 0.5) * dx[d][i];
          xl[d][i] = ps + ii * dx[d][i];
          xr[d][i] = ps + (ii + 1) * dx[d][i];
        } else {
          dx[d][i] = 0.5 - (-0.5);
          x[d][i] = 0.5 * (-0.5 + 0.5);
          xl[d][i] = -0.5;
          xr[d][i] = 0.5;
        }
      }
    }
    const size_t n3 = (size_t)np_tot[0] * np_tot[1] * np_tot[2];
    par_ = n3 >= 16384;
    // passive tracers follow the fluid variables (fluid.hpp:529-545)
    ntot = V.nvar + c.ntracer;
    Vc.assign(ntot * n3, 0.0);
    Uc.assign(ntot * n3, 0.0);
    Uc0.assign(ntot * n3, 0.0);
    Flux.assign(ntot * n3, 0.0);
    InvDt.assign(n3, 0.0);
    cMax.assign(n3, 0.0);
    shock.assign(n3, 0);
    if (c.mhd) {
      Vs.assign((size_t)c.dim * vsn[0] * vsn[1] * vsn[2], 0.0);
      Vs0.assign(Vs.size(), 0.0);
    }
    for (auto *a : {&ex, &ey, &ez, &ezi, &ezj, &exj, &exk, &eyi, &eyk, &svx, &svy, &svz,
                    &axL, &axR, &ayL, &ayR, &azL, &azR, &dxL, &dxR, &dyL, &dyR, &dzL,
                    &dzR, &Ex1, &Ex2, &Ex3})
      a->assign(n3, 0.0);
    if (c.nstages == 2) { wc[0] = 0.5; w0[0] = 0.5; }
    if (c.nstages == 3) { wc[0] = 0.25; w0[0] = 0.75; wc[1] = 2.0 / 3.0; w0[1] = 1.0 / 3.0; }
    dt = c.first_dt;
  }

  // index helpers (LayoutRight)
  inline size_t i3(int k, int j, int i) const {
    return ((size_t)k * np_tot[1] + j) * np_tot[0] + i;
  }
  inline size_t i4(int n, int k, int j, int i) const {
    return (((size_t)n * np_tot[2] + k) * np_tot[1] + j) * np_tot[0] + i;
  }
  inline size_t is(int n, int k, int j, int i) const {
    return (((size_t)n * vsn[2] + k) * vsn[1] + j) * vsn[0] + i;
  }
  double &vc(int n, int k, int j, int i) { return Vc[i4(n, k, j, i)]; }
  double &uc(int n, int k, int j, int i) { return Uc[i4(n, k, j, i)]; }
  double &vs(int n, int k, int j, int i) { return Vs[is(n, k, j, i)]; }

  // Cartesian geometry, makeGeometry.cpp:67-70, :146-147, :178-179, :201-202
  double dV(int k, int j, int i) const {
    double v = dx[0][i];
    if (c.dim >= 2) v = v * dx[1][j];
    if (c.dim >= 3) v = v * dx[2][k];
    return v;
  }
  double Ax(int dir, int k, int j, int i) const {
    // extents of the A arrays are np_tot (+1 along their own direction); the
    // transverse coordinates only index dx of the other directions.
    double a;
    if (dir == IDIR) {
      a = 1.0;
      if (c.dim >= 2) a = a * dx[1][j];
      if (c.dim >= 3) a = a * dx[2][k];
    } else if (dir == JDIR) {
      a = dx[0][i];
      if (c.dim >= 2) a = a * 1.0;
      if (c.dim >= 3) a = a * dx[2][k];
    } else {
      a = dx[0][i];
      if (c.dim >= 2) a = a * dx[1][j];
      if (c.dim >= 3) a = a * 1.0;
    }
    return a;
  }

  // ---------------- Boundary (src/fluid/boundary/boundary.hpp) ----------------
  // BoundaryForAll / BoundaryForX?s loop ranges, boundary.hpp:969-1099
  void side_range(int dir, int side, int facedir, int lo[3], int hi[3]) const {
    for (int d = 0; d < 3; d++) {
      int nxd = np_int[d] + ((facedir == d) ? 1 : 0);
      if (d == dir) {
        lo[d] = side * (nghost[d] + nxd);
        hi[d] = nghost[d] + side * (nghost[d] + nxd);
      } else {
        lo[d] = 0;
        hi[d] = np_tot[d] + ((facedir == d) ? 1 : 0);
      }
    }
  }
  // EnforcePeriodic, boundary.hpp:576-673
  void enforce_periodic(int dir, int side) {
    int lo[3], hi[3];
    side_range(dir, side, -1, lo, hi);
    auto ref = [&](int d, int q) {
      return (d == dir) ? nghost[d] + (q + nghost[d] * (np_int[d] - 1)) % np_int[d] : q;
    };
    for (int n = 0; n < ntot; n++)  // boundary.hpp:135-136: tracers included
      for (int k = lo[2]; k < hi[2]; k++)
        for (int j = lo[1]; j < hi[1]; j++)
          for (int i = lo[0]; i < hi[0]; i++)
            vc(n, k, j, i) = vc(n, ref(2, k), ref(1, j), ref(0, i));
    if (c.mhd) {
      for (int f = 0; f < c.dim; f++) {
        side_range(dir, side, f, lo, hi);
        for (int k = lo[2]; k < hi[2]; k++)
          for (int j = lo[1]; j < hi[1]; j++)
            for (int i = lo[0]; i < hi[0]; i++)
              vs(f, k, j, i) = vs(f, ref(2, k), ref(1, j), ref(0, i));
      }
    }
  }
  // EnforceOutflow, boundary.hpp:739-809
  void enforce_outflow(int dir, int side) {
    int lo[3], hi[3];
    side_range(dir, side, -1, lo, hi);
    auto ref = [&](int d, int q) { return (d == dir) ? nghost[d] + side * (np_int[d] - 1) : q; };
    const int sign = 1 - 2 * side;
    for (int n = 0; n < ntot; n++)
      for (int k = lo[2]; k < hi[2]; k++)
        for (int j = lo[1]; j < hi[1]; j++)
          for (int i = lo[0]; i < hi[0]; i++) {
            const double r = vc(n, ref(2, k), ref(1, j), ref(0, i));
            if ((n == V.MX1 + dir) && (sign * r >= 0.0))
              vc(n, k, j, i) = 0.0;
            else
              vc(n, k, j, i) = r;
          }
    if (c.mhd) {
      for (int f = 0; f < c.dim; f++) {
        if (f == dir) continue;  // normal component: ReconstructNormalField
        side_range(dir, side, f, lo, hi);
        for (int k = lo[2]; k < hi[2]; k++)
          for (int j = lo[1]; j < hi[1]; j++)
            for (int i = lo[0]; i < hi[0]; i++)
              vs(f, k, j, i) = vs(f, ref(2, k), ref(1, j), ref(0, i));
      }
    }
  }
  // EnforceReflective, boundary.hpp:676-737
  void enforce_reflective(int dir, int side) {
    int lo[3], hi[3];
    side_range(dir, side, -1, lo, hi);
    auto ref = [&](int d, int q) {
      return (d == dir) ? 2 * (nghost[d] + side * np_int[d]) - q - 1 : q;
    };
    // with MHD the sign test n >= BX1 also flips tracers (boundary.hpp:697)
    for (int n = 0; n < ntot; n++) {
      const double sign =
          (n == V.MX1 + dir || (V.BX1 >= 0 && n >= V.BX1 && n != V.BX1 + dir)) ? -1.0 : 1.0;
      for (int k = lo[2]; k < hi[2]; k++)
        for (int j = lo[1]; j < hi[1]; j++)
          for (int i = lo[0]; i < hi[0]; i++)
            vc(n, k, j, i) = sign * vc(n, ref(2, k), ref(1, j), ref(0, i));
    }
    if (c.mhd) {
      for (int f = 0; f < c.dim; f++) {
        if (f == dir) continue;
        side_range(dir, side, f, lo, hi);
        for (int k = lo[2]; k < hi[2]; k++)
          for (int j = lo[1]; j < hi[1]; j++)
            for (int i = lo[0]; i < hi[0]; i++)
              vs(f, k, j, i) = -vs(f, ref(2, k), ref(1, j), ref(0, i));
      }
    }
  }
  // ReconstructNormalField, boundary.hpp:407-530 (Cartesian, no axis).  The
  // transverse divergence keeps the reference's association per direction:
  // IDIR/KDIR ((P-Q)+R)-S, JDIR (P-Q)+(R-S) (signLeft/signRight = 1).
  void reconstruct_normal_field(int dir) {
    const bool recL = !(c.lbound[dir] == ORC_BC_PERIODIC);
    const bool recR = !(c.rbound[dir] == ORC_BC_PERIODIC);
    if (!recL && !recR) return;
    const int nstart = beg[dir] - 1;
    const int nend = end[dir];
    const int nd = np_tot[dir];
    auto PQ = [&](int f, int k, int j, int i) {
      const int kp = k + (f == 2), jp = j + (f == 1), ip = i + (f == 0);
      return Ax(f, kp, jp, ip) * vs(f, kp, jp, ip) - Ax(f, k, j, i) * vs(f, k, j, i);
    };
    auto trans = [&](int k, int j, int i) {
      if (dir == IDIR) {
        double s = 0.0;
        if (c.dim >= 2) {
          s = s + Ax(1, k, j + 1, i) * vs(1, k, j + 1, i);
          s = s - Ax(1, k, j, i) * vs(1, k, j, i);
        }
        if (c.dim >= 3) {
          s = s + Ax(2, k + 1, j, i) * vs(2, k + 1, j, i);
          s = s - Ax(2, k, j, i) * vs(2, k, j, i);
        }
        return s;
      } else if (dir == JDIR) {
        double s = PQ(0, k, j, i);
        if (c.dim >= 3) s = s + 1 * PQ(2, k, j, i);
        return s;
      } else {
        double s = Ax(0, k, j, i + 1) * vs(0, k, j, i + 1) - Ax(0, k, j, i) * vs(0, k, j, i);
        s = s + Ax(1, k, j + 1, i) * vs(1, k, j + 1, i);
        s = s - Ax(1, k, j, i) * vs(1, k, j, i);
        return s;
      }
    };
    const int o1 = (dir == 0) ? 1 : 0;  // the two other directions
    const int o2 = (dir == 2) ? 1 : 2;
    for (int q2 = 0; q2 < np_tot[o2]; q2++)
      for (int q1 = 0; q1 < np_tot[o1]; q1++) {
        int idx[3];
        idx[o1] = q1;
        idx[o2] = q2;
        if (recL) {
          for (int q = nstart; q >= 0; q--) {
            idx[dir] = q;
            const int k = idx[2], j = idx[1], i = idx[0];
            const int kp = k + (dir == 2), jp = j + (dir == 1), ip = i + (dir == 0);
            vs(dir, k, j, i) = 1.0 / Ax(dir, k, j, i) *
                               (Ax(dir, kp, jp, ip) * vs(dir, kp, jp, ip) + (trans(k, j, i)));
          }
        }
        if (recR) {
          for (int q = nend; q < nd; q++) {
            idx[dir] = q;
            const int k = idx[2], j = idx[1], i = idx[0];
            const int kp = k + (dir == 2), jp = j + (dir == 1), ip = i + (dir == 0);
            vs(dir, kp, jp, ip) = 1.0 / Ax(dir, kp, jp, ip) *
                                  (Ax(dir, k, j, i) * vs(dir, k, j, i) - (trans(k, j, i)));
          }
        }
      }
  }
  // ReconstructVcField, boundary.hpp:388-404 (applied to Vc or Uc)
  void reconstruct_vc_field(std::vector<double> &X) {
    if (!c.mhd) return;
#pragma omp parallel for collapse(2) if (par_)
    for (int k = 0; k < np_tot[2]; k++)
      for (int j = 0; j < np_tot[1]; j++)
        for (int i = 0; i < np_tot[0]; i++) {
          X[i4(V.BX1, k, j, i)] = 0.5 * (vs(0, k, j, i) + vs(0, k, j, i + 1));
          if (c.dim >= 2) X[i4(V.BX2, k, j, i)] = 0.5 * (vs(1, k, j, i) + vs(1, k, j + 1, i));
          if (c.dim >= 3) X[i4(V.BX3, k, j, i)] = 0.5 * (vs(2, k, j, i) + vs(2, k + 1, j, i));
        }
  }
  // SetBoundaries, boundary.hpp:239-282
  void set_boundaries() {
    for (int dir = 0; dir < c.dim; dir++) {
      for (int side = 0; side < 2; side++) {
        const int bc = side == 0 ? c.lbound[dir] : c.rbound[dir];
        if (bc == ORC_BC_PERIODIC) enforce_periodic(dir, side);
        else if (bc == ORC_BC_OUTFLOW) enforce_outflow(dir, side);
        else if (bc == ORC_BC_REFLECTIVE) enforce_reflective(dir, side);
        else if (bc == ORC_BC_USERDEF && user_fn) user_fn(user_arg, dir, side, t);
      }
      if (c.mhd) reconstruct_normal_field(dir);
    }
    if (c.mhd) reconstruct_vc_field(Vc);
  }

  // ---------------- Pointwise conversions over all cells ----------------
  // ConvertConsToPrim, convertConsToPrim.hpp:105-149
  void cons_to_prim() {
    if (c.mhd) reconstruct_vc_field(Uc);
#pragma omp parallel for collapse(2) if (par_)
    for (int k = 0; k < np_tot[2]; k++)
      for (int j = 0; j < np_tot[1]; j++)
        for (int i = 0; i < np_tot[0]; i++) {
          double u[8], v[8];
          for (int n = 0; n < V.nvar; n++) u[n] = uc(n, k, j, i);
          cons2prim(c, V, u, v);
          for (int n = 0; n < V.nvar; n++) vc(n, k, j, i) = v[n];
          // Tracer::ConvertConsToPrim, tracer.cpp:13-28
          for (int n = V.nvar; n < ntot; n++) vc(n, k, j, i) = uc(n, k, j, i) / uc(V.RHO, k, j, i);
        }
  }
  // ConvertPrimToCons, convertConsToPrim.hpp:152-189
  void prim_to_cons() {
#pragma omp parallel for collapse(2) if (par_)
    for (int k = 0; k < np_tot[2]; k++)
      for (int j = 0; j < np_tot[1]; j++)
        for (int i = 0; i < np_tot[0]; i++) {
          double u[8], v[8];
          for (int n = 0; n < V.nvar; n++) v[n] = vc(n, k, j, i);
          prim2cons(c, V, v, u);
          for (int n = 0; n < V.nvar; n++) uc(n, k, j, i) = u[n];
          // Tracer::ConvertPrimToCons, tracer.cpp:32-47
          for (int n = V.nvar; n < ntot; n++) uc(n, k, j, i) = vc(n, k, j, i) * vc(V.RHO, k, j, i);
        }
  }

  // ---------------- Shock flattening (shockFlattening.hpp:62-168) ----------------
  bool flagged(int dir, int k, int j, int i, int s) const {
    if (!c.shock_flattening) return false;
    const int io = dir == IDIR, jo = dir == JDIR, ko = dir == KDIR;
    return shock[i3(k + s * ko, j + s * jo, i + s * io)] != 0;
  }
  // ShockFlattening::FindShock, Cartesian with an energy equation: a cell is a shock
  // where the flow converges and the pressure jump across it exceeds smoothing times the
  // smallest neighbouring pressure
  void find_shock() {
    const int ko = c.dim == 3 ? 1 : 0, jo = c.dim >= 2 ? 1 : 0;
#pragma omp parallel for collapse(2) if (par_)
    for (int k = beg[2] - ko; k < end[2] + ko; k++)
      for (int j = beg[1] - jo; j < end[1] + jo; j++)
        for (int i = beg[0] - 1; i < end[0] + 1; i++) {
          unsigned char f = 0;
          double divV = (vc(V.MX1, k, j, i + 1) - vc(V.MX1, k, j, i - 1)) / (2 * dx[0][i]);
          if (c.dim >= 2)
            divV = divV + (vc(V.MX2, k, j + 1, i) - vc(V.MX2, k, j - 1, i)) / (2 * dx[1][j]);
          if (c.dim == 3)
            divV = divV + (vc(V.MX3, k + 1, j, i) - vc(V.MX3, k - 1, j, i)) / (2 * dx[2][k]);
          if (divV < 0.0) {
            // pressure of a cell: PRS, or rho cs^2 when isothermal (shockFlattening.hpp:92-128)
            auto P = [&](int kk, int jj, int ii) {
              if (c.isothermal) {
                const double cs = cs_cell.empty() ? c.csiso : cs_cell[i3(kk, jj, ii)];
                return vc(V.RHO, kk, jj, ii) * cs * cs;
              }
              return vc(V.ENG, kk, jj, ii);
            };
            double pmin = P(k, j, i);
            const double pR = P(k, j, i + 1), pL = P(k, j, i - 1);
            pmin = FMIN(pmin, c.isothermal ? pL : pR);
            pmin = FMIN(pmin, c.isothermal ? pR : pL);
            double gradP = std::fabs(pR - pL);
            if (c.dim >= 2) {
              const double qR = P(k, j + 1, i), qL = P(k, j - 1, i);
              pmin = FMIN(pmin, qL);
              pmin = FMIN(pmin, qR);
              gradP += std::fabs(qR - qL);
            }
            if (c.dim == 3) {
              const double qR = P(k + 1, j, i), qL = P(k - 1, j, i);
              pmin = FMIN(pmin, qL);
              pmin = FMIN(pmin, qR);
              gradP += std::fabs(qR - qL);
            }
            if (gradP > c.shock_smoothing * pmin) f = 1;
          }
          shock[i3(k, j, i)] = f;
        }
  }

  // ---------------- Reconstruction (extrapolateToFaces.hpp:75-309) ----------------
  void extrapolate(int dir, int k, int j, int i, double *vL, double *vR) {
    const int io = dir == IDIR, jo = dir == JDIR, ko = dir == KDIR;
    for (int nv = 0; nv < V.nvar; nv++) {
      auto q = [&](int s) { return vc(nv, k + s * ko, j + s * jo, i + s * io); };
      if (c.order == 1) {
        vL[nv] = q(-1);
        vR[nv] = q(0);
      } else if (c.order == 2) {
        // extrapolateToFaces.hpp:89-124: a flagged cell takes MinModLim, which as built
        // returns 0
        double dvm = q(-1) - q(-2);
        double dvp = q(0) - q(-1);
        double dv = flagged(dir, k, j, i, -1) ? minmod_as_built(dvp, dvm) : vanleer(dvp, dvm);
        vL[nv] = q(-1) + 0.5 * dv;
        dvm = dvp;
        dvp = q(1) - q(0);
        dv = flagged(dir, k, j, i, 0) ? minmod_as_built(dvp, dvm) : vanleer(dvp, dvm);
        vR[nv] = q(0) - 0.5 * dv;
      } else if (c.order == 3) {
        // LimO3, extrapolateToFaces.hpp:170-246 (no shock flattening)
        const int idx = io * i + jo * j + ko * k;
        double dvm = q(-1) - q(-2);
        double dvp = q(0) - q(-1);
        double dv = flagged(dir, k, j, i, -1) ? minmod_as_built(dvp, dvm)
                                              : dvp * limo3lim(dvp, dvm, dx[dir][idx - 1]);
        vL[nv] = q(-1) + 0.5 * dv;
        if ((nv == V.RHO || nv == V.ENG) && vL[nv] <= 0.0) {
          dv = minmod_as_built(dvp, dvm);
          vL[nv] = q(-1) + 0.5 * dv;
        }
        dvm = dvp;
        dvp = q(1) - q(0);
        dv = flagged(dir, k, j, i, 0) ? minmod_as_built(dvp, dvm)
                                      : dvm * limo3lim(dvm, dvp, dx[dir][idx]);
        vR[nv] = q(0) - 0.5 * dv;
        if ((nv == V.RHO || nv == V.ENG) && vR[nv] <= 0.0) {
          dv = minmod_as_built(dvp, dvm);
          vR[nv] = q(0) - 0.5 * dv;
        }
      } else {
        double vm2 = q(-3), vm1 = q(-2), v0 = q(-1), vp1 = q(0), vp2 = q(1);
        double vr, vl;
        ppm_states(vm2, vm1, v0, vp1, vp2, vl, vr);
        if (nv == V.RHO || nv == V.ENG) {
          if (vr <= 0.0) {
            double dv = vanleer(vp1 - v0, v0 - vm1);
            vr = v0 + 0.5 * dv;
          }
        }
        vL[nv] = vr;
        vm2 = vm1;
        vm1 = v0;
        v0 = vp1;
        vp1 = vp2;
        vp2 = q(2);
        ppm_states(vm2, vm1, v0, vp1, vp2, vl, vr);
        if (nv == V.RHO || nv == V.ENG) {
          if (vl <= 0.0) {
            double dv = vanleer(vp1 - v0, v0 - vm1);
            vl = v0 - 0.5 * dv;
          }
        }
        vR[nv] = vl;
      }
    }
  }

  // ---------------- Riemann fluxes (calcFlux.hpp:27-85) ----------------
  std::vector<double> *et_arr(int dir) { return dir == 0 ? &ezi : dir == 1 ? &ezj : &eyk; }
  std::vector<double> *eb_arr(int dir) { return dir == 0 ? &eyi : dir == 1 ? &exj : &exk; }
  std::vector<double> *sv_arr(int dir) { return dir == 0 ? &svx : dir == 1 ? &svy : &svz; }
  std::vector<double> *aL_arr(int dir) { return dir == 0 ? &axL : dir == 1 ? &ayL : &azL; }
  std::vector<double> *aR_arr(int dir) { return dir == 0 ? &axR : dir == 1 ? &ayR : &azR; }
  std::vector<double> *dL_arr(int dir) { return dir == 0 ? &dxL : dir == 1 ? &dyL : &dzL; }
  std::vector<double> *dR_arr(int dir) { return dir == 0 ? &dxR : dir == 1 ? &dyR : &dzR; }

  void calc_flux(int dir) {
    // RiemannSolver::CalcFlux<IDIR> refreshes the shock flags (calcFlux.hpp:30-33)
    if (dir == IDIR && c.shock_flattening) find_shock();
    const int io = dir == IDIR, jo = dir == JDIR, ko = dir == KDIR;
    int ext[3] = {0, 0, 0};
    if (c.mhd) {
      // hlldMHD.hpp:32-49 and hllMHD.hpp:32-49 extend by nghost for the uct_hll(d)
      // averages; RoeMHD and TvdlfMHD always extend by one cell (roeMHD.hpp:62-73,
      // tvdlfMHD.hpp:30-41)
      int perp = 1;
      if ((c.emf == ORC_EMF_UCT_HLL || c.emf == ORC_EMF_UCT_HLLD) &&
          (c.solver == ORC_SOLVER_HLLD || c.solver == ORC_SOLVER_HLL))
        perp = nghost[dir];
      for (int d = 0; d < 3; d++) ext[d] = (d == dir || d >= c.dim) ? 0 : perp;
    }
    const Dir D(V, dir);
    last_flux_unwritten = false;
    bool unwritten = false;
#pragma omp parallel for collapse(2) reduction(|| : unwritten) if (par_)
    for (int k = beg[2] - ext[2]; k < end[2] + ko + ext[2]; k++)
      for (int j = beg[1] - ext[1]; j < end[1] + jo + ext[1]; j++)
        for (int i = beg[0] - ext[0]; i < end[0] + io + ext[0]; i++) {
          double vL[8], vR[8], F[8], cmax;
          FaceEmf E;
          E.hasEt = E.hasEb = E.hasSv = E.hasAD = false;
          extrapolate(dir, k, j, i, vL, vR);
          // a per-cell sound speed enters as the face average of the two cells'
          // (hllcHD.hpp:67-70, hlldMHD.hpp:188-191): a configuration whose constant is
          // that average, since every solver takes 0.5 (cs + cs) = cs of it
          orc_cfg cface;
          const orc_cfg *cf = &c;
          if (!cs_cell.empty()) {
            cface = c;
            cface.csiso = 0.5 * (cs_cell[i3(k, j, i)] + cs_cell[i3(k - ko, j - jo, i - io)]);
            cf = &cface;
          }
          bool wrote = true;
          if (c.mhd) {
            vL[D.BXn] = vs(dir, k, j, i);
            vR[D.BXn] = vL[D.BXn];
            if (c.solver == ORC_SOLVER_ROE) roe_mhd(*cf, V, D, dir, vL, vR, F, cmax, E);
            else if (c.solver == ORC_SOLVER_HLL) hll_mhd(*cf, V, D, dir, vL, vR, F, cmax, E);
            else if (c.solver == ORC_SOLVER_TVDLF) tvdlf_mhd(*cf, V, D, dir, vL, vR, F, cmax, E);
            else hlld_mhd(*cf, V, D, dir, vL, vR, F, cmax, E);
          } else if (c.solver == ORC_SOLVER_HLLC) {
            hllc_hd(*cf, V, D, vL, vR, F, cmax);
          } else if (c.solver == ORC_SOLVER_HLL) {
            hll_hd(*cf, V, D, vL, vR, F, cmax);
          } else if (c.solver == ORC_SOLVER_TVDLF) {
            tvdlf_hd(*cf, V, D, vL, vR, F, cmax);
          } else {
            wrote = roe_hd(*cf, V, D, vL, vR, F, cmax);
          }
          if (wrote) {
            for (int n = 0; n < V.nvar; n++) Flux[i4(n, k, j, i)] = F[n];
          } else {
            unwritten = true;
          }
          const size_t p = i3(k, j, i);
          cMax[p] = cmax;
          if (E.hasEt) (*et_arr(dir))[p] = E.Et;
          if (E.hasEb) (*eb_arr(dir))[p] = E.Eb;
          if (E.hasSv) (*sv_arr(dir))[p] = E.sv;
          if (E.hasAD) {
            (*aL_arr(dir))[p] = E.aL;
            (*aR_arr(dir))[p] = E.aR;
            (*dL_arr(dir))[p] = E.dL;
            (*dR_arr(dir))[p] = E.dR;
          }
        }
    last_flux_unwritten = unwritten;
  }

  // ---------------- CalcRightHandSide (calcRightHandSide.hpp:556-596) ----------------
  // Tracer::CalcFlux, tracer.hpp:58-87: face value upwinded by the sign of the Riemann
  // mass flux, van Leer slope; flux = F_rho * vface * A
  void tracer_flux(int dir) {
    const int io = dir == IDIR, jo = dir == JDIR, ko = dir == KDIR;
    for (int nv = V.nvar; nv < ntot; nv++)
#pragma omp parallel for collapse(2) if (par_)
      for (int k = beg[2]; k < end[2] + ko; k++)
        for (int j = beg[1]; j < end[1] + jo; j++)
          for (int i = beg[0]; i < end[0] + io; i++) {
            auto q = [&](int s) { return vc(nv, k + s * ko, j + s * jo, i + s * io); };
            const double frho = Flux[i4(V.RHO, k, j, i)];
            double vface;
            if (frho > 0) {
              const double dvm = q(-1) - q(-2);
              const double dvp = q(0) - q(-1);
              const double dv = vanleer(dvp, dvm);
              vface = q(-1) + 0.5 * dv;
            } else {
              const double dvm = q(0) - q(-1);
              const double dvp = q(1) - q(0);
              const double dv = vanleer(dvp, dvm);
              vface = q(0) - 0.5 * dv;
            }
            Flux[i4(nv, k, j, i)] = frho * vface * Ax(dir, k, j, i);
          }
  }
  // Tracer::CalcRightHandSide, tracer.hpp:89-111
  void tracer_rhs(int dir, double dtl) {
    const int io = dir == IDIR, jo = dir == JDIR, ko = dir == KDIR;
    for (int nv = V.nvar; nv < ntot; nv++)
#pragma omp parallel for collapse(2) if (par_)
      for (int k = beg[2]; k < end[2]; k++)
        for (int j = beg[1]; j < end[1]; j++)
          for (int i = beg[0]; i < end[0]; i++)
            uc(nv, k, j, i) = uc(nv, k, j, i) + -dtl / dV(k, j, i) *
                              (Flux[i4(nv, k + ko, j + jo, i + io)] - Flux[i4(nv, k, j, i)]);
  }
  void calc_rhs(int dir, double dtl) {
    const int io = dir == IDIR, jo = dir == JDIR, ko = dir == KDIR;
    // "Correct Flux": Flux *= A (Cartesian, no Fargo), :156-194
#pragma omp parallel for collapse(2) if (par_)
    for (int k = beg[2]; k < end[2] + ko; k++)
      for (int j = beg[1]; j < end[1] + jo; j++)
        for (int i = beg[0]; i < end[0] + io; i++) {
          const double a = Ax(dir, k, j, i);
          for (int n = 0; n < V.nvar; n++) Flux[i4(n, k, j, i)] = Flux[i4(n, k, j, i)] * a;
        }
    // "CalcRightHandSide", :288-549
#pragma omp parallel for collapse(2) if (par_)
    for (int k = beg[2]; k < end[2]; k++)
      for (int j = beg[1]; j < end[1]; j++)
        for (int i = beg[0]; i < end[0]; i++) {
          const double dtdV = dtl / dV(k, j, i);
          double rhs[8];
          for (int n = 0; n < V.nvar; n++)
            rhs[n] = -dtdV * (Flux[i4(n, k + ko, j + jo, i + io)] - Flux[i4(n, k, j, i)]);
          const int ig = io * i + jo * j + ko * k;
          const double dl = dx[dir][ig];
          InvDt[i3(k, j, i)] =
              InvDt[i3(k, j, i)] + 0.5 * (cMax[i3(k + ko, j + jo, i + io)] + cMax[i3(k, j, i)]) / (dl);
          for (int n = 0; n < V.nvar; n++) {
            if (c.mhd) {
              if (n == V.BX1) continue;
              if (c.dim >= 2 && n == V.BX2) continue;
              if (c.dim >= 3 && n == V.BX3) continue;
            }
            uc(n, k, j, i) = uc(n, k, j, i) + rhs[n];
          }
        }
  }

  // ---------------- Constrained transport ----------------
  // CalcCellCenteredEMF, calcCornerEmf.hpp:84-126
  void calc_cell_emf() {
#pragma omp parallel for collapse(2) if (par_)
    for (int k = 0; k < np_tot[2]; k++)
      for (int j = 0; j < np_tot[1]; j++)
        for (int i = 0; i < np_tot[0]; i++) {
          double vx1 = vc(V.MX1, k, j, i);
          double vx2 = c.comps >= 2 ? vc(V.MX2, k, j, i) : 0.0;
          double vx3 = c.comps >= 3 ? vc(V.MX3, k, j, i) : 0.0;
          double bx1 = vc(V.BX1, k, j, i);
          double bx2 = c.comps >= 2 ? vc(V.BX2, k, j, i) : 0.0;
          double bx3 = c.comps >= 3 ? vc(V.BX3, k, j, i) : 0.0;
          const size_t p = i3(k, j, i);
          if (c.dim == 3) {
            Ex1[p] = (vx3 * bx2 - vx2 * bx3);
            Ex2[p] = (vx1 * bx3 - vx3 * bx1);
          }
          Ex3[p] = (vx2 * bx1 - vx1 * bx2);
        }
  }
  // CalcArithmeticAverage / CalcUCT0CornerEMF corner average, calcCornerEmf.hpp:62-78
  void arithmetic_corner() {
    const double w = 0.25;
#pragma omp parallel for collapse(2) if (par_)
    for (int k = beg[2]; k < end[2] + offs[2]; k++)
      for (int j = beg[1]; j < end[1] + offs[1]; j++)
        for (int i = beg[0]; i < end[0] + offs[0]; i++) {
          const size_t p = i3(k, j, i);
          if (c.dim == 3) {
            ex[p] = w * (exj[p] + exj[i3(k - 1, j, i)] + exk[p] + exk[i3(k, j - 1, i)]);
            ey[p] = w * (eyi[p] + eyi[i3(k - 1, j, i)] + eyk[p] + eyk[i3(k, j, i - 1)]);
          }
          ez[p] = w * (ezi[p] + ezi[i3(k, j - 1, i)] + ezj[p] + ezj[i3(k, j, i - 1)]);
        }
  }
  // CalcUCT0Average, calcCornerEmf.hpp:128-194
  void uct0_average() {
#pragma omp parallel for collapse(2) if (par_)
    for (int k = beg[2] - offs[2]; k < end[2] + offs[2]; k++)
      for (int j = beg[1] - offs[1]; j < end[1] + offs[1]; j++)
        for (int i = beg[0] - offs[0]; i < end[0] + offs[0]; i++) {
          const size_t p = i3(k, j, i);
          if (c.dim == 3) {
            exj[p] *= 2.0;
            exk[p] *= 2.0;
            eyi[p] *= 2.0;
            eyk[p] *= 2.0;
            exj[p] -= 0.5 * (Ex1[i3(k, j - 1, i)] + Ex1[p]);
            exk[p] -= 0.5 * (Ex1[i3(k - 1, j, i)] + Ex1[p]);
            eyi[p] -= 0.5 * (Ex2[i3(k, j, i - 1)] + Ex2[p]);
            eyk[p] -= 0.5 * (Ex2[i3(k - 1, j, i)] + Ex2[p]);
          }
          ezi[p] *= 2.0;
          ezj[p] *= 2.0;
          if (g_plant == ORC_PLANT_UCT0_INDEX) ezi[p] -= 0.5 * (Ex3[i3(k, j - 1, i)] + Ex3[p]);
          else ezi[p] -= 0.5 * (Ex3[i3(k, j, i - 1)] + Ex3[p]);
          ezj[p] -= 0.5 * (Ex3[i3(k, j - 1, i)] + Ex3[p]);
        }
    arithmetic_corner();
  }
  // CalcContactAverage, calcCornerEmf.hpp:196-282
  void contact_average() {
#pragma omp parallel for collapse(2) if (par_)
    for (int k = beg[2]; k < end[2] + offs[2]; k++)
      for (int j = beg[1]; j < end[1] + offs[1]; j++)
        for (int i = beg[0]; i < end[0] + offs[0]; i++) {
          auto A = [&](std::vector<double> &a, int kk, int jj, int ii) {
            const double v = a[i3(kk, jj, ii)];
            const bool sv = &a == &svx || &a == &svy || &a == &svz;
            return (sv && g_plant == ORC_PLANT_CONTACT_SV) ? 1.0 - v : v;
          };
          const double ez_l2 = (1 - A(svx, k, j - 1, i)) * (A(ezj, k, j, i) - A(Ex3, k, j - 1, i)) +
                               (A(svx, k, j - 1, i)) * (A(ezj, k, j, i - 1) - A(Ex3, k, j - 1, i - 1));
          const double ez_r2 = (1 - A(svx, k, j, i)) * (A(ezj, k, j, i) - A(Ex3, k, j, i)) +
                               (A(svx, k, j, i)) * (A(ezj, k, j, i - 1) - A(Ex3, k, j, i - 1));
          const double ez_l1 = (1 - A(svy, k, j, i - 1)) * (A(ezi, k, j, i) - A(Ex3, k, j, i - 1)) +
                               (A(svy, k, j, i - 1)) * (A(ezi, k, j - 1, i) - A(Ex3, k, j - 1, i - 1));
          const double ez_r1 = (1 - A(svy, k, j, i)) * (A(ezi, k, j, i) - A(Ex3, k, j, i)) +
                               (A(svy, k, j, i)) * (A(ezi, k, j - 1, i) - A(Ex3, k, j - 1, i));
          ez[i3(k, j, i)] = 0.25 * (ez_l2 + ez_r2 + ez_l1 + ez_r1 + A(ezi, k, j, i) +
                                    A(ezi, k, j - 1, i) + A(ezj, k, j, i) + A(ezj, k, j, i - 1));
          if (c.dim == 3) {
            const double ex_l3 = (1 - A(svy, k - 1, j, i)) * (A(exk, k, j, i) - A(Ex1, k - 1, j, i)) +
                                 (A(svy, k - 1, j, i)) * (A(exk, k, j - 1, i) - A(Ex1, k - 1, j - 1, i));
            const double ex_r3 = (1 - A(svy, k, j, i)) * (A(exk, k, j, i) - A(Ex1, k, j, i)) +
                                 (A(svy, k, j, i)) * (A(exk, k, j - 1, i) - A(Ex1, k, j - 1, i));
            const double ex_l2 = (1 - A(svz, k, j - 1, i)) * (A(exj, k, j, i) - A(Ex1, k, j - 1, i)) +
                                 (A(svz, k, j - 1, i)) * (A(exj, k - 1, j, i) - A(Ex1, k - 1, j - 1, i));
            const double ex_r2 = (1 - A(svz, k, j, i)) * (A(exj, k, j, i) - A(Ex1, k, j, i)) +
                                 (A(svz, k, j, i)) * (A(exj, k - 1, j, i) - A(Ex1, k - 1, j, i));
            ex[i3(k, j, i)] = 0.25 * (ex_l3 + ex_r3 + ex_l2 + ex_r2 + A(exj, k, j, i) +
                                      A(exj, k - 1, j, i) + A(exk, k, j, i) + A(exk, k, j - 1, i));
            const double ey_l3 = (1 - A(svx, k - 1, j, i)) * (A(eyk, k, j, i) - A(Ex2, k - 1, j, i)) +
                                 (A(svx, k - 1, j, i)) * (A(eyk, k, j, i - 1) - A(Ex2, k - 1, j, i - 1));
            const double ey_r3 = (1 - A(svx, k, j, i)) * (A(eyk, k, j, i) - A(Ex2, k, j, i)) +
                                 (A(svx, k, j, i)) * (A(eyk, k, j, i - 1) - A(Ex2, k, j, i - 1));
            const double ey_l1 = (1 - A(svz, k, j, i - 1)) * (A(eyi, k, j, i) - A(Ex2, k, j, i - 1)) +
                                 (A(svz, k, j, i - 1)) * (A(eyi, k - 1, j, i) - A(Ex2, k - 1, j, i - 1));
            const double ey_r1 = (1 - A(svz, k, j, i)) * (A(eyi, k, j, i) - A(Ex2, k, j, i)) +
                                 (A(svz, k, j, i)) * (A(eyi, k - 1, j, i) - A(Ex2, k - 1, j, i));
            ey[i3(k, j, i)] = 0.25 * (ey_l3 + ey_r3 + ey_l1 + ey_r1 + A(eyi, k, j, i) +
                                      A(eyi, k - 1, j, i) + A(eyk, k, j, i) + A(eyk, k, j, i - 1));
          }
        }
  }
  // One edge term of CalcRiemannAverage (calcRiemannEmf.hpp:62-428): B and the
  // transverse velocity reconstructed along `along` at the two sides of the edge.
  // `bq(s)` and `vq(s)` return the face value at offset s from the edge's left cell.
  template <class FB, class FV>
  double riemann_edge_term(FB bq, FV vq, double aL, double aR, double dL, double dR,
                           double sgn_phi) {
    double bL, bR, vL, vR, db, dv;
    if (c.order == 4) {
      ppm_states(bq(-2), bq(-1), bq(0), bq(1), bq(2), db, bL);
      ppm_states(bq(-1), bq(0), bq(1), bq(2), bq(3), bR, db);
      ppm_states(vq(-2), vq(-1), vq(0), vq(1), vq(2), dv, vL);
      ppm_states(vq(-1), vq(0), vq(1), vq(2), vq(3), vR, dv);
    } else {
      db = mclim(bq(1) - bq(0), bq(0) - bq(-1));
      bL = bq(0) + 0.5 * db;
      db = mclim(bq(2) - bq(1), bq(1) - bq(0));
      bR = bq(1) - 0.5 * db;
      dv = mclim(vq(1) - vq(0), vq(0) - vq(-1));
      vL = vq(0) + 0.5 * dv;
      dv = mclim(vq(2) - vq(1), vq(1) - vq(0));
      vR = vq(1) - 0.5 * dv;
    }
    double phi = dR * bR - dL * bL;
    if (g_plant == ORC_PLANT_RIEMANN_PHI) phi = -phi;
    return sgn_phi > 0 ? (aL * vL * bL + aR * vR * bR) + phi : (aL * vL * bL + aR * vR * bR) - phi;
  }
  void riemann_average() {
#pragma omp parallel for collapse(2) if (par_)
    for (int k = beg[2]; k < end[2] + offs[2]; k++)
      for (int j = beg[1]; j < end[1] + offs[1]; j++)
        for (int i = beg[0]; i < end[0] + offs[0]; i++) {
          const int im = i - 1, jm = j - 1, km = k - 1;
          auto A = [&](std::vector<double> &a, int kk, int jj, int ii) { return a[i3(kk, jj, ii)]; };
          double aL, aR, dL, dR;
          // --- IDIR faces ---
          // ez at (i-1/2, j-1/2, k)   calcRiemannEmf.hpp:74-130
          aL = 0.5 * (A(axL, k, jm, i) + A(axL, k, jm + 1, i));
          aR = 0.5 * (A(axR, k, jm, i) + A(axR, k, jm + 1, i));
          dL = 0.5 * (A(dxL, k, jm, i) + A(dxL, k, jm + 1, i));
          dR = 0.5 * (A(dxR, k, jm, i) + A(dxR, k, jm + 1, i));
          double ezv = riemann_edge_term([&](int s) { return vs(1, k, j, im + s); },
                                         [&](int s) { return A(ezj, k, j, im + s); }, aL, aR, dL, dR, +1);
          double eyv = 0.0, exv = 0.0;
          if (c.dim == 3) {
            // ey at (i-1/2, j, k-1/2)   :133-190
            aL = 0.5 * (A(axL, km, j, i) + A(axL, km + 1, j, i));
            aR = 0.5 * (A(axR, km, j, i) + A(axR, km + 1, j, i));
            dL = 0.5 * (A(dxL, km, j, i) + A(dxL, km + 1, j, i));
            dR = 0.5 * (A(dxR, km, j, i) + A(dxR, km + 1, j, i));
            eyv = riemann_edge_term([&](int s) { return vs(2, k, j, im + s); },
                                    [&](int s) { return A(eyk, k, j, im + s); }, aL, aR, dL, dR, -1);
            // ex at (i, j-1/2, k-1/2) from JDIR faces   :194-251
            aL = 0.5 * (A(ayL, km, j, i) + A(ayL, km + 1, j, i));
            aR = 0.5 * (A(ayR, km, j, i) + A(ayR, km + 1, j, i));
            dL = 0.5 * (A(dyL, km, j, i) + A(dyL, km + 1, j, i));
            dR = 0.5 * (A(dyR, km, j, i) + A(dyR, km + 1, j, i));
            exv = riemann_edge_term([&](int s) { return vs(2, k, jm + s, i); },
                                    [&](int s) { return A(exk, k, jm + s, i); }, aL, aR, dL, dR, +1);
          }
          // ez (JDIR faces)   :253-308
          aL = 0.5 * (A(ayL, k, j, im) + A(ayL, k, j, im + 1));
          aR = 0.5 * (A(ayR, k, j, im) + A(ayR, k, j, im + 1));
          dL = 0.5 * (A(dyL, k, j, im) + A(dyL, k, j, im + 1));
          dR = 0.5 * (A(dyR, k, j, im) + A(dyR, k, j, im + 1));
          ezv += riemann_edge_term([&](int s) { return vs(0, k, jm + s, i); },
                                   [&](int s) { return A(ezi, k, jm + s, i); }, aL, aR, dL, dR, -1);
          if (c.dim == 3) {
            // ey (KDIR faces)   :313-368
            aL = 0.5 * (A(azL, k, j, im) + A(azL, k, j, im + 1));
            aR = 0.5 * (A(azR, k, j, im) + A(azR, k, j, im + 1));
            dL = 0.5 * (A(dzL, k, j, im) + A(dzL, k, j, im + 1));
            dR = 0.5 * (A(dzR, k, j, im) + A(dzR, k, j, im + 1));
            eyv += riemann_edge_term([&](int s) { return vs(0, km + s, j, i); },
                                     [&](int s) { return A(eyi, km + s, j, i); }, aL, aR, dL, dR, +1);
            // ex (KDIR faces)   :370-425
            aL = 0.5 * (A(azL, k, jm, i) + A(azL, k, jm + 1, i));
            aR = 0.5 * (A(azR, k, jm, i) + A(azR, k, jm + 1, i));
            dL = 0.5 * (A(dzL, k, jm, i) + A(dzL, k, jm + 1, i));
            dR = 0.5 * (A(dzR, k, jm, i) + A(dzR, k, jm + 1, i));
            exv += riemann_edge_term([&](int s) { return vs(1, km + s, j, i); },
                                     [&](int s) { return A(exj, km + s, j, i); }, aL, aR, dL, dR, -1);
            ex[i3(k, j, i)] = exv;
            ey[i3(k, j, i)] = eyv;
          }
          ez[i3(k, j, i)] = ezv;
        }
  }
  // CalcCornerEMF, calcCornerEmf.hpp:14-38
  void calc_corner_emf() {
    if (!c.mhd || c.dim < 2) return;
    if (c.emf == ORC_EMF_ARITHMETIC) arithmetic_corner();
    if (c.emf == ORC_EMF_UCT_CONTACT || c.emf == ORC_EMF_UCT0) {
      calc_cell_emf();
      if (c.emf == ORC_EMF_UCT0) uct0_average();
      else contact_average();
    }
    if (c.emf == ORC_EMF_UCT_HLL || c.emf == ORC_EMF_UCT_HLLD) riemann_average();
  }
  // EvolveMagField, evolveMagField.hpp:53-132 (Cartesian)
  void evolve_mag_field(double dtl) {
    if (!c.mhd || c.dim < 2) return;
#pragma omp parallel for collapse(2) if (par_)
    for (int k = beg[2]; k < end[2] + offs[2]; k++)
      for (int j = beg[1]; j < end[1] + offs[1]; j++)
        for (int i = beg[0]; i < end[0] + offs[0]; i++) {
          auto E = [&](std::vector<double> &a, int kk, int jj, int ii) { return a[i3(kk, jj, ii)]; };
          double rhsx1 = 0.0 - dtl / dx[1][j] * (E(ez, k, j + 1, i) - E(ez, k, j, i));
          double rhsx2 = dtl / dx[0][i] * (E(ez, k, j, i + 1) - E(ez, k, j, i));
          double rhsx3 = 0.0;
          if (c.dim == 3) {
            rhsx1 = rhsx1 + dtl / dx[2][k] * (E(ey, k + 1, j, i) - E(ey, k, j, i));
            rhsx2 = rhsx2 - dtl / dx[2][k] * (E(ex, k + 1, j, i) - E(ex, k, j, i));
            rhsx3 = -dtl / dx[0][i] * (E(ey, k, j, i + 1) - E(ey, k, j, i)) +
                    dtl / dx[1][j] * (E(ex, k, j + 1, i) - E(ex, k, j, i));
          }
          vs(0, k, j, i) = vs(0, k, j, i) + rhsx1;
          vs(1, k, j, i) = vs(1, k, j, i) + rhsx2;
          if (c.dim == 3) vs(2, k, j, i) = vs(2, k, j, i) + rhsx3;
        }
  }

  // ---------------- Stage and cycle ----------------
  // Fluid::EvolveStage, evolveStage.hpp:40-90 (ideal MHD / HD, Cartesian)
  void evolve_stage(double tl, double dtl) {
    (void)tl;
    for (int dir = 0; dir < c.dim; dir++) {  // Fluid::LoopDir, evolveStage.hpp:14-33
      calc_flux(dir);
      if (c.ntracer > 0) tracer_flux(dir);
      calc_rhs(dir, dtl);
      if (c.ntracer > 0) tracer_rhs(dir, dtl);
    }
    if (c.mhd && c.dim >= 2) {
      calc_corner_emf();
      evolve_mag_field(dtl);
      reconstruct_vc_field(Uc);
    }
  }
  // DataBlock::ComputeTimestep, dataBlock.cpp:333-364
  double compute_timestep() {
    double dtmin = INFINITY;
    for (int k = beg[2]; k < end[2]; k++)
      for (int j = beg[1]; j < end[1]; j++)
        for (int i = beg[0]; i < end[0]; i++) dtmin = FMIN(1.0 / InvDt[i3(k, j, i)], dtmin);
    return dtmin;
  }
  void reset_stage() { std::fill(InvDt.begin(), InvDt.end(), 0.0); }
  // StateContainer::CopyFrom / AddAndStore, stateContainer.cpp:19-37, :75-104
  void copy_begin() {
    Uc0 = Uc;
    if (c.mhd) Vs0 = Vs;
  }
  void add_and_store(double wl, double wr) {
    for (size_t p = 0; p < Uc.size(); p++) Uc[p] = wl * Uc[p] + wr * Uc0[p];
    if (c.mhd)
      for (size_t p = 0; p < Vs.size(); p++) Vs[p] = wl * Vs[p] + wr * Vs0[p];
  }
  // Fluid::CheckDivB, checkDivB.hpp:14-68
  double check_divb() {
    if (!c.mhd) return 0.0;
    double divB = 0.0;
    for (int k = beg[2]; k < end[2]; k++)
      for (int j = beg[1]; j < end[1]; j++)
        for (int i = beg[0]; i < end[0]; i++) {
          double dB1 = 0, dB2 = 0, dB3 = 0, d1 = 0, d2 = 0, d3 = 0, B1 = 0, B2 = 0, B3 = 0;
          dB1 = (Ax(0, k, j, i + 1) * vs(0, k, j, i + 1) - Ax(0, k, j, i) * vs(0, k, j, i));
          d1 = 0.5 * (Ax(0, k, j, i + 1) + Ax(0, k, j, i));
          B1 = 0.5 * (vs(0, k, j, i + 1) + vs(0, k, j, i));
          if (c.dim >= 2) {
            dB2 = (Ax(1, k, j + 1, i) * vs(1, k, j + 1, i) - Ax(1, k, j, i) * vs(1, k, j, i));
            d2 = 0.5 * (Ax(1, k, j + 1, i) + Ax(1, k, j, i));
            B2 = 0.5 * (vs(1, k, j + 1, i) + vs(1, k, j, i));
          }
          if (c.dim >= 3) {
            dB3 = (Ax(2, k + 1, j, i) * vs(2, k + 1, j, i) - Ax(2, k, j, i) * vs(2, k, j, i));
            d3 = 0.5 * (Ax(2, k + 1, j, i) + Ax(2, k, j, i));
            B3 = 0.5 * (vs(2, k + 1, j, i) + vs(2, k, j, i));
          }
          double amplitude = 1e-40;
          double am = std::fabs(B1) * d1;
          if (c.dim >= 2) am = am + std::fabs(B2) * d2;
          if (c.dim >= 3) am = am + std::fabs(B3) * d3;
          amplitude += am;
          double s = dB1;
          if (c.dim >= 2) s = s + dB2;
          if (c.dim >= 3) s = s + dB3;
          divB = FMAX(std::fabs(s) / amplitude, divB);
        }
    return divB;
  }
  // TimeIntegrator::Cycle, timeIntegrator.cpp:246-414 (no RKL, Fargo, gravity)
  // returns 0 ok, 1 dt too small (runtime_error), 2 first_dt too large (IDEFIX_ERROR)
  int cycle() {
    const double t0 = t;
    double newdt = dt;
    reset_stage();
    for (int stage = 0; stage < c.nstages; stage++) {
      set_boundaries();
      prim_to_cons();
      if (c.nstages > 1 && stage == 0) copy_begin();
      evolve_stage(t, dt);
      t += dt;
      if (stage == 0) newdt = c.cfl * compute_timestep();
      if (stage > 0) {
        const double wcs = wc[stage - 1], w0s = w0[stage - 1];
        add_and_store(wcs, w0s);
        t = wcs * t + w0s * t0;
      }
      cons_to_prim();
    }
    t = t0 + dt;
    int status = 0;
    if (newdt > c.cfl_max_var * dt) {
      dt = c.cfl_max_var * dt;
    } else {
      if (ncycles == 0 && newdt < 0.5 * dt) status = 2;
      dt = newdt;
    }
    if (status == 0 && dt < 1e-15) status = 1;
    ncycles++;
    return status;
  }
  // main loop, main.cpp:190-225
  int run(double tstop, long maxcycles) {
    while (t < tstop) {
      if (tstop - t < dt) dt = tstop - t;
      int s = cycle();
      if (s) return s;
      if (maxcycles >= 0 && ncycles >= maxcycles) break;
    }
    return 0;
  }
};

}  // namespace

// ======================================================================================
// C API
// ======================================================================================
extern "C" {

void *orc_new(const orc_cfg *cfg) {
  if (cfg->dim < 1 || cfg->dim > 3 || cfg->comps < cfg->dim || cfg->comps > 3) return nullptr;
  if (cfg->mhd && cfg->solver == ORC_SOLVER_HLLC) return nullptr;
  if (!cfg->mhd && cfg->solver == ORC_SOLVER_HLLD) return nullptr;
  /* HllMHD and TvdlfMHD have no uct_hlld face store (hllMHD.hpp:449-452,
     tvdlfMHD.hpp:230-233) */
  if (cfg->mhd && (cfg->solver == ORC_SOLVER_HLL || cfg->solver == ORC_SOLVER_TVDLF) &&
      cfg->emf == ORC_EMF_UCT_HLLD)
    return nullptr;
  return new Oracle(*cfg);
}
void orc_free(void *o) { delete static_cast<Oracle *>(o); }
int orc_nthreads(int n) {
#ifdef _OPENMP
  int prev = omp_get_max_threads();
  if (n > 0) omp_set_num_threads(n);
  return prev;
#else
  (void)n;
  return 1;
#endif
}

void orc_layout(void *p, int *np_tot, int *nghost, int *nvar, int *vs_shape) {
  auto *o = static_cast<Oracle *>(p);
  for (int d = 0; d < 3; d++) {
    np_tot[d] = o->np_tot[d];
    nghost[d] = o->nghost[d];
    vs_shape[d] = o->vsn[d];
  }
  *nvar = o->ntot;
}
double *orc_array(void *p, int which, long *nelem) {
  auto *o = static_cast<Oracle *>(p);
  std::vector<double> *a = nullptr;
  switch (which) {
    case ORC_VC: a = &o->Vc; break;
    case ORC_VS: a = &o->Vs; break;
    case ORC_UC: a = &o->Uc; break;
    case ORC_INVDT: a = &o->InvDt; break;
    case ORC_CMAX: a = &o->cMax; break;
    case ORC_FLUX: a = &o->Flux; break;
    case ORC_EX: a = &o->ex; break;
    case ORC_EY: a = &o->ey; break;
    case ORC_EZ: a = &o->ez; break;
    case ORC_EZI: a = &o->ezi; break;
    case ORC_EZJ: a = &o->ezj; break;
    case ORC_EXJ: a = &o->exj; break;
    case ORC_EXK: a = &o->exk; break;
    case ORC_EYI: a = &o->eyi; break;
    case ORC_EYK: a = &o->eyk; break;
    case ORC_SVX: a = &o->svx; break;
    case ORC_SVY: a = &o->svy; break;
    case ORC_SVZ: a = &o->svz; break;
    case ORC_AXL: a = &o->axL; break;
    case ORC_AXR: a = &o->axR; break;
    case ORC_AYL: a = &o->ayL; break;
    case ORC_AYR: a = &o->ayR; break;
    case ORC_AZL: a = &o->azL; break;
    case ORC_AZR: a = &o->azR; break;
    case ORC_DXL: a = &o->dxL; break;
    case ORC_DXR: a = &o->dxR; break;
    case ORC_DYL: a = &o->dyL; break;
    case ORC_DYR: a = &o->dyR; break;
    case ORC_DZL: a = &o->dzL; break;
    case ORC_DZR: a = &o->dzR; break;
    case ORC_UC0: a = &o->Uc0; break;
    case ORC_VS0: a = &o->Vs0; break;
    default: break;
  }
  if (!a) { *nelem = 0; return nullptr; }
  *nelem = (long)a->size();
  return a->data();
}
double *orc_coord(void *p, int dir, int kind) {
  auto *o = static_cast<Oracle *>(p);
  switch (kind) {
    case 0: return o->x[dir].data();
    case 1: return o->xl[dir].data();
    case 2: return o->xr[dir].data();
    default: return o->dx[dir].data();
  }
}
void orc_set_boundaries(void *p) { static_cast<Oracle *>(p)->set_boundaries(); }
void orc_set_userdef(void *p, orc_userdef_fn fn, void *arg) {
  static_cast<Oracle *>(p)->user_fn = fn;
  static_cast<Oracle *>(p)->user_arg = arg;
}
int orc_set_iso_sound_speed(void *p, const double *cs) {
  Oracle *o = static_cast<Oracle *>(p);
  if (!o->c.isothermal) return 1;
  if (!cs) {
    o->cs_cell.clear();
    return 0;
  }
  o->cs_cell.assign(cs, cs + (size_t)o->np_tot[0] * o->np_tot[1] * o->np_tot[2]);
  return 0;
}
void orc_prim_to_cons(void *p) { static_cast<Oracle *>(p)->prim_to_cons(); }
void orc_cons_to_prim(void *p) { static_cast<Oracle *>(p)->cons_to_prim(); }
void orc_reset_stage(void *p) { static_cast<Oracle *>(p)->reset_stage(); }
void orc_calc_flux(void *p, int dir) { static_cast<Oracle *>(p)->calc_flux(dir); }
void orc_find_shock(void *p, unsigned char *out) {
  Oracle *o = static_cast<Oracle *>(p);
  o->find_shock();
  std::copy(o->shock.begin(), o->shock.end(), out);
}
void orc_calc_rhs(void *p, int dir, double dt) { static_cast<Oracle *>(p)->calc_rhs(dir, dt); }
void orc_calc_corner_emf(void *p) { static_cast<Oracle *>(p)->calc_corner_emf(); }
void orc_evolve_mag_field(void *p, double dt) { static_cast<Oracle *>(p)->evolve_mag_field(dt); }
void orc_reconstruct_vc_field_uc(void *p) {
  auto *o = static_cast<Oracle *>(p);
  o->reconstruct_vc_field(o->Uc);
}
void orc_evolve_stage(void *p, double t, double dt) { static_cast<Oracle *>(p)->evolve_stage(t, dt); }
double orc_compute_timestep(void *p) { return static_cast<Oracle *>(p)->compute_timestep(); }
double orc_check_divb(void *p) { return static_cast<Oracle *>(p)->check_divb(); }
void orc_add_and_store(void *p, double wl, double wr) { static_cast<Oracle *>(p)->add_and_store(wl, wr); }
void orc_copy_begin(void *p) { static_cast<Oracle *>(p)->copy_begin(); }
int orc_cycle(void *p) { return static_cast<Oracle *>(p)->cycle(); }
int orc_run(void *p, double tstop, long maxcycles) { return static_cast<Oracle *>(p)->run(tstop, maxcycles); }
double orc_get_t(void *p) { return static_cast<Oracle *>(p)->t; }
double orc_get_dt(void *p) { return static_cast<Oracle *>(p)->dt; }
void orc_set_t(void *p, double t) { static_cast<Oracle *>(p)->t = t; }
void orc_set_dt(void *p, double dt) { static_cast<Oracle *>(p)->dt = dt; }
long orc_get_ncycles(void *p) { return static_cast<Oracle *>(p)->ncycles; }

int orc_face_flux(const orc_cfg *cfg, int dir, const double *vL_in, const double *vR_in,
                  double *flux, double *cmax, double *emf7) {
  Vars V(*cfg);
  Dir D(V, dir);
  double vL[8], vR[8];
  for (int n = 0; n < V.nvar; n++) { vL[n] = vL_in[n]; vR[n] = vR_in[n]; }
  FaceEmf E;
  E.hasEt = E.hasEb = E.hasSv = E.hasAD = false;
  E.Et = E.Eb = E.sv = E.aL = E.aR = E.dL = E.dR = 0.0;
  bool wrote = true;
  if (cfg->mhd) {
    if (cfg->solver == ORC_SOLVER_ROE) roe_mhd(*cfg, V, D, dir, vL, vR, flux, *cmax, E);
    else if (cfg->solver == ORC_SOLVER_HLL) hll_mhd(*cfg, V, D, dir, vL, vR, flux, *cmax, E);
    else if (cfg->solver == ORC_SOLVER_TVDLF) tvdlf_mhd(*cfg, V, D, dir, vL, vR, flux, *cmax, E);
    else hlld_mhd(*cfg, V, D, dir, vL, vR, flux, *cmax, E);
  } else if (cfg->solver == ORC_SOLVER_HLLC) {
    hllc_hd(*cfg, V, D, vL, vR, flux, *cmax);
  } else if (cfg->solver == ORC_SOLVER_HLL) {
    hll_hd(*cfg, V, D, vL, vR, flux, *cmax);
  } else if (cfg->solver == ORC_SOLVER_TVDLF) {
    tvdlf_hd(*cfg, V, D, vL, vR, flux, *cmax);
  } else {
    wrote = roe_hd(*cfg, V, D, vL, vR, flux, *cmax);
  }
  if (emf7) {
    emf7[0] = E.Et; emf7[1] = E.Eb; emf7[2] = E.sv; emf7[3] = E.aL;
    emf7[4] = E.aR; emf7[5] = E.dL; emf7[6] = E.dR;
  }
  return wrote ? 1 : 0;
}
int orc_plant(int which) {
  const int prev = g_plant;
  g_plant = which;
  return prev;
}
double orc_vanleer(double dvp, double dvm) { return vanleer(dvp, dvm); }
double orc_mclim(double dvp, double dvm) { return mclim(dvp, dvm); }
double orc_limo3lim(double dvp, double dvm, double dx) { return limo3lim(dvp, dvm, dx); }
void orc_ppm_states(double vm2, double vm1, double v0, double vp1, double vp2, double *vl,
                    double *vr) {
  ppm_states(vm2, vm1, v0, vp1, vp2, *vl, *vr);
}
void orc_prim2cons_pt(const orc_cfg *cfg, const double *v, double *u) {
  Vars V(*cfg);
  prim2cons(*cfg, V, v, u);
}
void orc_cons2prim_pt(const orc_cfg *cfg, const double *u, double *v) {
  Vars V(*cfg);
  cons2prim(*cfg, V, u, v);
}

}  // extern "C"
