// Host-side launchers for the kernels of kernels.hpp, gathered per compile-time
// configuration into a type-erased operation table that the C ABI dispatches on.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "fused2d.hpp"
#include "kernels.hpp"
#include "mlp.hpp"

namespace idfx {

// Type-erased operations of one (configuration, compute type) pair.
struct Ops {
  bool valid = false;
  int dim = 0, nvar = 0;
  hipError_t (*sweep)(int dir, bool fused, const GridDesc &g, const SweepArgs &a,
                      const void *emf, int target_waves, hipStream_t s) = nullptr;
  hipError_t (*corner_emf)(const GridDesc &g, const double *Vc, const double *Vs,
                           const void *emf, int scheme, hipStream_t s) = nullptr;
  hipError_t (*evolve_b)(const GridDesc &g, double *Vs, double *Vs0, const void *emf,
                         double dt, int mode, double wc, double w0, hipStream_t s) = nullptr;
  hipError_t (*prim_to_cons)(const GridDesc &g, const double *Vc, double *Uc,
                             hipStream_t s) = nullptr;
  hipError_t (*cons_to_prim)(const GridDesc &g, double *Uc, const double *Uc0, double *Vc,
                             const double *Vs, const int *lo, const int *hi, int flags,
                             double wc, double w0, hipStream_t s) = nullptr;
  hipError_t (*vc_field)(const GridDesc &g, const double *Vs, double *X,
                         hipStream_t s) = nullptr;
  // all-periodic box: every ghost zone in one pass (see k_boundary_periodic)
  hipError_t (*boundary_periodic)(const GridDesc &g, double *Vc, double *Vs,
                                  hipStream_t s) = nullptr;
  hipError_t (*boundary)(const GridDesc &g, double *Vc, double *Vs, int dir, int lbc, int rbc,
                         hipStream_t s) = nullptr;
  hipError_t (*normal_field)(const GridDesc &g, double *Vs, int dir, int recL, int recR,
                             hipStream_t s) = nullptr;
  hipError_t (*correct_flux)(int dir, const GridDesc &g, double *Flux, hipStream_t s) = nullptr;
  hipError_t (*rhs)(int dir, const GridDesc &g, const double *Flux, const double *cMax,
                    double *Uc, double *InvDt, double dt, hipStream_t s) = nullptr;
  hipError_t (*divb)(const GridDesc &g, const double *Vs, unsigned long long *out,
                     hipStream_t s) = nullptr;
  // 3-D adiabatic MHD only: pressure and cell field after a SW_FINAL sweep
  hipError_t (*prs_from_faces)(const GridDesc &g, double *Vc, const double *Vs, double *Uc,
                               int write_uc, hipStream_t s) = nullptr;
  // 3-D MHD only: CalcCornerEMF (uct_contact) + EvolveMagField in one kernel (k_ct3d)
  hipError_t (*ct3d)(const GridDesc &g, const double *Vc, double *Vs, double *Vs0,
                     const void *emf, double dt, int mode, double wc, double w0,
                     hipStream_t s) = nullptr;
  // 2-D MHD only: y sweep + corner EMF + face-B update + ConsToPrim in one kernel
  hipError_t (*fused2d_y)(const GridDesc &g, const SweepArgs &a, const void *emf,
                          Fused2dArgs f, int target_waves, hipStream_t s) = nullptr;
  // learned limiter (RECON_MLP): interface weights of direction dir
  hipError_t (*mlp_infer)(int dir, const GridDesc &g, const double *Vc, const MlpNet &net,
                          float *W, hipStream_t s) = nullptr;
  // shock flattening: flags of the cells around the interior
  hipError_t (*find_shock)(const GridDesc &g, const double *Vc, unsigned char *flags,
                           double smoothing, hipStream_t s) = nullptr;
};

inline unsigned blocks_for(long long n, int bs = 256) {
  return (unsigned)std::max<long long>(1, (n + bs - 1) / bs);
}

// Resident waves of a 256-thread kernel over the device (CUs x blocks per CU x 4), queried
// once per kernel.  target_waves carries 32 waves per CU.
template <typename K>
long long resident_waves(K kernel, int target_waves) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess || nb < 1)
    nb = 1;
  return (long long)nb * 4 * std::max(1, target_waves / 32);
}

// Segment length of a marching kernel: minimises the modelled duration
// ceil(waves / slots) x (ts + 2), waves = columns x ceil(nd / ts), where ts + 2 counts the
// ts cells, the extra face and the window priming of one segment.
inline int march_segment(long long columns, int nd, long long slots) {
  double best = 1e300;
  int bt = nd;
  for (int ts = std::min(nd, 4); ts <= nd; ts++) {
    const long long waves = columns * ((nd + ts - 1) / ts);
    const double cost = (double)((waves + slots - 1) / slots) * (ts + 2);
    if (cost < best) {
      best = cost;
      bt = ts;
    }
  }
  return bt;
}

template <class C, typename T, int DIR, bool FUSED, int EMF, bool XTRA, bool FIN = false>
hipError_t launch_sweep_xtra(const GridDesc &g, const SweepArgs &a, const EmfArrays<T> &e,
                            int target_waves, hipStream_t s) {
  if constexpr (DIR == IDIR) {
    const int nint = g.end[0] - g.beg[0];
    const int nseg = (nint + 61) / 62;
    const long long nj = g.end[1] - g.beg[1] + 2 * a.ext[1];
    const long long nk = g.end[2] - g.beg[2] + 2 * a.ext[2];
    const long long waves = (long long)nseg * nj * nk;
    hipLaunchKernelGGL((k_sweep_x<C, T, FUSED, EMF, XTRA>), dim3(blocks_for(waves, 4)), dim3(256), 0, s, g,
                       a, e, nseg);
  } else {
    MarchGeom m;
    m.ilo = g.beg[0] - a.ext[0];
    m.ihi = g.end[0] + a.ext[0];
    constexpr int OT = (DIR == JDIR) ? 2 : 1;  // the other transverse direction
    m.tlo = g.beg[OT] - a.ext[OT];
    m.thi = g.end[OT] + a.ext[OT];
    m.nxb = (m.ihi - m.ilo + 63) / 64;
    const int nd = g.end[DIR] - g.beg[DIR];
    const long long columns = (long long)m.nxb * (m.thi - m.tlo);
    static const long long slots =
        resident_waves(k_sweep_march<C, T, DIR, FUSED, EMF, XTRA, FIN>, target_waves);
    m.ts = march_segment(columns, nd, slots);
    m.nseg = (nd + m.ts - 1) / m.ts;
    const long long waves = columns * m.nseg;
    hipLaunchKernelGGL((k_sweep_march<C, T, DIR, FUSED, EMF, XTRA, FIN>), dim3(blocks_for(waves, 4)),
                       dim3(256), 0, s, g, a, e, m);
  }
  return hipGetLastError();
}

// The options that add per-cell work (shock flags, passive tracers) select the XTRA
// instantiation; runs without them use kernels compiled without that code.
template <class C, typename T, int DIR, bool FUSED, int EMF>
hipError_t launch_sweep_dir(const GridDesc &g, const SweepArgs &a, const EmfArrays<T> &e,
                            int target_waves, hipStream_t s) {
  if (a.flags & SW_FINAL) {
    // the z sweep of the 3-D ping-pong schedule (adiabatic MHD, uct_contact, no per-cell
    // options)
    if constexpr (DIR == KDIR && C::DIM == 3 && C::MHD && !C::ISO && FUSED &&
                  EMF == EMF_UCT_CONTACT) {
      if (a.shock != nullptr || g.ntr > 0) return hipErrorInvalidValue;
      return launch_sweep_xtra<C, T, DIR, FUSED, EMF, false, true>(g, a, e, target_waves, s);
    } else {
      return hipErrorInvalidValue;
    }
  }
  if (a.shock != nullptr || g.ntr > 0)
    return launch_sweep_xtra<C, T, DIR, FUSED, EMF, true>(g, a, e, target_waves, s);
  return launch_sweep_xtra<C, T, DIR, FUSED, EMF, false>(g, a, e, target_waves, s);
}

// One direction of the sweep with the corner-EMF scheme as a compile-time parameter (the
// face data a sweep stores depend on it); hydro configurations have a single variant.
template <class C, typename T, int DIR, bool FUSED>
hipError_t launch_sweep_emf(const GridDesc &g, const SweepArgs &a, const EmfArrays<T> &e,
                            int target_waves, hipStream_t s) {
  if constexpr (C::MHD) {
    switch (a.emf) {
      case EMF_ARITHMETIC:
        return launch_sweep_dir<C, T, DIR, FUSED, EMF_ARITHMETIC>(g, a, e, target_waves, s);
      case EMF_UCT0: return launch_sweep_dir<C, T, DIR, FUSED, EMF_UCT0>(g, a, e, target_waves, s);
      case EMF_UCT_CONTACT:
        return launch_sweep_dir<C, T, DIR, FUSED, EMF_UCT_CONTACT>(g, a, e, target_waves, s);
      case EMF_UCT_HLL:
        return launch_sweep_dir<C, T, DIR, FUSED, EMF_UCT_HLL>(g, a, e, target_waves, s);
      case EMF_UCT_HLLD:
        return launch_sweep_dir<C, T, DIR, FUSED, EMF_UCT_HLLD>(g, a, e, target_waves, s);
      default: return hipErrorInvalidValue;
    }
  } else {
    return launch_sweep_dir<C, T, DIR, FUSED, EMF_ARITHMETIC>(g, a, e, target_waves, s);
  }
}

template <class C, typename T>
hipError_t launch_sweep(int dir, bool fused, const GridDesc &g, const SweepArgs &a,
                        const void *emf, int target_waves, hipStream_t s) {
  const EmfArrays<T> &e = *static_cast<const EmfArrays<T> *>(emf);
  if (dir == 0) {
    if (fused) return launch_sweep_emf<C, T, 0, true>(g, a, e, target_waves, s);
    if constexpr (std::is_same<T, double>::value)
      return launch_sweep_emf<C, T, 0, false>(g, a, e, target_waves, s);
  }
  if constexpr (C::DIM >= 2) {
    if (dir == 1) {
      if (fused) return launch_sweep_emf<C, T, 1, true>(g, a, e, target_waves, s);
      if constexpr (std::is_same<T, double>::value)
        return launch_sweep_emf<C, T, 1, false>(g, a, e, target_waves, s);
    }
  }
  if constexpr (C::DIM >= 3) {
    if (dir == 2) {
      if (fused) return launch_sweep_emf<C, T, 2, true>(g, a, e, target_waves, s);
      if constexpr (std::is_same<T, double>::value)
        return launch_sweep_emf<C, T, 2, false>(g, a, e, target_waves, s);
    }
  }
  return hipErrorInvalidValue;
}

template <class C, typename T>
hipError_t launch_corner_emf(const GridDesc &g, const double *Vc, const double *Vs,
                             const void *emf, int scheme, hipStream_t s) {
  if constexpr (C::MHD && C::DIM >= 2) {
    const long long n = (long long)(g.end[0] + 1 - g.beg[0]) * (g.end[1] + 1 - g.beg[1]) *
                        (g.end[2] + (C::DIM == 3) - g.beg[2]);
    hipLaunchKernelGGL((k_corner_emf<C, T>), dim3(blocks_for(n)), dim3(256), 0, s, g, Vc, Vs,
                       *static_cast<const EmfArrays<T> *>(emf), scheme);
    return hipGetLastError();
  }
  return hipSuccess;
}

template <class C, typename T>
hipError_t launch_evolve_b(const GridDesc &g, double *Vs, double *Vs0, const void *emf, double dt,
                           int mode, double wc, double w0, hipStream_t s) {
  if constexpr (C::MHD && C::DIM >= 2) {
    const long long n = (long long)(g.end[0] + 1 - g.beg[0]) * (g.end[1] + 1 - g.beg[1]) *
                        (g.end[2] + (C::DIM == 3) - g.beg[2]);
    hipLaunchKernelGGL((k_evolve_b<C, T>), dim3(blocks_for(n)), dim3(256), 0, s, g, Vs, Vs0,
                       *static_cast<const EmfArrays<T> *>(emf), dt, mode, wc, w0);
    return hipGetLastError();
  }
  return hipSuccess;
}

template <class C, typename T>
hipError_t launch_ct3d(const GridDesc &g, const double *Vc, double *Vs, double *Vs0,
                       const void *emf, double dt, int mode, double wc, double w0,
                       hipStream_t s) {
  if constexpr (C::MHD && C::DIM == 3) {
    // tiles of 63 x 7 faces in (i, j); k segments so that the grid holds about 2048
    // blocks of 8 waves (several rounds of the chip's resident blocks)
    const int nfi = g.end[0] - g.beg[0] + 1, nfj = g.end[1] - g.beg[1] + 1;
    const int nfk = g.end[2] - g.beg[2] + 1;
    const int nti = (nfi + 62) / 63, ntj = (nfj + 6) / 7;
    const long long tiles = (long long)nti * ntj;
    int nseg = (int)std::max<long long>(1, std::min<long long>(nfk, (2048 + tiles - 1) / tiles));
    const int kseg = (nfk + nseg - 1) / nseg;
    nseg = (nfk + kseg - 1) / kseg;
    hipLaunchKernelGGL((k_ct3d<C, T>), dim3((unsigned)(tiles * nseg)), dim3(512), 0, s, g, Vc,
                       Vs, Vs0, *static_cast<const EmfArrays<T> *>(emf), dt, mode, wc, w0, ntj,
                       nseg, kseg);
    return hipGetLastError();
  }
  return hipErrorInvalidValue;
}

template <class C>
hipError_t launch_prim_to_cons(const GridDesc &g, const double *Vc, double *Uc, hipStream_t s) {
  const long long n = (long long)g.nt[0] * g.nt[1] * g.nt[2];
  hipLaunchKernelGGL((k_prim_to_cons<C>), dim3(blocks_for(n)), dim3(256), 0, s, g, Vc, Uc);
  return hipGetLastError();
}

template <class C>
hipError_t launch_cons_to_prim(const GridDesc &g, double *Uc, const double *Uc0, double *Vc,
                               const double *Vs, const int *lo, const int *hi, int flags,
                               double wc, double w0, hipStream_t s) {
  const int n0 = hi[0] - lo[0], n1 = hi[1] - lo[1], n2 = hi[2] - lo[2];
  if (n0 <= 0 || n1 <= 0 || n2 <= 0) return hipSuccess;
  const int nxb = (n0 + 255) / 256;  // runs of 256 cells along x per row
  hipLaunchKernelGGL((k_cons_to_prim<C>), dim3(nxb * n1, n2), dim3(256), 0, s, g, Uc, Uc0, Vc,
                     Vs, lo[0], lo[1], lo[2], n0, nxb, flags, wc, w0);
  return hipGetLastError();
}

template <class C>
hipError_t launch_prs_from_faces(const GridDesc &g, double *Vc, const double *Vs, double *Uc,
                                 int write_uc, hipStream_t s) {
  const int n0 = g.end[0] - g.beg[0], n1 = g.end[1] - g.beg[1], n2 = g.end[2] - g.beg[2];
  const int nxb = (n0 + 255) / 256;
  hipLaunchKernelGGL((k_prs_from_faces<C>), dim3(nxb * n1, n2), dim3(256), 0, s, g, Vc, Vs, Uc,
                     n0, nxb, write_uc);
  return hipGetLastError();
}

template <class C>
hipError_t launch_vc_field(const GridDesc &g, const double *Vs, double *X, hipStream_t s) {
  if constexpr (C::MHD) {
    const long long n = (long long)g.nt[0] * g.nt[1] * g.nt[2];
    hipLaunchKernelGGL((k_vc_field<C>), dim3(blocks_for(n)), dim3(256), 0, s, g, Vs, X);
    return hipGetLastError();
  }
  return hipSuccess;
}

template <class C>
hipError_t launch_boundary(const GridDesc &g, double *Vc, double *Vs, int dir, int lbc, int rbc,
                           hipStream_t s) {
  constexpr int NF = C::MHD ? C::DIM : 0;
  long long total = 0;
  for (int f = -1; f < NF; f++) {
    long long sz = 2;
    for (int d = 0; d < 3; d++) sz *= (d == dir) ? g.beg[dir] : g.nt[d] + ((f == d) ? 1 : 0);
    total += sz;
  }
  hipLaunchKernelGGL((k_boundary<C>), dim3(blocks_for(total)), dim3(256), 0, s, g, Vc, Vs, dir,
                     lbc, rbc);
  return hipGetLastError();
}

template <class C>
hipError_t launch_boundary_periodic(const GridDesc &g, double *Vc, double *Vs, hipStream_t s) {
  constexpr int NF = C::MHD ? C::DIM : 0;
  long long total = 0;
  for (int f = -1; f < NF; f++)
    for (int d = 0; d < C::DIM; d++) {
      long long sz = 2LL * g.beg[d];
      for (int d2 = 0; d2 < 3; d2++) {
        if (d2 == d) continue;
        const int hi = g.end[d2] + ((f == d2) ? 1 : 0);
        sz *= (d2 < d) ? (hi - g.beg[d2]) : (g.nt[d2] + ((f == d2) ? 1 : 0));
      }
      total += sz;
    }
  hipLaunchKernelGGL((k_boundary_periodic<C>), dim3(blocks_for(total)), dim3(256), 0, s, g, Vc, Vs);
  return hipGetLastError();
}

template <class C>
hipError_t launch_normal_field(const GridDesc &g, double *Vs, int dir, int recL, int recR,
                               hipStream_t s) {
  if constexpr (C::MHD) {
    const int o1 = (dir == 0) ? 1 : 0;
    const int o2 = (dir == 2) ? 1 : 2;
    const long long n = (long long)g.nt[o1] * g.nt[o2];
    hipLaunchKernelGGL((k_normal_field<C>), dim3(blocks_for(n)), dim3(256), 0, s, g, Vs, dir,
                       recL, recR);
    return hipGetLastError();
  }
  return hipSuccess;
}

template <class C, int DIR>
hipError_t launch_correct_rhs_dir(bool rhs, const GridDesc &g, double *Flux, const double *cMax,
                                  double *Uc, double *InvDt, double dt, hipStream_t s) {
  constexpr int io = DIR == 0, jo = DIR == 1, ko = DIR == 2;
  if (!rhs) {
    const long long n = (long long)(g.end[0] + io - g.beg[0]) * (g.end[1] + jo - g.beg[1]) *
                        (g.end[2] + ko - g.beg[2]);
    hipLaunchKernelGGL((k_correct_flux<C, DIR>), dim3(blocks_for(n)), dim3(256), 0, s, g, Flux);
  } else {
    const long long n = (long long)(g.end[0] - g.beg[0]) * (g.end[1] - g.beg[1]) *
                        (g.end[2] - g.beg[2]);
    hipLaunchKernelGGL((k_rhs<C, DIR>), dim3(blocks_for(n)), dim3(256), 0, s, g, Flux, cMax, Uc,
                       InvDt, dt);
  }
  return hipGetLastError();
}

template <class C>
hipError_t launch_correct_flux(int dir, const GridDesc &g, double *Flux, hipStream_t s) {
  if (dir == 0) return launch_correct_rhs_dir<C, 0>(false, g, Flux, nullptr, nullptr, nullptr, 0, s);
  if constexpr (C::DIM >= 2)
    if (dir == 1) return launch_correct_rhs_dir<C, 1>(false, g, Flux, nullptr, nullptr, nullptr, 0, s);
  if constexpr (C::DIM >= 3)
    if (dir == 2) return launch_correct_rhs_dir<C, 2>(false, g, Flux, nullptr, nullptr, nullptr, 0, s);
  return hipErrorInvalidValue;
}

template <class C>
hipError_t launch_rhs(int dir, const GridDesc &g, const double *Flux, const double *cMax,
                      double *Uc, double *InvDt, double dt, hipStream_t s) {
  double *F = const_cast<double *>(Flux);
  if (dir == 0) return launch_correct_rhs_dir<C, 0>(true, g, F, cMax, Uc, InvDt, dt, s);
  if constexpr (C::DIM >= 2)
    if (dir == 1) return launch_correct_rhs_dir<C, 1>(true, g, F, cMax, Uc, InvDt, dt, s);
  if constexpr (C::DIM >= 3)
    if (dir == 2) return launch_correct_rhs_dir<C, 2>(true, g, F, cMax, Uc, InvDt, dt, s);
  return hipErrorInvalidValue;
}

template <class C>
hipError_t launch_divb(const GridDesc &g, const double *Vs, unsigned long long *out,
                       hipStream_t s) {
  if constexpr (C::MHD) {
    hipLaunchKernelGGL((k_divb_max<C>), dim3(1024), dim3(256), 0, s, g, Vs, out);
    return hipGetLastError();
  }
  return hipSuccess;
}

// Lanes of halo a corner EMF at column x draws y-face data from: columns x-L .. x+R.
inline void fused2d_halo(int order, int emf, int &L, int &R) {
  if (emf == EMF_UCT_HLL || emf == EMF_UCT_HLLD) {
    L = order == RECON_PPM ? 3 : 2;
    R = order == RECON_PPM ? 2 : 1;
  } else {
    L = 1;
    R = 0;
  }
}

template <class C, typename T, int EMF>
hipError_t launch_fused2d_y_emf(const GridDesc &g, const SweepArgs &a, const void *emf,
                                Fused2dArgs f, int target_waves, hipStream_t s) {
  int L, R;
  fused2d_halo(C::ORDER, EMF, L, R);
  // output columns hl .. hl+w-1 of each wave; corners at hl .. hl+w; y faces at
  // hl-L .. hl+w+R, all inside the 64 lanes
  f.hl = L;
  f.w = 63 - L - R;
  MarchGeom m{};
  const int ncols = g.end[0] + 1 - g.beg[0];  // face columns beg .. end
  m.nxb = (ncols + f.w - 1) / f.w;
  const int nd = g.end[1] - g.beg[1];
  if (a.shock != nullptr) {
    static const long long slots = resident_waves(k_fused2d_y<C, T, EMF, true>, target_waves);
    m.ts = march_segment(m.nxb, nd, slots);
    m.nseg = (nd + m.ts - 1) / m.ts;
    const long long waves = (long long)m.nxb * m.nseg;
    hipLaunchKernelGGL((k_fused2d_y<C, T, EMF, true>), dim3(blocks_for(waves, 4)), dim3(256), 0,
                       s, g, a, *static_cast<const EmfArrays<T> *>(emf), f, m);
  } else {
    static const long long slots = resident_waves(k_fused2d_y<C, T, EMF, false>, target_waves);
    m.ts = march_segment(m.nxb, nd, slots);
    m.nseg = (nd + m.ts - 1) / m.ts;
    const long long waves = (long long)m.nxb * m.nseg;
    hipLaunchKernelGGL((k_fused2d_y<C, T, EMF, false>), dim3(blocks_for(waves, 4)), dim3(256), 0,
                       s, g, a, *static_cast<const EmfArrays<T> *>(emf), f, m);
  }
  return hipGetLastError();
}

template <class C, typename T>
hipError_t launch_fused2d_y(const GridDesc &g, const SweepArgs &a, const void *emf,
                            Fused2dArgs f, int target_waves, hipStream_t s) {
  if constexpr (C::MHD && C::DIM == 2) {
    switch (a.emf) {
      case EMF_ARITHMETIC: return launch_fused2d_y_emf<C, T, EMF_ARITHMETIC>(g, a, emf, f, target_waves, s);
      case EMF_UCT0: return launch_fused2d_y_emf<C, T, EMF_UCT0>(g, a, emf, f, target_waves, s);
      case EMF_UCT_CONTACT: return launch_fused2d_y_emf<C, T, EMF_UCT_CONTACT>(g, a, emf, f, target_waves, s);
      case EMF_UCT_HLL: return launch_fused2d_y_emf<C, T, EMF_UCT_HLL>(g, a, emf, f, target_waves, s);
      case EMF_UCT_HLLD: return launch_fused2d_y_emf<C, T, EMF_UCT_HLLD>(g, a, emf, f, target_waves, s);
      default: break;
    }
  }
  return hipErrorInvalidValue;
}

// Interface weights for every face of direction DIR whose stencil lies in the array:
// faces beg .. end along DIR, every transverse index (the sweeps' extended rows included).
template <class C, int DIR>
hipError_t launch_mlp_dir(const GridDesc &g, const double *Vc, const MlpNet &net, float *W,
                          hipStream_t s) {
  int lo[3], n[3];
  for (int d = 0; d < 3; d++) {
    if (d == DIR) {
      lo[d] = g.beg[d];
      n[d] = g.end[d] + 1 - g.beg[d];
    } else {
      lo[d] = 0;
      n[d] = g.nt[d];
    }
  }
  const long long waves = (long long)((n[0] + 63) / 64) * n[1] * n[2];
  hipLaunchKernelGGL((k_mlp_infer<C, DIR>), dim3(blocks_for(waves, 4)), dim3(256), 0, s, g, Vc,
                     net, W, lo[0], lo[1], lo[2], n[0], n[1], n[2]);
  return hipGetLastError();
}

template <class C>
hipError_t launch_mlp(int dir, const GridDesc &g, const double *Vc, const MlpNet &net,
                      float *W, hipStream_t s) {
  if constexpr (C::ORDER == RECON_MLP) {
    if (dir == 0) return launch_mlp_dir<C, 0>(g, Vc, net, W, s);
    if constexpr (C::DIM >= 2)
      if (dir == 1) return launch_mlp_dir<C, 1>(g, Vc, net, W, s);
    if constexpr (C::DIM >= 3)
      if (dir == 2) return launch_mlp_dir<C, 2>(g, Vc, net, W, s);
  }
  return hipErrorInvalidValue;
}

template <class C>
hipError_t launch_find_shock(const GridDesc &g, const double *Vc, unsigned char *flags,
                             double smoothing, hipStream_t s) {
  const long long n = (long long)(g.end[0] - g.beg[0] + 2) *
                      (g.end[1] - g.beg[1] + (C::DIM >= 2 ? 2 : 0)) *
                      (g.end[2] - g.beg[2] + (C::DIM == 3 ? 2 : 0));
  hipLaunchKernelGGL((k_find_shock<C>), dim3(blocks_for(n)), dim3(256), 0, s, g, Vc, flags,
                     smoothing);
  return hipGetLastError();
}

template <class C, typename T>
Ops make_ops() {
  Ops o;
  o.valid = true;
  o.dim = C::DIM;
  o.nvar = C::NVAR;
  o.sweep = &launch_sweep<C, T>;
  o.corner_emf = &launch_corner_emf<C, T>;
  o.evolve_b = &launch_evolve_b<C, T>;
  o.prim_to_cons = &launch_prim_to_cons<C>;
  o.cons_to_prim = &launch_cons_to_prim<C>;
  o.vc_field = &launch_vc_field<C>;
  o.boundary = &launch_boundary<C>;
  o.boundary_periodic = &launch_boundary_periodic<C>;
  o.normal_field = &launch_normal_field<C>;
  o.correct_flux = &launch_correct_flux<C>;
  o.rhs = &launch_rhs<C>;
  o.divb = &launch_divb<C>;
  if constexpr (C::MHD && C::DIM == 2) o.fused2d_y = &launch_fused2d_y<C, T>;
  if constexpr (C::MHD && C::DIM == 3) o.ct3d = &launch_ct3d<C, T>;
  if constexpr (C::MHD && C::DIM == 3 && !C::ISO) o.prs_from_faces = &launch_prs_from_faces<C>;
  if constexpr (C::ORDER == RECON_MLP) o.mlp_infer = &launch_mlp<C>;
  o.find_shock = &launch_find_shock<C>;
  return o;
}

// Registry key: one entry per built configuration and precision.
constexpr int ops_key(int dim, int comps, int mhd, int order, int solver, int fp32,
                      int iso = 0) {
  return (((((dim * 4 + comps) * 2 + mhd) * 16 + order) * 8 + solver) * 2 + fp32) * 2 + iso;
}
constexpr int kOpsTableSize = ops_key(3, 3, 1, 15, 7, 1, 1) + 1;
Ops *ops_table();  // defined in capi.hip

}  // namespace idfx
