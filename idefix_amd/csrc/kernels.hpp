// gfx950 kernels of the per-stage finite-volume update.
//
// Layout is the reference's LayoutRight ([var][k][j][i], i fastest) for Vc/Uc and
// [dir][k][j][i] with one extra face along each active direction for Vs
// (src/fluid/fluid.hpp:529-560).  Kernels index with the ghost-inclusive np_tot extents.
//
// Sweeps.  The x sweep gives one face per lane: a wave covers 64 faces and updates the
// 63 cells between them, taking the right-hand flux from the neighbouring lane with a
// cross-lane shuffle.  The y and z sweeps march: a lane owns one x column and walks a
// segment of cells along the sweep direction, carrying the stencil window, the
// reconstructed state of the previous cell and the previous face flux in registers, so
// every cell is reconstructed once and every face solved once (plus one face per
// segment).  In FUSED mode the flux divergence goes straight into Uc in FP64 and the
// intercell flux never reaches HBM; in flux-only mode the kernel writes FluxRiemann
// like the reference (needed when a user flux hook is enrolled).
#pragma once

#include "physics.hpp"

namespace idfx {

struct GridDesc {
  int nt[3];    // np_tot per direction (i, j, k)
  int beg[3];   // first interior index
  int end[3];   // one past the last interior index
  int vsn[3];   // Vs extents (np_tot + 1 along active directions)
  long long sv;   // stride between variables of Vc/Uc
  long long svs;  // stride between components of Vs
  const double *dx[3];
  double gamma;  // adiabatic index, or the isothermal sound speed (isothermal configurations)
  int ntr;      // passive tracers, stored after the fluid variables of Vc/Uc
  // isothermal configurations with a per-cell sound speed ("csiso userdef",
  // eos_isothermal.hpp:29-33, :70-75): the caller's [np_tot cells] array; else nullptr
  const double *csiso;
};

template <typename T>
struct EmfArrays {
  T *ex, *ey, *ez;               // corner (edge) EMFs
  T *ezi, *ezj, *exj, *exk, *eyi, *eyk;  // face EMFs (Et/Eb of each sweep)
  unsigned char *svx, *svy, *svz;  // uct_contact upwind weights, as codes (sv_code)
  T *axL, *axR, *ayL, *ayR, *azL, *azR;  // uct_hll(d) coefficients
  T *dxL, *dxR, *dyL, *dyR, *dzL, *dzR;
};

// sweep flags
enum : int {
  SW_INIT_UC = 1,      // Uc = PrimToCons(Vc) + rhs   (else Uc += rhs)
  SW_SAVE_UC0 = 2,     // Uc0 = PrimToCons(Vc)        (the RK "begin" state)
  SW_INVDT_SET = 4,    // InvDt = term                (first direction after ResetStage)
  SW_INVDT_ADD = 8,    // InvDt += term
  SW_DT_REDUCE = 16,   // fold max(InvDt) of the updated cells into *dtmax
  SW_NO_INVDT_STORE = 32,  // keep InvDt in registers only (requires SW_INVDT_SET)
  // last sweep of a 3-D stage on the ping-pong schedule (the FIN kernels): the updated
  // cell state goes through AddAndStore and the field-independent part of ConsToPrim
  // into VcOut instead of back into Uc
  SW_FINAL = 64,
  SW_RK = 128,         // with SW_FINAL: U = wc U + w0 U0 (stages after the first)
  SW_WRITE_UC = 256,   // with SW_FINAL: also store the combined U into Uc
};

struct SweepArgs {
  const double *Vc;
  const double *Vs;
  double *Uc;
  double *Uc0;
  double *InvDt;
  double *Flux;   // flux-only mode
  double *cMax;   // flux-only mode
  double dt;
  int emf;
  int flags;
  int ext[3];     // transverse extension of the face range (CT needs)
  unsigned long long *dtmax;  // bit pattern of max(InvDt) (non-negative doubles)
  int uniform;    // all cells share one volume: dtdV below equals dt/dV(k,j,i)
  double dtdV;
  const float *mlpw;  // learned-limiter interface weights of this direction (RECON_MLP)
  const unsigned char *shock;  // shock flags per cell (shock flattening) or nullptr
  double *VcOut;      // SW_FINAL: output primitive state
  double wc, w0;      // SW_RK weights
};

// Element i of an array whose base is uniform across the wave, addressed as base + 32-bit
// byte offset: the compiler keeps the base in scalar registers and one 32-bit VGPR per
// access (global_load ... saddr), instead of a 64-bit VGPR address pair.  Arrays handled
// this way hold fewer than 2^29 elements (checked at context creation).
template <typename T>
__device__ __forceinline__ T ld32(const T *base, unsigned i) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                      (unsigned)(i * (unsigned)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void st32(T *base, unsigned i, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (unsigned)(i * (unsigned)sizeof(T))) = v;
}

// The uct_contact upwind weight of a face is 0, 1/2 or 1 (K_StoreContact,
// storeFlux.hpp:39-43); it is stored as the byte 2 w (1 byte per face instead of 8) and
// decoded exactly.
template <typename T>
__device__ __forceinline__ unsigned char sv_code(T w) {
  return (unsigned char)(T(2) * w);
}
template <typename T>
__device__ __forceinline__ T sv_val(unsigned char c) {
  return T(0.5) * T(c);
}

// The equation-of-state constant a Riemann solver takes at the face between cells pl and
// pr: gamma, or in isothermal configurations the face sound speed, the average of the two
// cells' values (hllcHD.hpp:67-70, hlldMHD.hpp:188-191); a constant cs gives
// 0.5 (cs + cs) = cs.
template <class C, typename T>
__device__ __forceinline__ T eos_face(const GridDesc &g, unsigned pl, unsigned pr) {
  if constexpr (C::ISO) {
    if (g.csiso != nullptr)
      return (T)(0.5 * (ld32(g.csiso, pr) + ld32(g.csiso, pl)));
  }
  return (T)g.gamma;
}

__device__ __forceinline__ long long idx3(const GridDesc &g, int k, int j, int i) {
  return ((long long)k * g.nt[1] + j) * g.nt[0] + i;
}
__device__ __forceinline__ long long idxs(const GridDesc &g, int n, int k, int j, int i) {
  return (((long long)n * g.vsn[2] + k) * g.vsn[1] + j) * g.vsn[0] + i;
}

// Work-groups are dispatched to the 8 XCDs in turn (block b on XCD b % 8), each with its
// own L2.  This maps the dispatched block to a logical block so that each XCD takes a
// contiguous run of logical blocks: kernels whose threads read the neighbouring rows and
// planes of the blocks before them then find them in their own L2.  A bijection on
// [0, gridDim.x).
__device__ __forceinline__ int xcd_contiguous_block() {
  const int nb = gridDim.x, xcd = blockIdx.x % 8, q8 = nb / 8, r8 = nb % 8;
  return (xcd < r8 ? xcd * (q8 + 1) : r8 * (q8 + 1) + (xcd - r8) * q8) + blockIdx.x / 8;
}

// Cartesian volume and areas, makeGeometry.cpp:67-70, :146-147, :178-179, :201-202
template <int DIM>
__device__ __forceinline__ double cell_dV(const GridDesc &g, int k, int j, int i) {
  double v = g.dx[0][i];
  if constexpr (DIM >= 2) v = v * g.dx[1][j];
  if constexpr (DIM >= 3) v = v * g.dx[2][k];
  return v;
}
template <int DIM, int DIR>
__device__ __forceinline__ double face_A(const GridDesc &g, int k, int j, int i) {
  double a;
  if constexpr (DIR == IDIR) {
    a = 1.0;
    if constexpr (DIM >= 2) a = a * g.dx[1][j];
    if constexpr (DIM >= 3) a = a * g.dx[2][k];
  } else if constexpr (DIR == JDIR) {
    a = g.dx[0][i];
    if constexpr (DIM >= 2) a = a * 1.0;
    if constexpr (DIM >= 3) a = a * g.dx[2][k];
  } else {
    a = g.dx[0][i];
    if constexpr (DIM >= 2) a = a * g.dx[1][j];
    if constexpr (DIM >= 3) a = a * 1.0;
  }
  return a;
}

template <typename T, int DIR>
__device__ __forceinline__ T *et_of(const EmfArrays<T> &e) {
  return DIR == 0 ? e.ezi : DIR == 1 ? e.ezj : e.eyk;
}
template <typename T, int DIR>
__device__ __forceinline__ T *eb_of(const EmfArrays<T> &e) {
  return DIR == 0 ? e.eyi : DIR == 1 ? e.exj : e.exk;
}

// Store the face-EMF data of one face (the K_Store* of storeFlux.hpp).
template <class C, int DIR, typename T>
__device__ __forceinline__ void store_face_emf(const EmfArrays<T> &e, unsigned p, int emf,
                                               const FaceEmf<T> &E) {
  if constexpr (C::MHD) {
    if (emf <= EMF_UCT_CONTACT) {
      if constexpr (C::COMP > 1) {
        st32(et_of<T, DIR>(e), p, E.Et);
        if constexpr (C::DIM == 3) st32(eb_of<T, DIR>(e), p, E.Eb);
      }
      if (emf == EMF_UCT_CONTACT)
        st32(DIR == 0 ? e.svx : DIR == 1 ? e.svy : e.svz, p, sv_code(E.sv));
    } else {
      if constexpr (C::COMP > 1) {
        st32(et_of<T, DIR>(e), p, E.Et);
        if constexpr (C::COMP >= 3) st32(eb_of<T, DIR>(e), p, E.Eb);
      }
      st32(DIR == 0 ? e.axL : DIR == 1 ? e.ayL : e.azL, p, E.aL);
      st32(DIR == 0 ? e.axR : DIR == 1 ? e.ayR : e.azR, p, E.aR);
      st32(DIR == 0 ? e.dxL : DIR == 1 ? e.dyL : e.dzL, p, E.dL);
      st32(DIR == 0 ? e.dxR : DIR == 1 ? e.dyR : e.dzR, p, E.dR);
    }
  }
}

// Per-cell reconstruction from a stencil window centred on the cell (PLM: 3 cells,
// PPM: 5 cells).  Returns the cell's left-face value `vl` (the right state of its left
// face) and right-face value `vr` (the left state of its right face), exactly as
// ExtrapolatePrimVar computes them for the two faces (extrapolateToFaces.hpp:88-124,
// :247-306).
template <class C, typename T>
__device__ __forceinline__ void recon_cell(const T *w, int nv, T &vl, T &vr, T dxc = T(0),
                                           bool flat = false) {
  if constexpr (C::ORDER == RECON_CONSTANT) {
    vl = w[0];
    vr = w[0];
  } else if constexpr (C::ORDER == RECON_PLM) {
    // w[0] = V(c-1), w[1] = V(c), w[2] = V(c+1)
    // a cell flagged by the shock detector takes MinModLim, which as built returns 0
    // (slopeLimiter.hpp:16-23, extrapolateToFaces.hpp:98-120)
    const T dvm = w[1] - w[0];
    const T dvp = w[2] - w[1];
    const T dv = flat ? T(0) : vanleer(dvp, dvm);
    vl = w[1] - T(0.5) * dv;
    vr = w[1] + T(0.5) * dv;
  } else if constexpr (C::ORDER == RECON_LIMO3) {
    // extrapolateToFaces.hpp:170-246 (no shock flattening); dxc is the cell's width.  The
    // positivity fallback calls MinModLim, which as written always returns 0
    // (slopeLimiter.hpp:16-23): the state falls back to the cell value.
    const T dvm = w[1] - w[0];
    const T dvp = w[2] - w[1];
    T dv = flat ? T(0) : dvp * limo3_lim(dvp, dvm, dxc);
    T r = w[1] + T(0.5) * dv;
    dv = flat ? T(0) : dvm * limo3_lim(dvm, dvp, dxc);
    T l = w[1] - T(0.5) * dv;
    if (nv == C::RHO || nv == C::ENG) {
      r = (r <= T(0)) ? w[1] + T(0.5) * T(0) : r;
      l = (l <= T(0)) ? w[1] - T(0.5) * T(0) : l;
    }
    vl = l;
    vr = r;
  } else {
    // w[0..4] = V(c-2) .. V(c+2)
    T l, r;
    ppm_states(w[0], w[1], w[2], w[3], w[4], l, r);
    if (nv == C::RHO || nv == C::ENG) {
      // van Leer fallback of extrapolateToFaces.hpp:260-304 (value selects, see ppm_states)
      const T dv = vanleer(w[3] - w[2], w[2] - w[1]);
      r = (r <= T(0)) ? w[2] + T(0.5) * dv : r;
      l = (l <= T(0)) ? w[2] - T(0.5) * dv : l;
    }
    vl = l;
    vr = r;
  }
}
template <class C>
constexpr int window() {
  return C::ORDER == RECON_CONSTANT
             ? 1
             : ((C::ORDER == RECON_PLM || C::ORDER == RECON_LIMO3 || C::ORDER == RECON_MLP) ? 3 : 5);
}
// Width of cell `cell` along dir, the only geometric input of a reconstruction (LimO3).
template <class C, typename T>
__device__ __forceinline__ T recon_dx(const GridDesc &g, int dir, int cell) {
  if constexpr (C::ORDER == RECON_LIMO3) return (T)ld32(g.dx[dir], (unsigned)cell);
  else return T(0);
}
// Shock flag of the cell at linear index p (only PLM and LimO3 read it).
template <class C>
__device__ __forceinline__ bool recon_flat(const SweepArgs &a, unsigned p) {
  if constexpr (C::ORDER == RECON_PLM || C::ORDER == RECON_LIMO3)
    return a.shock != nullptr && ld32(a.shock, p) != 0;
  else return false;
}

// Learned-limiter reconstruction of one cell from the window V(c-1), V(c), V(c+1) and
// the interface weights W[q][face] (float, SoA with stride sv): vl, the right state of
// face c (index pc), from wR of that face; vr, the left state of face c+1 (index
// pc + stride), from its wL.  q is the face-frame index of variable n.
template <class C, int DIR, typename T>
__device__ __forceinline__ void recon_mlp(const T *w, int n, const float *W, long long sv,
                                          unsigned pc, unsigned stride, T &vl, T &vr) {
  int q = 0;
  if (n == C::RHO) q = 0;
  else if (n == C::template Xn<DIR>()) q = 1;
  else if (n == C::template Xt<DIR>()) q = 2;
  else if (n == C::template Xb<DIR>()) q = 3;
  else if (n == C::PRS) q = 4;
  else if (n == C::template BXn<DIR>()) q = 5;
  else if (n == C::template BXt<DIR>()) q = 6;
  else q = 7;
  const T wr = (T)ld32(W + (2 * q + 1) * sv, pc);
  const T wl = (T)ld32(W + (2 * q) * sv, pc + stride);
  vl = w[1] - wr * (w[1] - w[0]);
  vr = w[1] + wl * (w[2] - w[1]);
}

// Conserved state at the start of the sweep for one cell: PrimToCons(Vc) (FP64).
template <class C>
__device__ __forceinline__ void cell_p2c(const GridDesc &g, const double *Vc, unsigned p,
                                         double *u) {
  double v[C::NVAR];
#pragma unroll
  for (int n = 0; n < C::NVAR; n++) v[n] = ld32(Vc + n * g.sv, p);
  prim2cons<C>(v, u, g.gamma);
}

// Block-level max of non-negative doubles by bit pattern, one atomic per block.
__device__ __forceinline__ void block_max_atomic(unsigned long long v, unsigned long long *dst) {
  __shared__ unsigned long long red[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = red[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) m = red[q] > m ? red[q] : m;
    atomicMax(dst, m);
  }
}

// Passive tracers of one cell (tracer.hpp:58-111, tracer.cpp:13-47): the face value of
// each tracer is upwinded by the sign of the Riemann mass flux with a van Leer slope, its
// flux is F_rho * vface * A, and rho*tracer is updated like the fluid variables.  Both
// faces of the cell are evaluated here; the neighbouring cell evaluates the shared face
// from the same inputs, so the update stays conservative.
template <class C>
__device__ __forceinline__ double tracer_face(const double *V, unsigned p, unsigned sd,
                                              double frho) {
  // face between the cells at p - sd and p
  if (frho > 0) {
    const double vm2 = ld32(V, p - 2 * sd), vm1 = ld32(V, p - sd), v0 = ld32(V, p);
    const double dv = vanleer(v0 - vm1, vm1 - vm2);
    return vm1 + 0.5 * dv;
  }
  const double vm1 = ld32(V, p - sd), v0 = ld32(V, p), vp1 = ld32(V, p + sd);
  const double dv = vanleer(vp1 - v0, v0 - vm1);
  return v0 - 0.5 * dv;
}
template <class C, int DIR>
__device__ __forceinline__ void tracer_update(const GridDesc &g, const SweepArgs &a, unsigned p,
                                           double dtdV, double A, double fm, double fp) {
  const unsigned sd = DIR == 0 ? 1u : DIR == 1 ? (unsigned)g.nt[0] : (unsigned)(g.nt[0] * g.nt[1]);
  const double rho = ld32(a.Vc + C::RHO * g.sv, p);
  for (int t = 0; t < g.ntr; t++) {
    const long long n = C::NVAR + t;
    const double *V = a.Vc + n * g.sv;
    double u;
    if (a.flags & (SW_INIT_UC | SW_SAVE_UC0)) {
      const double u0 = ld32(V, p) * rho;  // Tracer::ConvertPrimToCons
      if (a.flags & SW_SAVE_UC0) st32(a.Uc0 + n * g.sv, p, u0);
      u = (a.flags & SW_INIT_UC) ? u0 : ld32((const double *)a.Uc + n * g.sv, p);
    } else {
      u = ld32((const double *)a.Uc + n * g.sv, p);
    }
    const double ftp = fp * tracer_face<C>(V, p + sd, sd, fp) * A;
    const double ftm = fm * tracer_face<C>(V, p, sd, fm) * A;
    st32(a.Uc + n * g.sv, p, u + -dtdV * (ftp - ftm));
  }
}

// Update of one cell by the flux difference along DIR (calcRightHandSide.hpp:288-549,
// Cartesian, with the area correction of :156-194 applied to both faces first).  XTRA
// compiles in the optional per-cell work (passive tracers); without it the kernels carry
// no trace of the options.
template <class C, int DIR, bool XTRA, bool FIN = false, typename T>
__device__ __forceinline__ double cell_update(const GridDesc &g, const SweepArgs &a, int k,
                                              int j, int i, const T *Fm, const T *Fp,
                                              T cmaxm, T cmaxp,
                                              const double *u0pre = nullptr,
                                              const double *upre = nullptr) {
  const unsigned p = (unsigned)idx3(g, k, j, i);
  const double dtdV = a.uniform ? a.dtdV : a.dt / cell_dV<C::DIM>(g, k, j, i);
  const double A = face_A<C::DIM, DIR>(g, k, j, i);
  double u[C::NVAR];
  if (a.flags & (SW_INIT_UC | SW_SAVE_UC0)) {
    double u0[C::NVAR];
    cell_p2c<C>(g, a.Vc, p, u0);
    if (a.flags & SW_SAVE_UC0) {
#pragma unroll
      for (int n = 0; n < C::NVAR; n++)
        if (!C::is_ct_field(n)) st32(a.Uc0 + n * g.sv, p, u0[n]);
    }
    if (a.flags & SW_INIT_UC) {
#pragma unroll
      for (int n = 0; n < C::NVAR; n++) u[n] = u0[n];
    } else {
#pragma unroll
      for (int n = 0; n < C::NVAR; n++)
        if (!C::is_ct_field(n)) u[n] = ld32((const double *)a.Uc + n * g.sv, p);
    }
  } else {
    // upre: the cell's accumulated state, loaded by the caller ahead of the Riemann solve
#pragma unroll
    for (int n = 0; n < C::NVAR; n++)
      if (!C::is_ct_field(n)) u[n] = upre ? upre[n] : ld32((const double *)a.Uc + n * g.sv, p);
  }
  // FIN: the stage's begin state (loaded by the caller ahead of the Riemann solve, or
  // here before any store of this cell)
  double u0[C::NVAR];
  if constexpr (FIN) {
    if (a.flags & SW_RK) {
#pragma unroll
      for (int n = 0; n < C::NVAR; n++)
        if (!C::is_ct_field(n))
          u0[n] = u0pre ? u0pre[n] : ld32((const double *)a.Uc0 + n * g.sv, p);
    }
  }
  if constexpr (FIN) {
    // the stage's last sweep (ping-pong schedule): the new state, its RK combination as
    // in AddAndStore, and rho, v and E - kinetic energy of ConsToPrim into VcOut; the
    // pressure and the cell field follow once the faces are updated (k_prs_from_faces)
    double un[C::NVAR], vo[C::NVAR];
#pragma unroll
    for (int n = 0; n < C::NVAR; n++) {
      if (C::is_ct_field(n)) continue;
      const double fp = (double)Fp[n] * A;
      const double fm = (double)Fm[n] * A;
      const double rhs = -dtdV * (fp - fm);
      un[n] = u[n] + rhs;
      if (a.flags & SW_RK) un[n] = a.wc * un[n] + a.w0 * u0[n];
    }
    if (a.flags & SW_WRITE_UC) {
#pragma unroll
      for (int n = 0; n < C::NVAR; n++)
        if (!C::is_ct_field(n)) st32(a.Uc + n * g.sv, p, un[n]);
    }
    cons2prim_part<C>(un, vo);
#pragma unroll
    for (int n = 0; n < C::NVAR; n++)
      if (!C::is_ct_field(n)) st32(a.VcOut + n * g.sv, p, vo[n]);
  } else {
#pragma unroll
    for (int n = 0; n < C::NVAR; n++) {
      if (C::is_ct_field(n)) continue;
      const double fp = (double)Fp[n] * A;
      const double fm = (double)Fm[n] * A;
      const double rhs = -dtdV * (fp - fm);
      st32(a.Uc + n * g.sv, p, u[n] + rhs);
    }
  }
  if constexpr (XTRA) {
    if (g.ntr > 0) tracer_update<C, DIR>(g, a, p, dtdV, A, (double)Fm[C::RHO], (double)Fp[C::RHO]);
  }
  const int ig = DIR == 0 ? i : DIR == 1 ? j : k;
  const double dl = g.dx[DIR][ig];
  const double term = 0.5 * ((double)cmaxp + (double)cmaxm) / (dl);
  double inv = 0.0;
  if (a.flags & SW_INVDT_SET) {
    inv = term;
    if (!(a.flags & SW_NO_INVDT_STORE)) st32(a.InvDt, p, inv);
  } else if (a.flags & SW_INVDT_ADD) {
    inv = ld32((const double *)a.InvDt, p) + term;
    st32(a.InvDt, p, inv);
  }
  return inv;
}

// ------------------------------------------------------------------------------------
// x sweep: one cell per lane.  Lane l reconstructs cell c = i0 - 1 + l once; face c
// takes its left state from lane l-1 (a cross-lane shuffle) and its right state from the
// lane itself, and cell c takes the flux of face c+1 from lane l+1.  A wave therefore
// solves 63 faces and updates the 62 cells between them.
// ------------------------------------------------------------------------------------
// occupancy target of the x sweep: 4 waves per SIMD (<= 128 VGPRs).  The 8-variable MHD
// states spill a little at 128 (2-D; 3-D PPM in FP64: 36 B per lane) but still run faster
// than at 3 waves (3-D PPM HLLD at 512^3: 11.7 against 12.4 ms); the other 3-D
// reconstructions in FP64 spill more (up to 64 B) and stay at 3.
template <class C>
constexpr int x_waves_per_eu() {
  return (C::DIM == 3 && C::ORDER != RECON_PPM) ? 3 : 4;
}
template <class C, typename T, bool FUSED, int EMF, bool XTRA>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(x_waves_per_eu<C>())))
void k_sweep_x(GridDesc g, SweepArgs a, EmfArrays<T> e,
                                                 int nseg) {
  constexpr int NV = C::NVAR;
  constexpr int NW = window<C>();
  constexpr int H = NW / 2;
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int nj = g.end[1] - g.beg[1] + 2 * a.ext[1];
  const int nk = g.end[2] - g.beg[2] + 2 * a.ext[2];
  const bool wave_ok = w < nseg * nj * nk;
  unsigned long long red = 0;
  if (wave_ok) {
    const int seg = w % nseg;
    const int r = w / nseg;
    const int j = g.beg[1] - a.ext[1] + r % nj;
    const int k = g.beg[2] - a.ext[2] + r / nj;
    const int c = g.beg[0] - 1 + seg * 62 + lane;  // this lane's cell; its left face is c-1/2
    const int cc = c <= g.end[0] ? c : g.end[0];     // clamped for the loads
    const unsigned row = (unsigned)idx3(g, k, j, 0);

    // reconstruct the lane's cell: vl = right state of face c, vr = left state of face c+1
    T vl[NV], vr[NV];
    const bool flat = XTRA && recon_flat<C>(a, row + (unsigned)cc);
#pragma unroll
    for (int n = 0; n < NV; n++) {
      T wv[NW];
#pragma unroll
      for (int s = 0; s < NW; s++) wv[s] = (T)ld32(a.Vc + n * g.sv, row + (unsigned)(cc - H + s));
      if constexpr (C::ORDER == RECON_MLP) {
        if (C::MHD && n == C::BX1) { vl[n] = vr[n] = wv[H]; continue; }  // normal field: Vs
        recon_mlp<C, IDIR>(wv, n, a.mlpw, g.sv, row + (unsigned)cc, 1u, vl[n], vr[n]);
      } else {
        recon_cell<C>(wv, n, vl[n], vr[n], recon_dx<C, T>(g, 0, cc), flat);
      }
    }
    T vL[NV], vR[NV];
#pragma unroll
    for (int n = 0; n < NV; n++) {
      vL[n] = __shfl_up(vr[n], 1);
      vR[n] = vl[n];
    }
    if constexpr (C::MHD) {
      vL[C::BX1] = (T)ld32(a.Vs, (unsigned)idxs(g, 0, k, j, cc));
      vR[C::BX1] = vL[C::BX1];
    }
    T F[NV], cmax;
    FaceEmf<T> E;
    const bool wrote = riemann<C, IDIR>(vL, vR, F, cmax, E, EMF,
                                        eos_face<C, T>(g, row + (unsigned)(cc - 1), row + (unsigned)cc));
    const bool face_ok = lane >= 1 && c <= g.end[0];
    const bool own_face = face_ok && (lane <= 62 || c == g.end[0]);
    const unsigned p = row + (unsigned)c;
    if (own_face) store_face_emf<C, IDIR>(e, p, EMF, E);
    if constexpr (!FUSED) {
      if (own_face) {
        if (wrote) {
#pragma unroll
          for (int n = 0; n < NV; n++) st32(a.Flux + n * g.sv, p, (double)F[n]);
        }
        st32(a.cMax, p, (double)cmax);
      }
    } else {
      T Fp[NV];
#pragma unroll
      for (int n = 0; n < NV; n++) Fp[n] = __shfl_down(F[n], 1);
      const T cmaxp = __shfl_down(cmax, 1);
      const bool cell_ok = lane >= 1 && lane <= 62 && c < g.end[0] && j >= g.beg[1] &&
                           j < g.end[1] && k >= g.beg[2] && k < g.end[2];
      if (cell_ok) {
        const double inv = cell_update<C, IDIR, XTRA>(g, a, k, j, c, F, Fp, cmax, cmaxp);
        red = __double_as_longlong(inv);
      }
    }
  }
  if constexpr (FUSED) {
    if (a.flags & SW_DT_REDUCE) block_max_atomic(red, a.dtmax);
  }
}

// ------------------------------------------------------------------------------------
// y / z sweeps: marching along DIR with the stencil window in registers
// ------------------------------------------------------------------------------------
struct MarchGeom {
  int ilo, ihi;   // x range (interior +- extension)
  int tlo, thi;   // range of the other transverse direction
  int ts;         // cells per segment along DIR
  int nseg;       // segments along DIR
  int nxb;        // 64-wide x blocks
};

// The stencil rows consecutive marching steps share are carried in LDS for the 3-cell
// window (PLM, learned limiter); PPM keeps its 5-row window in registers.
template <class C>
constexpr bool march_lds_window() {
  return window<C>() == 3;
}
// 3 waves per SIMD fit the 2-D march of the 6-variable MHD state without spills; the 3-D
// one and the 8-variable 2-D one would spill, so they stay at 2
template <class C>
constexpr int march_waves_per_eu() {
  return (march_lds_window<C>() && C::DIM == 2 && C::NVAR < 8) ? 3 : 2;
}
template <class C, typename T, int DIR, bool FUSED, int EMF, bool XTRA, bool FIN = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(march_waves_per_eu<C>())))
void k_sweep_march(GridDesc g, SweepArgs a, EmfArrays<T> e,
                                                     MarchGeom m) {
  static_assert(DIR == JDIR || DIR == KDIR, "marching sweep is for y and z");
  constexpr int NV = C::NVAR;
  constexpr int NW = window<C>();
  constexpr int H = NW / 2;
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int ntr = m.thi - m.tlo;
  unsigned long long red = 0;
  if (w < m.nxb * ntr * m.nseg) {
    const int xb = w % m.nxb;
    const int rest = w / m.nxb;
    const int seg = rest % m.nseg;
    const int tr = m.tlo + rest / m.nseg;
    const int i = m.ilo + xb * 64 + lane;
    const bool lane_ok = i < m.ihi;
    const int ic = lane_ok ? i : m.ihi - 1;
    const int c0 = g.beg[DIR] + seg * m.ts;
    const int c1 = min(c0 + m.ts, g.end[DIR]);
    // stride along DIR and base offset of the column
    const unsigned sd = (DIR == JDIR) ? g.nt[0] : (unsigned)g.nt[0] * g.nt[1];
    const int jj = (DIR == JDIR) ? 0 : tr;  // j for the z sweep
    const int kk = (DIR == JDIR) ? tr : 0;  // k for the y sweep
    const unsigned col = (unsigned)idx3(g, kk, jj, ic);  // index with DIR coordinate = 0
    const bool rhs_lane = lane_ok && i >= g.beg[0] && i < g.end[0] &&
                          ((DIR == JDIR) ? (tr >= g.beg[2] && tr < g.end[2])
                                         : (tr >= g.beg[1] && tr < g.end[1]));
    // Vs normal component: stride along DIR in the Vs array
    const unsigned vsd = (DIR == JDIR) ? g.vsn[0] : (unsigned)g.vsn[0] * g.vsn[1];
    const unsigned vcol = (unsigned)idxs(g, 0, kk, jj, ic);
    const double *VsDir = a.Vs + DIR * g.svs;  // the normal component

    constexpr bool CARRY = march_lds_window<C>();
    // window rows f-1, f of thread t at lds[(row & 1) * NV + n][t]
    __shared__ T lds[CARRY ? 2 * NV : 1][256];
    const int tid = threadIdx.x;
    T win[NW][NV];
#pragma unroll
    for (int s = 0; s < NW; s++)
#pragma unroll
      for (int n = 0; n < NV; n++) {
        win[s][n] = (T)ld32(a.Vc + n * g.sv, col + (unsigned)(c0 - 1 - H + s) * sd);
        if constexpr (CARRY) {
          if (s >= 1) lds[((c0 - 1 - H + s) & 1) * NV + n][tid] = win[s][n];
        }
      }
    // right-face value of cell c0-1: the left state of the first face
    T carry[NV], Fprev[NV], cprev = 0;
#pragma unroll
    for (int n = 0; n < NV; n++) Fprev[n] = 0;
#pragma unroll
    for (int n = 0; n < NV; n++) {
      T wv[NW];
#pragma unroll
      for (int s = 0; s < NW; s++) wv[s] = win[s][n];
      T dummy;
      if constexpr (C::ORDER == RECON_MLP) {
        if (C::MHD && n == C::template BXn<DIR>()) { carry[n] = wv[H]; continue; }
        recon_mlp<C, DIR>(wv, n, a.mlpw, g.sv, col + (unsigned)(c0 - 1) * sd, sd, dummy, carry[n]);
      } else {
        recon_cell<C>(wv, n, dummy, carry[n], recon_dx<C, T>(g, DIR, c0 - 1),
                      XTRA && recon_flat<C>(a, col + (unsigned)(c0 - 1) * sd));
      }
    }
    for (int f = c0; f <= c1; f++) {
      // shift the window to cell f and load its leading row
      if constexpr (CARRY) {
#pragma unroll
        for (int n = 0; n < NV; n++) {
          win[0][n] = lds[((f - 1) & 1) * NV + n][tid];
          win[1][n] = lds[(f & 1) * NV + n][tid];
          win[2][n] = (T)ld32(a.Vc + n * g.sv, col + (unsigned)(f + 1) * sd);
          lds[((f + 1) & 1) * NV + n][tid] = win[2][n];  // replaces row f-1, read above
        }
      } else {
#pragma unroll
        for (int s = 0; s < NW - 1; s++)
#pragma unroll
          for (int n = 0; n < NV; n++) win[s][n] = win[s + 1][n];
#pragma unroll
        for (int n = 0; n < NV; n++)
          win[NW - 1][n] = (T)ld32(a.Vc + n * g.sv, col + (unsigned)(f + H) * sd);
      }
      // FIN: the begin state of cell f - 1 (updated at the end of this step), loaded now
      // so that its latency overlaps the reconstruction and the Riemann solve
      double u0pre[NV];
      if constexpr (FIN) {
        if ((a.flags & SW_RK) && f > c0 && rhs_lane) {
#pragma unroll
          for (int n = 0; n < NV; n++)
            if (!C::is_ct_field(n))
              u0pre[n] = ld32((const double *)a.Uc0 + n * g.sv, col + (unsigned)(f - 1) * sd);
        }
      }
      // the 3-D sweeps without per-cell options, except Roe's and the final FP64 one: the
      // accumulated state of cell f - 1 (the register room this needs is there only in
      // those; the FP32 sweeps have it at their 3 waves per SIMD)
      double upre[NV];
      constexpr bool PRE_U = FUSED && !XTRA && C::DIM == 3 && C::SOLVER != SOLVER_ROE &&
                             (!FIN || std::is_same<T, float>::value);
      if constexpr (PRE_U) {
        if (f > c0 && rhs_lane) {
#pragma unroll
          for (int n = 0; n < NV; n++)
            if (!C::is_ct_field(n))
              upre[n] = ld32((const double *)a.Uc + n * g.sv, col + (unsigned)(f - 1) * sd);
        }
      }
      T vL[NV], vR[NV], vr_next[NV];
      const bool flat = XTRA && recon_flat<C>(a, col + (unsigned)f * sd);
#pragma unroll
      for (int n = 0; n < NV; n++) {
        T wv[NW];
#pragma unroll
        for (int s = 0; s < NW; s++) wv[s] = win[s][n];
        if constexpr (C::ORDER == RECON_MLP) {
          if (C::MHD && n == C::template BXn<DIR>()) {
            vR[n] = vr_next[n] = wv[H];
          } else {
            recon_mlp<C, DIR>(wv, n, a.mlpw, g.sv, col + (unsigned)f * sd, sd, vR[n], vr_next[n]);
          }
        } else {
          recon_cell<C>(wv, n, vR[n], vr_next[n], recon_dx<C, T>(g, DIR, f), flat);
        }
        vL[n] = carry[n];
      }
      if constexpr (C::MHD) {
        constexpr int BXn = C::template BXn<DIR>();
        vL[BXn] = (T)ld32(VsDir, vcol + (unsigned)f * vsd);
        vR[BXn] = vL[BXn];
      }
      T F[NV], cmax;
      FaceEmf<T> E;
      const bool wrote = riemann<C, DIR>(vL, vR, F, cmax, E, EMF,
                                         eos_face<C, T>(g, col + (unsigned)(f - 1) * sd, col + (unsigned)f * sd));
      const unsigned p = col + (unsigned)f * sd;
      const bool own_face = lane_ok && (f < c1 || c1 == g.end[DIR]);
      if (own_face) store_face_emf<C, DIR>(e, p, EMF, E);
      if constexpr (!FUSED) {
        if (own_face) {
          if (wrote) {
#pragma unroll
            for (int n = 0; n < NV; n++) st32(a.Flux + n * g.sv, p, (double)F[n]);
          }
          st32(a.cMax, p, (double)cmax);
        }
      } else {
        if (f > c0 && rhs_lane) {
          const int cf = f - 1;
          const int ci = ic, cj = (DIR == JDIR) ? cf : jj, ck = (DIR == JDIR) ? kk : cf;
          const double inv = cell_update<C, DIR, XTRA, FIN>(g, a, ck, cj, ci, Fprev, F, cprev,
                                                            cmax, u0pre, PRE_U ? upre : nullptr);
          const unsigned long long b = __double_as_longlong(inv);
          red = b > red ? b : red;
        }
      }
#pragma unroll
      for (int n = 0; n < NV; n++) {
        carry[n] = vr_next[n];
        Fprev[n] = F[n];
      }
      cprev = cmax;
    }
  }
  if constexpr (FUSED) {
    if (a.flags & SW_DT_REDUCE) block_max_atomic(red, a.dtmax);
  }
}

// ------------------------------------------------------------------------------------
// CalcRightHandSide from a materialised FluxRiemann (flux-only mode):
// "Correct Flux" then "CalcRightHandSide" (calcRightHandSide.hpp:556-596).
// ------------------------------------------------------------------------------------
template <class C, int DIR>
__global__ __launch_bounds__(256) void k_correct_flux(GridDesc g, double *Flux) {
  constexpr int io = DIR == 0, jo = DIR == 1, ko = DIR == 2;
  const int ni = g.end[0] + io - g.beg[0], nj = g.end[1] + jo - g.beg[1];
  const int nk = g.end[2] + ko - g.beg[2];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ni * nj * nk) return;
  const int i = g.beg[0] + t % ni;
  const int j = g.beg[1] + (t / ni) % nj;
  const int k = g.beg[2] + t / ((long long)ni * nj);
  const double A = face_A<C::DIM, DIR>(g, k, j, i);
  const long long p = idx3(g, k, j, i);
#pragma unroll
  for (int n = 0; n < C::NVAR; n++) Flux[n * g.sv + p] = Flux[n * g.sv + p] * A;
}

template <class C, int DIR>
__global__ __launch_bounds__(256) void k_rhs(GridDesc g, const double *Flux, const double *cMax,
                                             double *Uc, double *InvDt, double dt) {
  constexpr int io = DIR == 0, jo = DIR == 1, ko = DIR == 2;
  const int ni = g.end[0] - g.beg[0], nj = g.end[1] - g.beg[1], nk = g.end[2] - g.beg[2];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ni * nj * nk) return;
  const int i = g.beg[0] + t % ni;
  const int j = g.beg[1] + (t / ni) % nj;
  const int k = g.beg[2] + t / ((long long)ni * nj);
  const long long p = idx3(g, k, j, i);
  const long long pp = idx3(g, k + ko, j + jo, i + io);
  const double dtdV = dt / cell_dV<C::DIM>(g, k, j, i);
  const int ig = io * i + jo * j + ko * k;
  const double dl = g.dx[DIR][ig];
  InvDt[p] = InvDt[p] + 0.5 * (cMax[pp] + cMax[p]) / (dl);
#pragma unroll
  for (int n = 0; n < C::NVAR; n++) {
    if (C::is_ct_field(n)) continue;
    const double rhs = -dtdV * (Flux[n * g.sv + pp] - Flux[n * g.sv + p]);
    Uc[n * g.sv + p] = Uc[n * g.sv + p] + rhs;
  }
}

// ------------------------------------------------------------------------------------
// Constrained transport: corner EMFs (calcCornerEmf.hpp, calcRiemannEmf.hpp)
// ------------------------------------------------------------------------------------
// cell-centred E = -v x B of one cell (CalcCellCenteredEMF, calcCornerEmf.hpp:84-126)
template <class C, typename T>
__device__ __forceinline__ void cell_emf(const GridDesc &g, const double *Vc, int k, int j,
                                         int i, T &E1, T &E2, T &E3) {
  const long long p = idx3(g, k, j, i);
  T vx1 = (T)Vc[C::MX1 * g.sv + p], vx2 = 0, vx3 = 0;
  T bx1 = (T)Vc[C::BX1 * g.sv + p], bx2 = 0, bx3 = 0;
  if constexpr (C::COMP >= 2) {
    vx2 = (T)Vc[C::MX2 * g.sv + p];
    bx2 = (T)Vc[C::BX2 * g.sv + p];
  }
  if constexpr (C::COMP >= 3) {
    vx3 = (T)Vc[C::MX3 * g.sv + p];
    bx3 = (T)Vc[C::BX3 * g.sv + p];
  }
  E1 = (vx3 * bx2 - vx2 * bx3);
  E2 = (vx1 * bx3 - vx3 * bx1);
  E3 = (vx2 * bx1 - vx1 * bx2);
}

// The three uct_contact edge EMFs of a 3-D block at edge index (k, j, i): ex at
// (i, j-1/2, k-1/2), ey at (i-1/2, j, k-1/2), ez at (i-1/2, j-1/2, k)
// (CalcContactAverage, calcCornerEmf.hpp:228-282).  E1/E2/E3[dk][dj][di] are the cell
// EMFs of cells (k-1+dk, j-1+dj, i-1+di); the face data follow the argument names
// (km = k-1, jm = j-1, im = i-1).  Shared by k_corner_emf and k_ct3d, so both compute the
// same values with the same operations.
template <typename T>
struct ContactFaces {
  T svx, svx_jm, svx_km, svy, svy_im, svy_km, svz, svz_jm, svz_im;
  T ezi, ezi_jm, ezj, ezj_im, exj, exj_km, exk, exk_jm, eyi, eyi_km, eyk, eyk_im;
};
template <typename T>
__device__ __forceinline__ void contact_edges3(const ContactFaces<T> &f, const T (&E1)[2][2][2],
                                               const T (&E2)[2][2][2], const T (&E3)[2][2][2],
                                               T &ex, T &ey, T &ez) {
  const T ez_l2 = (1 - f.svx_jm) * (f.ezj - E3[1][0][1]) + (f.svx_jm) * (f.ezj_im - E3[1][0][0]);
  const T ez_r2 = (1 - f.svx) * (f.ezj - E3[1][1][1]) + (f.svx) * (f.ezj_im - E3[1][1][0]);
  const T ez_l1 = (1 - f.svy_im) * (f.ezi - E3[1][1][0]) + (f.svy_im) * (f.ezi_jm - E3[1][0][0]);
  const T ez_r1 = (1 - f.svy) * (f.ezi - E3[1][1][1]) + (f.svy) * (f.ezi_jm - E3[1][0][1]);
  ez = T(0.25) * (ez_l2 + ez_r2 + ez_l1 + ez_r1 + f.ezi + f.ezi_jm + f.ezj + f.ezj_im);
  const T ex_l3 = (1 - f.svy_km) * (f.exk - E1[0][1][1]) + (f.svy_km) * (f.exk_jm - E1[0][0][1]);
  const T ex_r3 = (1 - f.svy) * (f.exk - E1[1][1][1]) + (f.svy) * (f.exk_jm - E1[1][0][1]);
  const T ex_l2 = (1 - f.svz_jm) * (f.exj - E1[1][0][1]) + (f.svz_jm) * (f.exj_km - E1[0][0][1]);
  const T ex_r2 = (1 - f.svz) * (f.exj - E1[1][1][1]) + (f.svz) * (f.exj_km - E1[0][1][1]);
  ex = T(0.25) * (ex_l3 + ex_r3 + ex_l2 + ex_r2 + f.exj + f.exj_km + f.exk + f.exk_jm);
  const T ey_l3 = (1 - f.svx_km) * (f.eyk - E2[0][1][1]) + (f.svx_km) * (f.eyk_im - E2[0][1][0]);
  const T ey_r3 = (1 - f.svx) * (f.eyk - E2[1][1][1]) + (f.svx) * (f.eyk_im - E2[1][1][0]);
  const T ey_l1 = (1 - f.svz_im) * (f.eyi - E2[1][1][0]) + (f.svz_im) * (f.eyi_km - E2[0][1][0]);
  const T ey_r1 = (1 - f.svz) * (f.eyi - E2[1][1][1]) + (f.svz) * (f.eyi_km - E2[0][1][1]);
  ey = T(0.25) * (ey_l3 + ey_r3 + ey_l1 + ey_r1 + f.eyi + f.eyi_km + f.eyk + f.eyk_im);
}

// One edge term of CalcRiemannAverage (calcRiemannEmf.hpp:62-428): B and the upwinded
// transverse velocity reconstructed (MC limiter, or PPM for ORDER 4) from the left and
// right cells of the edge.  b[s], v[s] hold the face values at offset s-3 from the
// edge's left cell (s = 0..6 covers -3..3).
template <class C, typename T>
__device__ __forceinline__ T edge_term(const T *b, const T *v, T aL, T aR, T dL, T dR,
                                       bool plus) {
  T bL, bR, vL, vR, d;
  if constexpr (C::ORDER == RECON_PPM) {
    ppm_states(b[1], b[2], b[3], b[4], b[5], d, bL);
    ppm_states(b[2], b[3], b[4], b[5], b[6], bR, d);
    ppm_states(v[1], v[2], v[3], v[4], v[5], d, vL);
    ppm_states(v[2], v[3], v[4], v[5], v[6], vR, d);
  } else {
    d = mclim(b[4] - b[3], b[3] - b[2]);
    bL = b[3] + T(0.5) * d;
    d = mclim(b[5] - b[4], b[4] - b[3]);
    bR = b[4] - T(0.5) * d;
    d = mclim(v[4] - v[3], v[3] - v[2]);
    vL = v[3] + T(0.5) * d;
    d = mclim(v[5] - v[4], v[4] - v[3]);
    vR = v[4] - T(0.5) * d;
  }
  const T phi = dR * bR - dL * bL;
  return plus ? (aL * vL * bL + aR * vR * bR) + phi : (aL * vL * bL + aR * vR * bR) - phi;
}

template <class C, typename T>
__global__ __launch_bounds__(256) void k_corner_emf(GridDesc g, const double *Vc,
                                                    const double *Vs, EmfArrays<T> e, int emf) {
  constexpr int D3 = C::DIM == 3;
  const int ni = g.end[0] + 1 - g.beg[0], nj = g.end[1] + 1 - g.beg[1];
  const int nk = g.end[2] + D3 - g.beg[2];
  // edges in i, j, k order; the rows and planes below an edge are read by this block and
  // the ones just before it, kept on one XCD
  const long long t = (long long)xcd_contiguous_block() * blockDim.x + threadIdx.x;
  if (t >= (long long)ni * nj * nk) return;
  const int i = g.beg[0] + t % ni;
  const int j = g.beg[1] + (t / ni) % nj;
  const int k = g.beg[2] + t / ((long long)ni * nj);
  const long long p = idx3(g, k, j, i);
  auto at = [&](const T *arr, int kk, int jj, int ii) { return arr[idx3(g, kk, jj, ii)]; };
  auto sva = [&](const unsigned char *arr, int kk, int jj, int ii) {
    return sv_val<T>(arr[idx3(g, kk, jj, ii)]);
  };
  if (emf == EMF_ARITHMETIC || emf == EMF_UCT0) {
    // CalcArithmeticAverage (:62-78); uct0 first rewrites the face EMFs (:150-172)
    const T w = T(0.25);
    auto mod = [&](const T *arr, int kk, int jj, int ii, int comp, int dk, int dj, int di) {
      T v = at(arr, kk, jj, ii);
      if (emf == EMF_UCT0) {
        T a1, a2, a3, b1, b2, b3;
        cell_emf<C, T>(g, Vc, kk - dk, jj - dj, ii - di, a1, a2, a3);
        cell_emf<C, T>(g, Vc, kk, jj, ii, b1, b2, b3);
        const T ea = comp == 1 ? a1 : comp == 2 ? a2 : a3;
        const T eb = comp == 1 ? b1 : comp == 2 ? b2 : b3;
        v *= T(2);
        v -= T(0.5) * (ea + eb);
      }
      return v;
    };
    if constexpr (C::DIM == 3) {
      e.ex[p] = w * (mod(e.exj, k, j, i, 1, 0, 1, 0) + mod(e.exj, k - 1, j, i, 1, 0, 1, 0) +
                     mod(e.exk, k, j, i, 1, 1, 0, 0) + mod(e.exk, k, j - 1, i, 1, 1, 0, 0));
      e.ey[p] = w * (mod(e.eyi, k, j, i, 2, 0, 0, 1) + mod(e.eyi, k - 1, j, i, 2, 0, 0, 1) +
                     mod(e.eyk, k, j, i, 2, 1, 0, 0) + mod(e.eyk, k, j, i - 1, 2, 1, 0, 0));
    }
    e.ez[p] = w * (mod(e.ezi, k, j, i, 3, 0, 0, 1) + mod(e.ezi, k, j - 1, i, 3, 0, 0, 1) +
                   mod(e.ezj, k, j, i, 3, 0, 1, 0) + mod(e.ezj, k, j, i - 1, 3, 0, 1, 0));
  } else if (emf == EMF_UCT_CONTACT) {
    // CalcContactAverage, calcCornerEmf.hpp:228-282
    T E1[2][2][2], E2[2][2][2], E3[2][2][2];  // [dk][dj][di] for cells (k-1+dk, j-1+dj, i-1+di)
#pragma unroll
    for (int dk = 0; dk < 1 + D3; dk++)
#pragma unroll
      for (int dj = 0; dj < 2; dj++)
#pragma unroll
        for (int di = 0; di < 2; di++)
          cell_emf<C, T>(g, Vc, k - D3 + dk, j - 1 + dj, i - 1 + di, E1[dk][dj][di],
                         E2[dk][dj][di], E3[dk][dj][di]);
    if constexpr (C::DIM == 3) {
      ContactFaces<T> f;
      f.svx = sva(e.svx, k, j, i);
      f.svx_jm = sva(e.svx, k, j - 1, i);
      f.svx_km = sva(e.svx, k - 1, j, i);
      f.svy = sva(e.svy, k, j, i);
      f.svy_im = sva(e.svy, k, j, i - 1);
      f.svy_km = sva(e.svy, k - 1, j, i);
      f.svz = sva(e.svz, k, j, i);
      f.svz_jm = sva(e.svz, k, j - 1, i);
      f.svz_im = sva(e.svz, k, j, i - 1);
      f.ezi = at(e.ezi, k, j, i);
      f.ezi_jm = at(e.ezi, k, j - 1, i);
      f.ezj = at(e.ezj, k, j, i);
      f.ezj_im = at(e.ezj, k, j, i - 1);
      f.exj = at(e.exj, k, j, i);
      f.exj_km = at(e.exj, k - 1, j, i);
      f.exk = at(e.exk, k, j, i);
      f.exk_jm = at(e.exk, k, j - 1, i);
      f.eyi = at(e.eyi, k, j, i);
      f.eyi_km = at(e.eyi, k - 1, j, i);
      f.eyk = at(e.eyk, k, j, i);
      f.eyk_im = at(e.eyk, k, j, i - 1);
      T ex, ey, ez;
      contact_edges3(f, E1, E2, E3, ex, ey, ez);
      e.ex[p] = ex;
      e.ey[p] = ey;
      e.ez[p] = ez;
    } else {
      const int K = D3;  // index of cell k in the [dk] axis
      const T ez_l2 = (1 - sva(e.svx, k, j - 1, i)) * (at(e.ezj, k, j, i) - E3[K][0][1]) +
                      (sva(e.svx, k, j - 1, i)) * (at(e.ezj, k, j, i - 1) - E3[K][0][0]);
      const T ez_r2 = (1 - sva(e.svx, k, j, i)) * (at(e.ezj, k, j, i) - E3[K][1][1]) +
                      (sva(e.svx, k, j, i)) * (at(e.ezj, k, j, i - 1) - E3[K][1][0]);
      const T ez_l1 = (1 - sva(e.svy, k, j, i - 1)) * (at(e.ezi, k, j, i) - E3[K][1][0]) +
                      (sva(e.svy, k, j, i - 1)) * (at(e.ezi, k, j - 1, i) - E3[K][0][0]);
      const T ez_r1 = (1 - sva(e.svy, k, j, i)) * (at(e.ezi, k, j, i) - E3[K][1][1]) +
                      (sva(e.svy, k, j, i)) * (at(e.ezi, k, j - 1, i) - E3[K][0][1]);
      e.ez[p] = T(0.25) * (ez_l2 + ez_r2 + ez_l1 + ez_r1 + at(e.ezi, k, j, i) +
                           at(e.ezi, k, j - 1, i) + at(e.ezj, k, j, i) + at(e.ezj, k, j, i - 1));
    }
  } else {
    // uct_hll / uct_hlld: CalcRiemannAverage
    // stencil slots used: MC limiter -1..2, PPM -2..3 around the edge's left cell
    constexpr int SLO = C::ORDER == RECON_PPM ? 1 : 2;
    constexpr int SHI = C::ORDER == RECON_PPM ? 6 : 5;
    const int im = i - 1, jm = j - 1, km = k - 1;
    T b[7], v[7], aL, aR, dL, dR;
    auto vs = [&](int n, int kk, int jj, int ii) { return (T)Vs[idxs(g, n, kk, jj, ii)]; };
    // ez from x faces: B = BX2s along x at y-face j; v = ezj along x
    aL = T(0.5) * (at(e.axL, k, jm, i) + at(e.axL, k, jm + 1, i));
    aR = T(0.5) * (at(e.axR, k, jm, i) + at(e.axR, k, jm + 1, i));
    dL = T(0.5) * (at(e.dxL, k, jm, i) + at(e.dxL, k, jm + 1, i));
    dR = T(0.5) * (at(e.dxR, k, jm, i) + at(e.dxR, k, jm + 1, i));
#pragma unroll
    for (int s = SLO; s <= SHI; s++) {
      b[s] = vs(1, k, j, im + s - 3);
      v[s] = at(e.ezj, k, j, im + s - 3);
    }
    T ezv = edge_term<C>(b, v, aL, aR, dL, dR, true);
    T eyv = 0, exv = 0;
    if constexpr (C::DIM == 3) {
      aL = T(0.5) * (at(e.axL, km, j, i) + at(e.axL, km + 1, j, i));
      aR = T(0.5) * (at(e.axR, km, j, i) + at(e.axR, km + 1, j, i));
      dL = T(0.5) * (at(e.dxL, km, j, i) + at(e.dxL, km + 1, j, i));
      dR = T(0.5) * (at(e.dxR, km, j, i) + at(e.dxR, km + 1, j, i));
#pragma unroll
      for (int s = SLO; s <= SHI; s++) {
        b[s] = vs(2, k, j, im + s - 3);
        v[s] = at(e.eyk, k, j, im + s - 3);
      }
      eyv = edge_term<C>(b, v, aL, aR, dL, dR, false);
      aL = T(0.5) * (at(e.ayL, km, j, i) + at(e.ayL, km + 1, j, i));
      aR = T(0.5) * (at(e.ayR, km, j, i) + at(e.ayR, km + 1, j, i));
      dL = T(0.5) * (at(e.dyL, km, j, i) + at(e.dyL, km + 1, j, i));
      dR = T(0.5) * (at(e.dyR, km, j, i) + at(e.dyR, km + 1, j, i));
#pragma unroll
      for (int s = SLO; s <= SHI; s++) {
        b[s] = vs(2, k, jm + s - 3, i);
        v[s] = at(e.exk, k, jm + s - 3, i);
      }
      exv = edge_term<C>(b, v, aL, aR, dL, dR, true);
    }
    aL = T(0.5) * (at(e.ayL, k, j, im) + at(e.ayL, k, j, im + 1));
    aR = T(0.5) * (at(e.ayR, k, j, im) + at(e.ayR, k, j, im + 1));
    dL = T(0.5) * (at(e.dyL, k, j, im) + at(e.dyL, k, j, im + 1));
    dR = T(0.5) * (at(e.dyR, k, j, im) + at(e.dyR, k, j, im + 1));
#pragma unroll
    for (int s = SLO; s <= SHI; s++) {
      b[s] = vs(0, k, jm + s - 3, i);
      v[s] = at(e.ezi, k, jm + s - 3, i);
    }
    ezv += edge_term<C>(b, v, aL, aR, dL, dR, false);
    if constexpr (C::DIM == 3) {
      aL = T(0.5) * (at(e.azL, k, j, im) + at(e.azL, k, j, im + 1));
      aR = T(0.5) * (at(e.azR, k, j, im) + at(e.azR, k, j, im + 1));
      dL = T(0.5) * (at(e.dzL, k, j, im) + at(e.dzL, k, j, im + 1));
      dR = T(0.5) * (at(e.dzR, k, j, im) + at(e.dzR, k, j, im + 1));
#pragma unroll
      for (int s = SLO; s <= SHI; s++) {
        b[s] = vs(0, km + s - 3, j, i);
        v[s] = at(e.eyi, km + s - 3, j, i);
      }
      eyv += edge_term<C>(b, v, aL, aR, dL, dR, true);
      aL = T(0.5) * (at(e.azL, k, jm, i) + at(e.azL, k, jm + 1, i));
      aR = T(0.5) * (at(e.azR, k, jm, i) + at(e.azR, k, jm + 1, i));
      dL = T(0.5) * (at(e.dzL, k, jm, i) + at(e.dzL, k, jm + 1, i));
      dR = T(0.5) * (at(e.dzR, k, jm, i) + at(e.dzR, k, jm + 1, i));
#pragma unroll
      for (int s = SLO; s <= SHI; s++) {
        b[s] = vs(1, km + s - 3, j, i);
        v[s] = at(e.exj, km + s - 3, j, i);
      }
      exv += edge_term<C>(b, v, aL, aR, dL, dR, false);
      e.ex[p] = exv;
      e.ey[p] = eyv;
    }
    e.ez[p] = ezv;
  }
}

// EvolveMagField, evolveMagField.hpp:53-132 (Cartesian).  mode 0: Vs += rhs;
// mode 1: Vs0 = Vs then Vs += rhs (stage 0 of a multi-stage integrator);
// mode 2: Vs = wc*(Vs + rhs) + w0*Vs0 (AddAndStore of a later stage).
template <class C, typename T>
__global__ __launch_bounds__(256) void k_evolve_b(GridDesc g, double *Vs, double *Vs0,
                                                  EmfArrays<T> e, double dt, int mode, double wc,
                                                  double w0) {
  constexpr int D3 = C::DIM == 3;
  const int ni = g.end[0] + 1 - g.beg[0], nj = g.end[1] + 1 - g.beg[1];
  const int nk = g.end[2] + D3 - g.beg[2];
  // faces in i, j, k order; the edges of the next row and plane stay on one XCD
  const long long t = (long long)xcd_contiguous_block() * blockDim.x + threadIdx.x;
  if (t >= (long long)ni * nj * nk) return;
  const int i = g.beg[0] + t % ni;
  const int j = g.beg[1] + (t / ni) % nj;
  const int k = g.beg[2] + t / ((long long)ni * nj);
  auto E = [&](const T *arr, int kk, int jj, int ii) { return (double)arr[idx3(g, kk, jj, ii)]; };
  double rhsx1 = 0.0 - dt / g.dx[1][j] * (E(e.ez, k, j + 1, i) - E(e.ez, k, j, i));
  double rhsx2 = dt / g.dx[0][i] * (E(e.ez, k, j, i + 1) - E(e.ez, k, j, i));
  double rhsx3 = 0.0;
  if constexpr (D3) {
    rhsx1 = rhsx1 + dt / g.dx[2][k] * (E(e.ey, k + 1, j, i) - E(e.ey, k, j, i));
    rhsx2 = rhsx2 - dt / g.dx[2][k] * (E(e.ex, k + 1, j, i) - E(e.ex, k, j, i));
    rhsx3 = -dt / g.dx[0][i] * (E(e.ey, k, j, i + 1) - E(e.ey, k, j, i)) +
            dt / g.dx[1][j] * (E(e.ex, k, j + 1, i) - E(e.ex, k, j, i));
  }
  double rhs[3] = {rhsx1, rhsx2, rhsx3};
#pragma unroll
  for (int n = 0; n < C::DIM; n++) {
    const long long q = idxs(g, n, k, j, i);
    const double old = Vs[q];
    const double upd = old + rhs[n];
    if (mode == 1) Vs0[q] = old;
    Vs[q] = (mode == 2) ? wc * upd + w0 * Vs0[q] : upd;
  }
}

// CalcCornerEMF (uct_contact) and EvolveMagField of a 3-D block in one kernel, for the
// stage path (the reference-granularity entry points keep k_corner_emf + k_evolve_b).
// A block of 8 waves owns a tile of 63 face columns (i) x 7 face rows (j) and marches
// along k over a segment of face planes.  Lane l of wave w computes the three edge EMFs at
// edge (k, j0 + w, i0 + l) of each plane k with the arithmetic of k_corner_emf
// (contact_edges3), so the tile's 64 x 8 edges cover its faces and the one halo column
// and row they need; the edges of plane k stay in registers (own column), reach the
// neighbouring column by a shuffle and the neighbouring row through LDS (three plane
// buffers, one barrier per plane), and the face data and cell EMFs of plane k - 1 are
// the previous step's plane k.  After plane k, the faces of plane k - 1 are updated with
// the arithmetic of k_evolve_b (mode / wc / w0 as there), in place: every face has one
// owner and no other block reads Vs.  A lane loads the old values (and RK begin values)
// of the faces it updates in a step at the top of that step, so their latency overlaps
// the edge work and the barrier.  Only faces the next stage reads are updated: each
// face field over [beg, end] along its own direction and [beg, end) across; the edge
// arrays ex, ey, ez are not written.
// Cell EMFs of one cell from 32-bit offsets (the arithmetic of cell_emf).
template <class C, typename T>
__device__ __forceinline__ void cell_emf32(const double *Vc, long long sv, unsigned p, T &E1,
                                           T &E2, T &E3) {
  const T vx1 = (T)ld32(Vc + C::MX1 * sv, p), vx2 = (T)ld32(Vc + C::MX2 * sv, p);
  const T vx3 = (T)ld32(Vc + C::MX3 * sv, p);
  const T bx1 = (T)ld32(Vc + C::BX1 * sv, p), bx2 = (T)ld32(Vc + C::BX2 * sv, p);
  const T bx3 = (T)ld32(Vc + C::BX3 * sv, p);
  E1 = (vx3 * bx2 - vx2 * bx3);
  E2 = (vx1 * bx3 - vx3 * bx1);
  E3 = (vx2 * bx1 - vx1 * bx2);
}

// One face-field update of k_evolve_b's arithmetic (mode 0: Vs += rhs; 1: Vs0 = Vs, then
// Vs += rhs; 2: Vs = wc (Vs + rhs) + w0 Vs0).
__device__ __forceinline__ void evolve_face(double *Vs, double *Vs0, unsigned q, double rhs,
                                            int mode, double wc, double w0) {
  const double old = ld32(Vs, q);
  const double upd = old + rhs;
  if (mode == 1) st32(Vs0, q, old);
  st32(Vs, q, (mode == 2) ? wc * upd + w0 * ld32(Vs0, q) : upd);
}
// The same with the face's old value (and, in mode 2, its RK begin value) loaded earlier.
__device__ __forceinline__ void evolve_face_pre(double *Vs, double *Vs0, unsigned q, double old,
                                                double old0, double rhs, int mode, double wc,
                                                double w0) {
  const double upd = old + rhs;
  if (mode == 1) st32(Vs0, q, old);
  st32(Vs, q, (mode == 2) ? wc * upd + w0 * old0 : upd);
}

template <class C, typename T>
__global__ __launch_bounds__(512) void k_ct3d(GridDesc g, const double *Vc, double *Vs,
                                              double *Vs0, EmfArrays<T> e, double dt, int mode,
                                              double wc, double w0, int ntj, int nseg,
                                              int kseg) {
  static_assert(C::MHD && C::DIM == 3 && C::COMP == 3, "k_ct3d is the 3-D MHD CT");
  constexpr int TI = 63, TJ = 7;
  __shared__ T lez[3][8][64], lex[3][8][64];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  // each XCD takes a contiguous run of (tile, segment) indices, so that neighbouring
  // tiles, which read each other's halo rows and columns, share an L2
  const int blk = xcd_contiguous_block();
  const int seg = blk % nseg;
  const int tile = blk / nseg;
  const int ti = tile / ntj, tj = tile % ntj;
  const int i = g.beg[0] + TI * ti + l, j = g.beg[1] + TJ * tj + w;
  const int ks = g.beg[2] + kseg * seg;
  const int ke = min(ks + kseg, g.end[2] + 1);  // face planes [ks, ke)
  const int klast = min(ke, g.end[2]);          // last edge plane
  const bool edge_ok = i <= g.end[0] && j <= g.end[1];
  const unsigned sy = (unsigned)g.nt[0], sz = (unsigned)g.nt[0] * (unsigned)g.nt[1];
  const unsigned vy = (unsigned)g.vsn[0], vz = (unsigned)g.vsn[0] * (unsigned)g.vsn[1];
  double *Vs1 = Vs + g.svs, *Vs2 = Vs + 2 * g.svs;
  double *Vs01 = Vs0 + g.svs, *Vs02 = Vs0 + 2 * g.svs;
  // plane k - 1 data carried from the previous step
  T E1m0 = 0, E1m1 = 0, E2m0 = 0, E2m1 = 0;  // E1 of cells (k-1, j-1+dj, i), E2 of (k-1, j, i-1+di)
  T svx_m = 0, svy_m = 0, exj_m = 0, eyi_m = 0;
  if (edge_ok) {
    const unsigned p = (unsigned)idx3(g, ks - 1, j, i);
    T a1, a2, a3;
    cell_emf32<C, T>(Vc, g.sv, p - sy, E1m0, a2, a3);
    cell_emf32<C, T>(Vc, g.sv, p, E1m1, E2m1, a3);
    cell_emf32<C, T>(Vc, g.sv, p - 1, a1, E2m0, a3);
    svx_m = sv_val<T>(ld32(e.svx, p));
    svy_m = sv_val<T>(ld32(e.svy, p));
    exj_m = ld32(e.exj, p);
    eyi_m = ld32(e.eyi, p);
  }
  T pex = 0, pey = 0, pez = 0;  // edges of the previous plane, own column
  // dt / dx of the lane's column and row (clamped indices: lanes past the block's end
  // compute no face)
  const double rdx = dt / g.dx[0][min(i, g.nt[0] - 1)], rdy = dt / g.dx[1][min(j, g.nt[1] - 1)];
  const double mrdx = -dt / g.dx[0][min(i, g.nt[0] - 1)];
  const bool own = l < TI && w < TJ;
  for (int k = ks; k <= klast; k++) {
    // the faces of plane k - 1 this lane updates in this step: their old values (and RK
    // begin values) are loaded before the edge work and the barrier; only this lane
    // writes them, later in this step
    const int kk = k - 1;
    const unsigned q = (unsigned)kk * vz + (unsigned)j * vy + (unsigned)i;
    const bool u1 = k > ks && own && i <= g.end[0] && j < g.end[1] && kk < g.end[2];
    const bool u2 = k > ks && own && i < g.end[0] && j <= g.end[1] && kk < g.end[2];
    const bool u3 = k > ks && own && i < g.end[0] && j < g.end[1];
    double o1 = 0.0, o2 = 0.0, o3 = 0.0, z1 = 0.0, z2 = 0.0, z3 = 0.0;
    if (u1) {
      o1 = ld32(Vs, q);
      if (mode == 2) z1 = ld32(Vs0, q);
    }
    if (u2) {
      o2 = ld32(Vs1, q);
      if (mode == 2) z2 = ld32(Vs01, q);
    }
    if (u3) {
      o3 = ld32(Vs2, q);
      if (mode == 2) z3 = ld32(Vs02, q);
    }
    const double rdz = (k > ks && kk < g.end[2]) ? dt / g.dx[2][kk] : 0.0;
    T cex = 0, cey = 0, cez = 0;
    if (edge_ok) {
      const unsigned p = (unsigned)idx3(g, k, j, i);
      T E1[2][2][2], E2[2][2][2], E3[2][2][2];
      cell_emf32<C, T>(Vc, g.sv, p - sy - 1, E1[1][0][0], E2[1][0][0], E3[1][0][0]);
      cell_emf32<C, T>(Vc, g.sv, p - sy, E1[1][0][1], E2[1][0][1], E3[1][0][1]);
      cell_emf32<C, T>(Vc, g.sv, p - 1, E1[1][1][0], E2[1][1][0], E3[1][1][0]);
      cell_emf32<C, T>(Vc, g.sv, p, E1[1][1][1], E2[1][1][1], E3[1][1][1]);
      E1[0][0][1] = E1m0;
      E1[0][1][1] = E1m1;
      E2[0][1][0] = E2m0;
      E2[0][1][1] = E2m1;
      ContactFaces<T> f;
      f.svx = sv_val<T>(ld32(e.svx, p));
      f.svx_jm = sv_val<T>(ld32(e.svx, p - sy));
      f.svx_km = svx_m;
      f.svy = sv_val<T>(ld32(e.svy, p));
      f.svy_im = sv_val<T>(ld32(e.svy, p - 1));
      f.svy_km = svy_m;
      f.svz = sv_val<T>(ld32(e.svz, p));
      f.svz_jm = sv_val<T>(ld32(e.svz, p - sy));
      f.svz_im = sv_val<T>(ld32(e.svz, p - 1));
      f.ezi = ld32(e.ezi, p);
      f.ezi_jm = ld32(e.ezi, p - sy);
      f.ezj = ld32(e.ezj, p);
      f.ezj_im = ld32(e.ezj, p - 1);
      f.exj = ld32(e.exj, p);
      f.exj_km = exj_m;
      f.exk = ld32(e.exk, p);
      f.exk_jm = ld32(e.exk, p - sy);
      f.eyi = ld32(e.eyi, p);
      f.eyi_km = eyi_m;
      f.eyk = ld32(e.eyk, p);
      f.eyk_im = ld32(e.eyk, p - 1);
      contact_edges3(f, E1, E2, E3, cex, cey, cez);
      E1m0 = E1[1][0][1];
      E1m1 = E1[1][1][1];
      E2m0 = E2[1][1][0];
      E2m1 = E2[1][1][1];
      svx_m = f.svx;
      svy_m = f.svy;
      exj_m = f.exj;
      eyi_m = f.eyi;
    }
    lez[k % 3][w][l] = cez;
    lex[k % 3][w][l] = cex;
    __syncthreads();
    if (k > ks) {
      // faces of plane kk = k - 1 from the edges of planes kk (p*) and k (c*)
      const T ez_ip = __shfl_down(pez, 1);
      const T ey_ip = __shfl_down(pey, 1);
      const T ez_jp = lez[kk % 3][w < 7 ? w + 1 : w][l];
      const T ex_jp = lex[kk % 3][w < 7 ? w + 1 : w][l];
      if (u1) {  // BX1s
        double r = 0.0 - rdy * ((double)ez_jp - (double)pez);
        r = r + rdz * ((double)cey - (double)pey);
        evolve_face_pre(Vs, Vs0, q, o1, z1, r, mode, wc, w0);
      }
      if (u2) {  // BX2s
        double r = rdx * ((double)ez_ip - (double)pez);
        r = r - rdz * ((double)cex - (double)pex);
        evolve_face_pre(Vs1, Vs01, q, o2, z2, r, mode, wc, w0);
      }
      if (u3) {  // BX3s at face plane kk
        const double r = mrdx * ((double)ey_ip - (double)pey) + rdy * ((double)ex_jp - (double)pex);
        evolve_face_pre(Vs2, Vs02, q, o3, z3, r, mode, wc, w0);
      }
    }
    pex = cex;
    pey = cey;
    pez = cez;
  }
  if (ke == g.end[2] + 1) {
    // the top face plane of the block (BX3s at end) from the last edge plane
    const int kk = g.end[2];
    const T ey_ip = __shfl_down(pey, 1);
    const T ex_jp = lex[kk % 3][w < 7 ? w + 1 : w][l];
    if (l < TI && w < TJ && i < g.end[0] && j < g.end[1]) {
      const double r = -dt / g.dx[0][i] * ((double)ey_ip - (double)pey) +
                       dt / g.dx[1][j] * ((double)ex_jp - (double)pex);
      const unsigned q = (unsigned)kk * vz + (unsigned)j * vy + (unsigned)i;
      evolve_face(Vs2, Vs02, q, r, mode, wc, w0);
    }
  }
}

// ------------------------------------------------------------------------------------
// Pointwise conversions
// ------------------------------------------------------------------------------------
// ShockFlattening::FindShock (shockFlattening.hpp:62-168, Cartesian, with an energy
// equation) over the interior and one cell around it along the active directions: a cell
// is flagged where div v < 0 and the pressure jump across it exceeds smoothing times the
// smallest neighbouring pressure.
template <class C>
__global__ __launch_bounds__(256) void k_find_shock(GridDesc g, const double *Vc,
                                                   unsigned char *flags, double smoothing) {
  constexpr int jo = C::DIM >= 2 ? 1 : 0, ko = C::DIM == 3 ? 1 : 0;
  const int ni = g.end[0] - g.beg[0] + 2, nj = g.end[1] - g.beg[1] + 2 * jo;
  const int nk = g.end[2] - g.beg[2] + 2 * ko;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ni * nj * nk) return;
  const int i = g.beg[0] - 1 + t % ni;
  const int j = g.beg[1] - jo + (t / ni) % nj;
  const int k = g.beg[2] - ko + t / ((long long)ni * nj);
  const long long p = idx3(g, k, j, i);
  const long long sx = 1, sy = g.nt[0], sz = (long long)g.nt[0] * g.nt[1];
  // pressure: PRS, or rho cs^2 in isothermal configurations with the cell's own sound
  // speed (g.gamma holds a constant one)
  auto P = [&](long long q) {
    if constexpr (C::ISO) {
      const double cs = g.csiso != nullptr ? g.csiso[q] : g.gamma;
      return Vc[C::RHO * g.sv + q] * cs * cs;
    } else {
      return Vc[C::ENG * g.sv + q];
    }
  };
  double divV = (Vc[C::MX1 * g.sv + p + sx] - Vc[C::MX1 * g.sv + p - sx]) / (2 * g.dx[0][i]);
  if constexpr (C::DIM >= 2)
    divV = divV + (Vc[C::MX2 * g.sv + p + sy] - Vc[C::MX2 * g.sv + p - sy]) / (2 * g.dx[1][j]);
  if constexpr (C::DIM == 3)
    divV = divV + (Vc[C::MX3 * g.sv + p + sz] - Vc[C::MX3 * g.sv + p - sz]) / (2 * g.dx[2][k]);
  unsigned char f = 0;
  if (divV < 0.0) {
    double pmin = P(p);
    pmin = fmin(pmin, P(p + sx));
    pmin = fmin(pmin, P(p - sx));
    double gradP = fabs(P(p + sx) - P(p - sx));
    if constexpr (C::DIM >= 2) {
      pmin = fmin(pmin, P(p + sy));
      pmin = fmin(pmin, P(p - sy));
      gradP += fabs(P(p + sy) - P(p - sy));
    }
    if constexpr (C::DIM == 3) {
      pmin = fmin(pmin, P(p + sz));
      pmin = fmin(pmin, P(p - sz));
      gradP += fabs(P(p + sz) - P(p - sz));
    }
    if (gradP > smoothing * pmin) f = 1;
  }
  flags[p] = f;
}

// ConvertPrimToCons over a box (convertConsToPrim.hpp:152-189)
template <class C>
__global__ __launch_bounds__(256) void k_prim_to_cons(GridDesc g, const double *Vc, double *Uc) {
  const long long n3 = (long long)g.nt[0] * g.nt[1] * g.nt[2];
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  double u[C::NVAR];
  cell_p2c<C>(g, Vc, p, u);
#pragma unroll
  for (int n = 0; n < C::NVAR; n++) Uc[n * g.sv + p] = u[n];
  // Tracer::ConvertPrimToCons (tracer.cpp:32-47)
  for (int n = C::NVAR; n < C::NVAR + g.ntr; n++) Uc[n * g.sv + p] = Vc[n * g.sv + p] * Vc[C::RHO * g.sv + p];
}

// ReconstructVcField (boundary.hpp:388-404) + AddAndStore of the cell state
// (stateContainer.cpp:75-104) + ConvertConsToPrim (convertConsToPrim.hpp:105-149), over
// the box [lo, hi).  flags: 1 = combine Uc = wc*Uc + w0*Uc0 first, 2 = write Uc back.
// A block of 256 threads covers a run of 256 cells along x: blockIdx.x = run + nxb (j -
// lo1), blockIdx.y = k - lo2.  Offsets are 32-bit (a block holds fewer than 2^29 cells and
// faces per component).
template <class C>
__global__ __launch_bounds__(256) void k_cons_to_prim(GridDesc g, double *Uc, const double *Uc0,
                                                      double *Vc, const double *Vs, int lo0,
                                                      int lo1, int lo2, int n0, int nxb,
                                                      int flags, double wc, double w0) {
  const int run = blockIdx.x % nxb;
  const int i = lo0 + 256 * run + threadIdx.x;
  if (i >= lo0 + n0) return;
  const int j = lo1 + blockIdx.x / nxb;
  const int k = lo2 + blockIdx.y;
  const unsigned p = (unsigned)idx3(g, k, j, i);
  double u[C::NVAR], v[C::NVAR];
#pragma unroll
  for (int n = 0; n < C::NVAR; n++) {
    if (C::is_ct_field(n)) continue;
    u[n] = ld32(Uc + n * g.sv, p);
    if (flags & 1) u[n] = wc * u[n] + w0 * ld32(Uc0 + n * g.sv, p);
  }
  if constexpr (C::MHD) {
    const unsigned vy = (unsigned)g.vsn[0], vz = (unsigned)g.vsn[0] * (unsigned)g.vsn[1];
    const unsigned q = (unsigned)k * vz + (unsigned)j * vy + (unsigned)i;
    u[C::BX1] = 0.5 * (ld32(Vs, q) + ld32(Vs, q + 1));
    if constexpr (C::DIM >= 2)
      u[C::BX2] = 0.5 * (ld32(Vs + g.svs, q) + ld32(Vs + g.svs, q + vy));
    if constexpr (C::DIM >= 3)
      u[C::BX3] = 0.5 * (ld32(Vs + 2 * g.svs, q) + ld32(Vs + 2 * g.svs, q + vz));
  }
  cons2prim<C>(u, v, g.gamma);
#pragma unroll
  for (int n = 0; n < C::NVAR; n++) st32(Vc + n * g.sv, p, v[n]);
  if (flags & 2) {
#pragma unroll
    for (int n = 0; n < C::NVAR; n++) st32(Uc + n * g.sv, p, u[n]);
  }
  // tracers: AddAndStore, then Tracer::ConvertConsToPrim (tracer.cpp:13-28)
  for (int n = C::NVAR; n < C::NVAR + g.ntr; n++) {
    double ut = ld32(Uc + n * g.sv, p);
    if (flags & 1) ut = wc * ut + w0 * ld32(Uc0 + n * g.sv, p);
    st32(Vc + n * g.sv, p, ut / u[C::RHO]);
    if (flags & 2) st32(Uc + n * g.sv, p, ut);
  }
}

// Second half of ConsToPrim on the ping-pong schedule (adiabatic MHD): the cell field of
// the interior cells from the updated faces (ReconstructVcField, boundary.hpp:388-404),
// the pressure from E - kinetic energy left in the energy slot by the last sweep
// (prs_from_part), and with write_uc the cell field into Uc as k_cons_to_prim does.
// Grid as k_cons_to_prim's.
template <class C>
__global__ __launch_bounds__(256) void k_prs_from_faces(GridDesc g, double *Vc, const double *Vs,
                                                        double *Uc, int n0, int nxb,
                                                        int write_uc) {
  if constexpr (C::MHD && !C::ISO && C::DIM == 3) {
    const int run = blockIdx.x % nxb;
    const int i = g.beg[0] + 256 * run + threadIdx.x;
    if (i >= g.beg[0] + n0) return;
    const int j = g.beg[1] + blockIdx.x / nxb;
    const int k = g.beg[2] + blockIdx.y;
    const unsigned p = (unsigned)idx3(g, k, j, i);
    const unsigned vy = (unsigned)g.vsn[0], vz = (unsigned)g.vsn[0] * (unsigned)g.vsn[1];
    const unsigned q = (unsigned)k * vz + (unsigned)j * vy + (unsigned)i;
    double b[3];
    b[0] = 0.5 * (ld32(Vs, q) + ld32(Vs, q + 1));
    b[1] = 0.5 * (ld32(Vs + g.svs, q) + ld32(Vs + g.svs, q + vy));
    b[2] = 0.5 * (ld32(Vs + 2 * g.svs, q) + ld32(Vs + 2 * g.svs, q + vz));
    st32(Vc + C::ENG * g.sv, p, prs_from_part<C>(ld32(Vc + C::ENG * g.sv, p), b, g.gamma));
    st32(Vc + C::BX1 * g.sv, p, b[0]);
    st32(Vc + C::BX2 * g.sv, p, b[1]);
    st32(Vc + C::BX3 * g.sv, p, b[2]);
    if (write_uc) {
      st32(Uc + C::BX1 * g.sv, p, b[0]);
      st32(Uc + C::BX2 * g.sv, p, b[1]);
      st32(Uc + C::BX3 * g.sv, p, b[2]);
    }
  }
}

// ReconstructVcField over all cells (boundary.hpp:388-404)
template <class C>
__global__ __launch_bounds__(256) void k_vc_field(GridDesc g, const double *Vs, double *X) {
  const long long n3 = (long long)g.nt[0] * g.nt[1] * g.nt[2];
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n3) return;
  const int i = p % g.nt[0];
  const int j = (p / g.nt[0]) % g.nt[1];
  const int k = p / ((long long)g.nt[0] * g.nt[1]);
  X[C::BX1 * g.sv + p] = 0.5 * (Vs[idxs(g, 0, k, j, i)] + Vs[idxs(g, 0, k, j, i + 1)]);
  if constexpr (C::DIM >= 2)
    X[C::BX2 * g.sv + p] = 0.5 * (Vs[idxs(g, 1, k, j, i)] + Vs[idxs(g, 1, k, j + 1, i)]);
  if constexpr (C::DIM >= 3)
    X[C::BX3 * g.sv + p] = 0.5 * (Vs[idxs(g, 2, k, j, i)] + Vs[idxs(g, 2, k + 1, j, i)]);
}

// Ghost zones of a box periodic in every active direction, all directions at once: every
// ghost cell of Vc and ghost face of each Vs component takes the value at the periodic
// image of its ghost coordinates (faces along their own direction wrap with period
// nint, face end being interior).  Equals the direction-by-direction fill of k_boundary for periodic
// sides (boundary.hpp:576-673), whose corner values are the images in every direction,
// and the following ReconstructVcField: ghost cells take the average of their two faces as
// the fill leaves them, interior cells keep theirs, which equals the face average when
// the previous stage's ConsToPrim wrote it.  The ghost sites are
// enumerated slab by slab: slab d holds the sites whose index is a ghost along d and
// interior along every direction before d.
template <class C>
__global__ __launch_bounds__(256) void k_boundary_periodic(GridDesc g, double *Vc, double *Vs) {
  constexpr int NF = C::MHD ? C::DIM : 0;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (int f = -1; f < NF; f++) {  // f = -1: cells, else face component f
    int lo[3], hi[3], ext[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = g.beg[d];
      hi[d] = g.end[d] + ((f == d) ? 1 : 0);      // interior index range [lo, hi)
      ext[d] = g.nt[d] + ((f == d) ? 1 : 0);      // full index range [0, ext)
    }
    for (int d = 0; d < C::DIM; d++) {
      long long sz = 2LL * g.beg[d];  // ghosts along d: beg below and beg above
      for (int d2 = 0; d2 < 3; d2++)
        if (d2 != d) sz *= (d2 < d) ? (hi[d2] - lo[d2]) : ext[d2];
      if (t >= sz) {
        t -= sz;
        continue;
      }
      // decode t into the slab: along d the ghost index, before d interior, after d any
      int q[3];
      long long r = t;
      for (int d2 = 0; d2 < 3; d2++) {
        const long long n = (d2 == d) ? (2LL * g.beg[d]) : (d2 < d) ? (hi[d2] - lo[d2]) : ext[d2];
        const int v = (int)(r % n);
        r /= n;
        if (d2 == d) q[d2] = v < g.beg[d] ? v : hi[d] + (v - g.beg[d]);
        else q[d2] = (d2 < d) ? lo[d2] + v : v;
      }
      // the image wraps the ghost coordinates only: an interior index keeps its own value
      // (the last face along a face component's direction is interior and need not
      // equal the first one, e.g. where the sweeps leave face data unwritten)
      int src[3];
      for (int d2 = 0; d2 < 3; d2++) {
        const int n = g.end[d2] - g.beg[d2];
        const bool ghost = d2 < C::DIM && (q[d2] < lo[d2] || q[d2] >= hi[d2]);
        src[d2] = ghost ? g.beg[d2] + (((q[d2] - g.beg[d2]) % n) + n) % n : q[d2];
      }
      if (f < 0) {
        const long long ps = idx3(g, src[2], src[1], src[0]);
        const long long pd = idx3(g, q[2], q[1], q[0]);
        for (int n = 0; n < C::NVAR + g.ntr; n++) {
          if (C::MHD && n >= C::BX1 && n < C::BX1 + C::DIM) {
            // ReconstructVcField: the average of the two faces as the fill leaves them
            // (each face at the image of its ghost coordinates)
            const int fd = n - C::BX1;
            double v[2];
            for (int side = 0; side < 2; side++) {
              int fi[3];
              for (int d2 = 0; d2 < 3; d2++) {
                const int qq = q[d2] + ((d2 == fd) ? side : 0);
                const int fhi = g.end[d2] + ((d2 == fd) ? 1 : 0);
                const int nn = g.end[d2] - g.beg[d2];
                const bool gh = d2 < C::DIM && (qq < g.beg[d2] || qq >= fhi);
                fi[d2] = gh ? g.beg[d2] + (((qq - g.beg[d2]) % nn) + nn) % nn : qq;
              }
              v[side] = Vs[idxs(g, fd, fi[2], fi[1], fi[0])];
            }
            Vc[n * g.sv + pd] = 0.5 * (v[0] + v[1]);
          } else {
            Vc[n * g.sv + pd] = Vc[n * g.sv + ps];
          }
        }
      } else {
        Vs[idxs(g, f, q[2], q[1], q[0])] = Vs[idxs(g, f, src[2], src[1], src[0])];
      }
      return;
    }
  }
}

// AddAndStore over whole arrays (stateContainer.cpp:91-98)
static __global__ __launch_bounds__(256) void k_add_and_store(double *U, const double *U0, long long n,
                                                       double wl, double wr) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) U[p] = wl * U[p] + wr * U0[p];
}

// max over the interior of InvDt (bit pattern), for ComputeTimestep (dataBlock.cpp:333-364)
static __global__ __launch_bounds__(256) void k_invdt_max(GridDesc g, const double *InvDt,
                                                   unsigned long long *out) {
  const int ni = g.end[0] - g.beg[0], nj = g.end[1] - g.beg[1], nk = g.end[2] - g.beg[2];
  const long long n = (long long)ni * nj * nk;
  unsigned long long m = 0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = g.beg[0] + t % ni;
    const int j = g.beg[1] + (t / ni) % nj;
    const int k = g.beg[2] + t / ((long long)ni * nj);
    const unsigned long long b = __double_as_longlong(InvDt[idx3(g, k, j, i)]);
    m = b > m ? b : m;
  }
  block_max_atomic(m, out);
}

// CheckDivB (checkDivB.hpp:14-68): max over the interior, bit pattern
template <class C>
__global__ __launch_bounds__(256) void k_divb_max(GridDesc g, const double *Vs,
                                                  unsigned long long *out) {
  const int ni = g.end[0] - g.beg[0], nj = g.end[1] - g.beg[1], nk = g.end[2] - g.beg[2];
  const long long n = (long long)ni * nj * nk;
  unsigned long long m = 0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = g.beg[0] + t % ni;
    const int j = g.beg[1] + (t / ni) % nj;
    const int k = g.beg[2] + t / ((long long)ni * nj);
    auto vs = [&](int c, int kk, int jj, int ii) { return Vs[idxs(g, c, kk, jj, ii)]; };
    const double A1 = face_A<C::DIM, 0>(g, k, j, i);
    double dB = A1 * vs(0, k, j, i + 1) - A1 * vs(0, k, j, i);
    double d1 = 0.5 * (A1 + A1);
    double B1 = 0.5 * (vs(0, k, j, i + 1) + vs(0, k, j, i));
    double am = fabs(B1) * d1;
    if constexpr (C::DIM >= 2) {
      const double A2 = face_A<C::DIM, 1>(g, k, j, i);
      const double dB2 = A2 * vs(1, k, j + 1, i) - A2 * vs(1, k, j, i);
      const double d2 = 0.5 * (A2 + A2);
      const double B2 = 0.5 * (vs(1, k, j + 1, i) + vs(1, k, j, i));
      am = am + fabs(B2) * d2;
      dB = dB + dB2;
    }
    if constexpr (C::DIM >= 3) {
      const double A3 = face_A<C::DIM, 2>(g, k, j, i);
      const double dB3 = A3 * vs(2, k + 1, j, i) - A3 * vs(2, k, j, i);
      const double d3 = 0.5 * (A3 + A3);
      const double B3 = 0.5 * (vs(2, k + 1, j, i) + vs(2, k, j, i));
      am = am + fabs(B3) * d3;
      dB = dB + dB3;
    }
    double amplitude = 1e-40;
    amplitude += am;
    const double r = fabs(dB) / amplitude;
    const unsigned long long b = __double_as_longlong(r);
    m = b > m ? b : m;
  }
  block_max_atomic(m, out);
}

// ------------------------------------------------------------------------------------
// Boundaries (boundary.hpp): one launch per direction fills both sides.
// ------------------------------------------------------------------------------------
enum BcType { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECTIVE = 2, BC_USERDEF = 3 };

// Ghost cells of Vc along `dir`, and the tangential Vs components (all components for
// periodic), for the sides whose type is not user-defined.
template <class C>
__global__ __launch_bounds__(256) void k_boundary(GridDesc g, double *Vc, double *Vs, int dir,
                                                  int lbc, int rbc) {
  // enumerate ghost sites of the cell array and of the three face arrays in one grid
  const int ng = g.beg[dir];
  const int nint = g.end[dir] - g.beg[dir];
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  // slab sizes (both sides) for cells and for each face component
  constexpr int NF = C::MHD ? C::DIM : 0;  // face-centred components to fill
  long long sz[4];
  int ext[4][3];
  for (int f = -1; f < NF; f++) {
    for (int d = 0; d < 3; d++) {
      const int extra = (f == d) ? 1 : 0;
      ext[f + 1][d] = (d == dir) ? ng : g.nt[d] + extra;
    }
    sz[f + 1] = 2LL * ext[f + 1][0] * ext[f + 1][1] * ext[f + 1][2];
  }
  int f = -1;
  for (; f < NF; f++) {
    if (t < sz[f + 1]) break;
    t -= sz[f + 1];
  }
  if (f >= NF) return;
  const int *e3 = ext[f + 1];
  const long long half = sz[f + 1] / 2;
  const int side = t >= half ? 1 : 0;
  if (side) t -= half;
  int q[3];
  q[0] = t % e3[0];
  q[1] = (t / e3[0]) % e3[1];
  q[2] = t / ((long long)e3[0] * e3[1]);
  const int bc = side ? rbc : lbc;
  if (bc == BC_USERDEF) return;
  // position along dir (BoundaryForAll / BoundaryForX?s ranges, boundary.hpp:969-1099)
  const int nxd = nint + ((f == dir) ? 1 : 0);
  q[dir] = side ? (ng + nxd + q[dir]) : q[dir];
  int r[3] = {q[0], q[1], q[2]};
  if (f < 0) {
    // tracers included (boundary.hpp:135-136); with MHD the reflective sign test n >= BX1
    // also flips them (boundary.hpp:697)
    for (int n = 0; n < C::NVAR + g.ntr; n++) {
      double val;
      if (bc == BC_PERIODIC) {
        r[dir] = ng + (q[dir] + ng * (nint - 1)) % nint;
        val = Vc[n * g.sv + idx3(g, r[2], r[1], r[0])];
      } else if (bc == BC_OUTFLOW) {
        r[dir] = ng + side * (nint - 1);
        val = Vc[n * g.sv + idx3(g, r[2], r[1], r[0])];
        const int sign = 1 - 2 * side;
        if (n == C::MX1 + dir && sign * val >= 0.0) val = 0.0;
      } else {
        r[dir] = 2 * (ng + side * nint) - q[dir] - 1;
        const double sign =
            (n == C::MX1 + dir || (C::MHD && n >= C::BX1 && n != C::BX1 + dir)) ? -1.0 : 1.0;
        val = sign * Vc[n * g.sv + idx3(g, r[2], r[1], r[0])];
      }
      Vc[n * g.sv + idx3(g, q[2], q[1], q[0])] = val;
    }
  } else {
    if (bc == BC_PERIODIC) {
      r[dir] = ng + (q[dir] + ng * (nint - 1)) % nint;
      Vs[idxs(g, f, q[2], q[1], q[0])] = Vs[idxs(g, f, r[2], r[1], r[0])];
    } else if (f != dir) {
      if (bc == BC_OUTFLOW) {
        r[dir] = ng + side * (nint - 1);
        Vs[idxs(g, f, q[2], q[1], q[0])] = Vs[idxs(g, f, r[2], r[1], r[0])];
      } else {
        r[dir] = 2 * (ng + side * nint) - q[dir] - 1;
        Vs[idxs(g, f, q[2], q[1], q[0])] = -Vs[idxs(g, f, r[2], r[1], r[0])];
      }
    }
  }
}

// ReconstructNormalField (boundary.hpp:407-530), Cartesian without axis: one thread per
// transverse site, sequential over the ghost layers.
template <class C>
__global__ __launch_bounds__(256) void k_normal_field(GridDesc g, double *Vs, int dir, int recL,
                                                      int recR) {
  const int o1 = (dir == 0) ? 1 : 0;
  const int o2 = (dir == 2) ? 1 : 2;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nt[o1] * g.nt[o2]) return;
  int idx[3];
  idx[o1] = t % g.nt[o1];
  idx[o2] = t / g.nt[o1];
  auto vs = [&](int c, int kk, int jj, int ii) -> double & { return Vs[idxs(g, c, kk, jj, ii)]; };
  auto A = [&](int d, int kk, int jj, int ii) {
    if (d == 0) return face_A<C::DIM, 0>(g, kk, jj, ii);
    if (d == 1) return face_A<C::DIM, 1>(g, kk, jj, ii);
    return face_A<C::DIM, 2>(g, kk, jj, ii);
  };
  auto trans = [&](int k, int j, int i) {
    if (dir == 0) {
      double s = 0.0;
      if constexpr (C::DIM >= 2) {
        s = s + A(1, k, j + 1, i) * vs(1, k, j + 1, i);
        s = s - A(1, k, j, i) * vs(1, k, j, i);
      }
      if constexpr (C::DIM >= 3) {
        s = s + A(2, k + 1, j, i) * vs(2, k + 1, j, i);
        s = s - A(2, k, j, i) * vs(2, k, j, i);
      }
      return s;
    } else if (dir == 1) {
      double s = A(0, k, j, i + 1) * vs(0, k, j, i + 1) - A(0, k, j, i) * vs(0, k, j, i);
      if constexpr (C::DIM >= 3)
        s = s + 1 * (A(2, k + 1, j, i) * vs(2, k + 1, j, i) - A(2, k, j, i) * vs(2, k, j, i));
      return s;
    } else {
      double s = A(0, k, j, i + 1) * vs(0, k, j, i + 1) - A(0, k, j, i) * vs(0, k, j, i);
      s = s + A(1, k, j + 1, i) * vs(1, k, j + 1, i);
      s = s - A(1, k, j, i) * vs(1, k, j, i);
      return s;
    }
  };
  const int nstart = g.beg[dir] - 1, nend = g.end[dir], nd = g.nt[dir];
  if (recL) {
    for (int q = nstart; q >= 0; q--) {
      idx[dir] = q;
      const int k = idx[2], j = idx[1], i = idx[0];
      const int kp = k + (dir == 2), jp = j + (dir == 1), ip = i + (dir == 0);
      vs(dir, k, j, i) = 1.0 / A(dir, k, j, i) * (A(dir, kp, jp, ip) * vs(dir, kp, jp, ip) + (trans(k, j, i)));
    }
  }
  if (recR) {
    for (int q = nend; q < nd; q++) {
      idx[dir] = q;
      const int k = idx[2], j = idx[1], i = idx[0];
      const int kp = k + (dir == 2), jp = j + (dir == 1), ip = i + (dir == 0);
      vs(dir, kp, jp, ip) = 1.0 / A(dir, kp, jp, ip) * (A(dir, k, j, i) * vs(dir, k, j, i) - (trans(k, j, i)));
    }
  }
}

}  // namespace idfx
