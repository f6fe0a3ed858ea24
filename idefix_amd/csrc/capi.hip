// C ABI of idefix_amd (declared in include/idefix_amd.h): context management, dispatch
// of the runtime configuration to the compiled kernel configurations, and the host-side
// sequencing of a Runge-Kutta stage and cycle (timeIntegrator.cpp:246-414).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/idefix_amd.h"
#include "launch.hpp"

namespace idfx {
// defined by the instantiation units
void register_hd1(Ops *t);
void register_hd2(Ops *t);
void register_hd3(Ops *t);
void register_mhd2_f64(Ops *t);
void register_mhd2_f32(Ops *t);
void register_mhd3_plm_f64(Ops *t);
void register_mhd3_ppm_f64(Ops *t);
void register_mhd3_f32(Ops *t);
void register_mhd_roe(Ops *t);
void register_mhd2c3(Ops *t);
void register_mlp(Ops *t);
void register_hll(Ops *t);
void register_tvdlf(Ops *t);
void register_limo3_hd(Ops *t);
void register_limo3_mhd(Ops *t);
void register_limo3_mhd_b(Ops *t);
void register_mhd1(Ops *t);
void register_iso(Ops *t);
void register_iso3(Ops *t);

Ops *ops_table() {
  static Ops *table = [] {
    Ops *t = new Ops[kOpsTableSize];
    register_hd1(t);
    register_hd2(t);
    register_hd3(t);
    register_mhd2_f64(t);
    register_mhd2_f32(t);
    register_mhd3_plm_f64(t);
    register_mhd3_ppm_f64(t);
    register_mhd3_f32(t);
    register_mhd_roe(t);
    register_mhd2c3(t);
    register_mlp(t);
    register_hll(t);
    register_tvdlf(t);
    register_limo3_hd(t);
    register_limo3_mhd(t);
    register_limo3_mhd_b(t);
    register_mhd1(t);
    register_iso(t);
    register_iso3(t);
    return t;
  }();
  return table;
}

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ULL; }
__global__ void k_fill(double *p, long long n, double v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
}  // namespace idfx

using namespace idfx;

// Interior cells from which the fused CT kernel replaces k_corner_emf + k_evolve_b at
// fusion level 1 (3-D uct_contact, MI355X, per stage, fused against the two kernels:
// 128^3 0.19 against 0.12 ms, 256^3 1.10 against 0.87-0.92 ms, 384^3 3.52 against 3.25
// ms, 512^3 7.84 against 8.34 ms)
constexpr long long CT3D_MIN_CELLS = 1LL << 26;

struct idfx_ctx {
  idfx_desc d;
  GridDesc g;
  const Ops *ops = nullptr;
  int np_tot[3], nghost[3], nvar, vsn[3];
  long long n3 = 0, nvs = 0;
  std::vector<void *> allocs;
  EmfArrays<double> e64{};
  EmfArrays<float> e32{};
  double *cMax = nullptr;      // flux-only mode
  double *Flux = nullptr;      // FluxRiemann scratch for the unfused stage path
  double *InvDt = nullptr;     // used when the caller passes no InvDt
  double *Uc0 = nullptr, *Vs0 = nullptr;  // Runge-Kutta "begin" state
  unsigned long long *dtmax = nullptr, *divb = nullptr;  // device scalars
  unsigned char *shock = nullptr;  // shock flags (shock flattening on)
  unsigned long long *host = nullptr;                    // pinned readback
  hipEvent_t ev_dt = nullptr;
  int target_waves = 8192;
  bool fp32 = false;
  bool uniform = false;  // every direction has one cell width
  double dx0[3] = {1.0, 1.0, 1.0};
  // fused 2-D MHD stage: the stage reads (Vc, Vs) and writes the new state into the other
  // buffer of a ping-pong pair, so idfx_cycle alternates between the caller's arrays and
  // these
  // region timer (idfx_ctx_set_profiling): HIP event pairs recorded on the launch stream
  bool prof = false;
  std::vector<hipEvent_t> ev_start, ev_stop;
  std::vector<int> ev_region;
  size_t ev_used = 0;
  double region_ms[IDFX_NREGIONS] = {};
  long region_n[IDFX_NREGIONS] = {};
  // learned limiter (IDFX_RECON_MLP): interface weights per direction and the BF16 network
  bool mlp = false;
  float *mlpW[3] = {nullptr, nullptr, nullptr};
  short *mlp_wt = nullptr;   // w1t | w2t | w3t, transposed BF16
  float *mlp_b = nullptr;    // b1 | b2 | b3
  MlpNet net{};
  bool fused2d_ok = false;   // configuration supports it
  bool fused2d_on = false;   // selected (idfx_ctx_set_fusion)
  bool fused3d_ok = false;   // 3-D MHD with uct_contact: fused CT kernel (k_ct3d)
  bool fused3d_on = false;
  // the fused CT kernel (k_ct3d) rather than k_corner_emf + k_evolve_b: fusion level 2, or
  // level 1 on blocks of at least CT3D_MIN_CELLS interior cells, where it is the faster one
  bool ct3d_use = false;
  // 3-D adiabatic MHD without per-cell options or user-defined sides: idfx_cycle's stages
  // ping-pong between Vc and Vc2, the z sweep writing the next state (SW_FINAL)
  bool fused3d_pp_ok = false;
  double *Vc2 = nullptr, *Vs2 = nullptr;
  double *Vc3 = nullptr;  // 3-D ping-pong with three stages: the first stage's output
  bool vc3_failed = false;  // its allocation failed once: RK3 keeps the in-place first stage
  // enrolled user boundary (IDFX_BC_USERDEF sides) and the time it is called with
  idfx_userdef_boundary_fn user_fn = nullptr;
  void *user_arg = nullptr;
  double t_bound = 0.0;
  const void *emf() const { return fp32 ? (const void *)&e32 : (const void *)&e64; }
};

namespace idfx {
// the launch grid of a context, for the other translation units of the library
const GridDesc &ctx_grid(const idfx_ctx *c) { return c->g; }
}  // namespace idfx

namespace {

int check(hipError_t e) { return e == hipSuccess ? IDFX_OK : IDFX_E_LAUNCH; }
#define IDFX_TRY(x)                  \
  do {                               \
    int _s = (x);                    \
    if (_s != IDFX_OK) return _s;    \
  } while (0)

template <typename T>
int alloc_n(idfx_ctx *c, T **p, long long n) {
  void *q = nullptr;
  if (hipMalloc(&q, sizeof(T) * (size_t)std::max<long long>(n, 1)) != hipSuccess) return IDFX_E_ALLOC;
  c->allocs.push_back(q);
  *p = static_cast<T *>(q);
  // scratch starts at zero, as Kokkos views do
  if (hipMemset(q, 0, sizeof(T) * (size_t)std::max<long long>(n, 1)) != hipSuccess) return IDFX_E_ALLOC;
  return IDFX_OK;
}

template <typename T>
int alloc_emf(idfx_ctx *c, EmfArrays<T> &e) {
  const int dim = c->d.dim;
  const int emf = c->d.emf;
  const long long n = c->n3;
  IDFX_TRY(alloc_n(c, &e.ez, n));
  IDFX_TRY(alloc_n(c, &e.ezi, n));
  IDFX_TRY(alloc_n(c, &e.ezj, n));
  if (dim == 3) {
    IDFX_TRY(alloc_n(c, &e.ex, n));
    IDFX_TRY(alloc_n(c, &e.ey, n));
    IDFX_TRY(alloc_n(c, &e.exj, n));
    IDFX_TRY(alloc_n(c, &e.exk, n));
    IDFX_TRY(alloc_n(c, &e.eyi, n));
    IDFX_TRY(alloc_n(c, &e.eyk, n));
  } else if (c->d.comps == 3) {
    IDFX_TRY(alloc_n(c, &e.eyi, n));
    IDFX_TRY(alloc_n(c, &e.exj, n));
  }
  if (emf == IDFX_EMF_UCT_CONTACT) {
    IDFX_TRY(alloc_n(c, &e.svx, n));
    if (dim >= 2) IDFX_TRY(alloc_n(c, &e.svy, n));
    if (dim >= 3) IDFX_TRY(alloc_n(c, &e.svz, n));
  }
  if (emf == IDFX_EMF_UCT_HLL || emf == IDFX_EMF_UCT_HLLD) {
    IDFX_TRY(alloc_n(c, &e.axL, n));
    IDFX_TRY(alloc_n(c, &e.axR, n));
    IDFX_TRY(alloc_n(c, &e.dxL, n));
    IDFX_TRY(alloc_n(c, &e.dxR, n));
    if (dim >= 2) {
      IDFX_TRY(alloc_n(c, &e.ayL, n));
      IDFX_TRY(alloc_n(c, &e.ayR, n));
      IDFX_TRY(alloc_n(c, &e.dyL, n));
      IDFX_TRY(alloc_n(c, &e.dyR, n));
    }
    if (dim >= 3) {
      IDFX_TRY(alloc_n(c, &e.azL, n));
      IDFX_TRY(alloc_n(c, &e.azR, n));
      IDFX_TRY(alloc_n(c, &e.dzL, n));
      IDFX_TRY(alloc_n(c, &e.dzR, n));
    }
  }
  return IDFX_OK;
}

SweepArgs sweep_args(idfx_ctx *c, int dir, double dt, const double *Vc, const double *Vs,
                     double *Uc, double *InvDt, int flags) {
  SweepArgs a{};
  a.Vc = Vc;
  a.Vs = Vs;
  a.Uc = Uc;
  a.Uc0 = c->Uc0;
  a.InvDt = InvDt ? InvDt : c->InvDt;
  a.Flux = nullptr;
  a.cMax = c->cMax;
  a.dt = dt;
  a.emf = c->d.emf;
  a.flags = flags;
  a.dtmax = c->dtmax;
  a.mlpw = c->mlp ? c->mlpW[dir] : nullptr;
  a.shock = c->shock;
  for (int d = 0; d < 3; d++) a.ext[d] = 0;
  // uniform grid: every cell has the volume of cell (beg); dt/dV computed once here with
  // the device's association (dx1 * dx2) * dx3
  a.uniform = c->uniform ? 1 : 0;
  if (c->uniform) {
    double v = c->dx0[0];
    if (c->d.dim >= 2) v = v * c->dx0[1];
    if (c->d.dim >= 3) v = v * c->dx0[2];
    a.dtdV = dt / v;
  } else {
    a.dtdV = 0.0;
  }
  if (c->d.mhd) {
    // face range extension across DIR required by CT: nghost for the uct_hll(d) averages
    // of HllD and Hll (hlldMHD.hpp:32-49, hllMHD.hpp:32-49); RoeMHD and TvdlfMHD always
    // extend by one cell (roeMHD.hpp:62-73, tvdlfMHD.hpp:30-41)
    int perp = 1;
    if ((c->d.emf == IDFX_EMF_UCT_HLL || c->d.emf == IDFX_EMF_UCT_HLLD) &&
        (c->d.solver == IDFX_SOLVER_HLLD || c->d.solver == IDFX_SOLVER_HLL))
      perp = c->nghost[dir];
    for (int d = 0; d < c->d.dim; d++) a.ext[d] = (d == dir) ? 0 : perp;
  }
  return a;
}

// Learned limiter: interface weights of direction dir from Vc into the context's array
// (or W when given), before a sweep of that direction; its own profiler region.
int mlp_faces(idfx_ctx *c, int dir, const double *Vc, float *W, hipStream_t s);

bool needs_unfused(const idfx_ctx *c) {
  // the 1-D Roe flux keeps the previous flux at strong shocks (roeHD.hpp:315-328)
  return c->d.dim == 1 && !c->d.mhd && c->d.solver == IDFX_SOLVER_ROE;
}

// Shock flags from Vc before the first direction's fluxes, as RiemannSolver::CalcFlux<IDIR>
// does (calcFlux.hpp:30-33); the other directions reuse them.
int shock_step(idfx_ctx *c, int dir, const double *Vc, hipStream_t s) {
  if (!c->shock || dir != 0) return IDFX_OK;
  return check(c->ops->find_shock(c->g, Vc, c->shock, c->d.shock_smoothing, s));
}

int do_sweep_unfused(idfx_ctx *c, int dir, double dt, const double *Vc, const double *Vs,
                     double *Uc, double *InvDt, double *Flux, hipStream_t s) {
  IDFX_TRY(shock_step(c, dir, Vc, s));
  SweepArgs a = sweep_args(c, dir, dt, Vc, Vs, Uc, InvDt, 0);
  a.Flux = Flux;
  IDFX_TRY(mlp_faces(c, dir, Vc, nullptr, s));
  IDFX_TRY(check(c->ops->sweep(dir, false, c->g, a, c->emf(), c->target_waves, s)));
  IDFX_TRY(check(c->ops->correct_flux(dir, c->g, Flux, s)));
  IDFX_TRY(check(c->ops->rhs(dir, c->g, Flux, c->cMax, Uc, a.InvDt, dt, s)));
  return IDFX_OK;
}

int read_u64(idfx_ctx *c, unsigned long long *dev, unsigned long long *out, hipStream_t s) {
  IDFX_TRY(check(hipMemcpyAsync(c->host, dev, sizeof(unsigned long long), hipMemcpyDeviceToHost, s)));
  IDFX_TRY(check(hipStreamSynchronize(s)));
  *out = *c->host;
  return IDFX_OK;
}

double bits_to_double(unsigned long long b) {
  double d;
  std::memcpy(&d, &b, sizeof(d));
  return d;
}

// Fold every recorded event pair into the per-region totals (blocks until they complete).
int resolve_regions(idfx_ctx *c) {
  for (size_t q = 0; q < c->ev_used; q++) {
    IDFX_TRY(check(hipEventSynchronize(c->ev_stop[q])));
    float ms = 0.0f;
    IDFX_TRY(check(hipEventElapsedTime(&ms, c->ev_start[q], c->ev_stop[q])));
    c->region_ms[c->ev_region[q]] += ms;
    c->region_n[c->ev_region[q]] += 1;
  }
  c->ev_used = 0;
  return IDFX_OK;
}

// Times the launches issued in its scope on stream s as one region (pushRegion/popRegion
// with the profiler enabled, global.cpp:66-95, without the fences: the events sit in the
// stream).
struct Region {
  idfx_ctx *c;
  hipStream_t s;
  long slot = -1;
  Region(idfx_ctx *ctx, int id, hipStream_t st) : c(ctx), s(st) {
    if (!c->prof) return;
    if (c->ev_used == c->ev_start.size()) {
      if (c->ev_start.size() >= 512) {
        if (resolve_regions(c) != IDFX_OK) return;
      } else {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return;
        if (hipEventCreate(&b) != hipSuccess) {
          hipEventDestroy(a);
          return;
        }
        c->ev_start.push_back(a);
        c->ev_stop.push_back(b);
        c->ev_region.push_back(0);
      }
    }
    slot = (long)c->ev_used++;
    c->ev_region[slot] = id;
    hipEventRecord(c->ev_start[slot], s);
  }
  ~Region() {
    if (slot >= 0) hipEventRecord(c->ev_stop[slot], s);
  }
};

int mlp_faces(idfx_ctx *c, int dir, const double *Vc, float *W, hipStream_t s) {
  if (!c->mlp) return IDFX_OK;
  Region r(c, IDFX_REGION_MLP, s);
  return c->ops->mlp_infer(dir, c->g, Vc, c->net, W ? W : c->mlpW[dir], s) == hipSuccess
             ? IDFX_OK : IDFX_E_LAUNCH;
}

}  // namespace

extern "C" {

const char *idfx_strerror(int s) {
  switch (s) {
    case IDFX_OK: return "ok";
    case IDFX_E_ARGS: return "invalid argument or unsupported configuration";
    case IDFX_E_LAUNCH: return "HIP launch or runtime error";
    case IDFX_E_NAN: return "Nan found after integration cycle";
    case IDFX_E_DT_SMALL: return "dt is too small";
    case IDFX_E_FIRST_DT: return "Your guessed first_dt is too large";
    case IDFX_E_ALLOC: return "device allocation failed";
    default: return "unknown status";
  }
}

int idfx_ctx_create(const idfx_desc *desc, const double *dx1, const double *dx2,
                    const double *dx3, idfx_ctx **out) {
  if (!desc || !out) return IDFX_E_ARGS;
  *out = nullptr;
  const idfx_desc &d = *desc;
  if (d.dim < 1 || d.dim > 3 || d.comps < d.dim || d.comps > 3) return IDFX_E_ARGS;
  if (d.recon < 1 || (d.recon > 4 && d.recon != IDFX_RECON_MLP)) return IDFX_E_ARGS;
  const bool any_mode = d.solver == IDFX_SOLVER_ROE || d.solver == IDFX_SOLVER_HLL ||
                        d.solver == IDFX_SOLVER_TVDLF;
  if (d.mhd && d.solver != IDFX_SOLVER_HLLD && !any_mode) return IDFX_E_ARGS;
  if (!d.mhd && d.solver != IDFX_SOLVER_HLLC && !any_mode) return IDFX_E_ARGS;
  if (d.emf < 0 || d.emf > 4) return IDFX_E_ARGS;
  // HllMHD and TvdlfMHD write no uct_hlld face data (hllMHD.hpp:449-452,
  // tvdlfMHD.hpp:230-233): the reference would average unwritten arrays
  if (d.mhd && (d.solver == IDFX_SOLVER_HLL || d.solver == IDFX_SOLVER_TVDLF) &&
      d.emf == IDFX_EMF_UCT_HLLD)
    return IDFX_E_ARGS;
  // passive tracers ride on the fused sweeps; the flux-array path of the 1-D Roe HD
  // solver has no tracer step
  if (d.ntracer < 0 || d.ntracer > 64) return IDFX_E_ARGS;
  // shock flattening acts on the PLM and LimO3 slopes (extrapolateToFaces.hpp:98-226)
  if (d.shock_flattening && d.recon != IDFX_RECON_PLM && d.recon != IDFX_RECON_LIMO3)
    return IDFX_E_ARGS;
  if (d.ntracer > 0 && d.dim == 1 && !d.mhd && d.solver == IDFX_SOLVER_ROE) return IDFX_E_ARGS;
  if (d.isothermal && !(d.csiso > 0.0)) return IDFX_E_ARGS;
  const int key = ops_key(d.dim, d.comps, d.mhd, d.recon, d.solver,
                          d.precision == IDFX_PREC_FP32, d.isothermal ? 1 : 0);
  if (key < 0 || key >= kOpsTableSize) return IDFX_E_ARGS;
  const Ops *ops = &ops_table()[key];
  if (!ops->valid) return IDFX_E_ARGS;
  idfx_ctx *c = new idfx_ctx();
  c->d = d;
  c->ops = ops;
  c->fp32 = d.precision == IDFX_PREC_FP32;
  const double *dxh[3] = {dx1, dx2, dx3};
  for (int q = 0; q < 3; q++) {
    const bool act = q < d.dim;
    if (act && d.nx[q] < 1) {
      delete c;
      return IDFX_E_ARGS;
    }
    c->nghost[q] = act ? (d.recon < 4 ? 2 : 3) : 0;
    c->np_tot[q] = (act ? d.nx[q] : 1) + 2 * c->nghost[q];
    c->vsn[q] = c->np_tot[q] + (act ? 1 : 0);
    c->g.nt[q] = c->np_tot[q];
    c->g.beg[q] = c->nghost[q];
    c->g.end[q] = c->nghost[q] + (act ? d.nx[q] : 1);
    c->g.vsn[q] = c->vsn[q];
  }
  c->nvar = ops->nvar + d.ntracer;
  c->g.ntr = d.ntracer;
  c->n3 = (long long)c->np_tot[0] * c->np_tot[1] * c->np_tot[2];
  // kernels address each variable (and each Vs component) with 32-bit byte offsets
  if ((long long)c->vsn[0] * c->vsn[1] * c->vsn[2] >= (1LL << 29)) {
    delete c;
    return IDFX_E_ARGS;
  }
  c->nvs = d.mhd ? (long long)d.dim * c->vsn[0] * c->vsn[1] * c->vsn[2] : 0;
  c->g.sv = c->n3;
  c->g.svs = (long long)c->vsn[0] * c->vsn[1] * c->vsn[2];
  // the kernels read the equation of state's one constant from here: gamma, or the
  // isothermal sound speed
  c->g.gamma = d.isothermal ? d.csiso : d.gamma;
  c->g.csiso = nullptr;
  int st = IDFX_OK;
  for (int q = 0; q < 3 && st == IDFX_OK; q++) {
    double *p = nullptr;
    st = alloc_n(c, &p, c->np_tot[q]);
    if (st != IDFX_OK) break;
    std::vector<double> h(c->np_tot[q], 1.0);
    if (dxh[q]) std::memcpy(h.data(), dxh[q], sizeof(double) * c->np_tot[q]);
    c->dx0[q] = h[0];
    bool same = true;
    for (double v : h) same = same && (v == h[0]);
    if (q == 0) c->uniform = same;
    else c->uniform = c->uniform && same;
    if (hipMemcpy(p, h.data(), sizeof(double) * c->np_tot[q], hipMemcpyHostToDevice) != hipSuccess)
      st = IDFX_E_LAUNCH;
    c->g.dx[q] = p;
  }
  if (st == IDFX_OK && d.mhd) st = c->fp32 ? alloc_emf(c, c->e32) : alloc_emf(c, c->e64);
  if (st == IDFX_OK) st = alloc_n(c, &c->cMax, c->n3);
  if (st == IDFX_OK) st = alloc_n(c, &c->InvDt, c->n3);
  if (st == IDFX_OK) st = alloc_n(c, &c->Uc0, c->n3 * c->nvar);
  if (st == IDFX_OK && d.mhd) st = alloc_n(c, &c->Vs0, c->nvs);
  if (st == IDFX_OK && needs_unfused(c)) st = alloc_n(c, &c->Flux, c->n3 * c->nvar);
  if (st == IDFX_OK) st = alloc_n(c, &c->dtmax, 1);
  if (st == IDFX_OK && d.shock_flattening) st = alloc_n(c, &c->shock, c->n3);
  if (st == IDFX_OK) st = alloc_n(c, &c->divb, 1);
  if (st == IDFX_OK && d.recon == IDFX_RECON_MLP) {
    c->mlp = true;
    for (int q = 0; q < d.dim && st == IDFX_OK; q++) st = alloc_n(c, &c->mlpW[q], 16 * c->n3);
    if (st == IDFX_OK) st = alloc_n(c, &c->mlp_wt, 64 * 48 + 64 * 64 + 16 * 64);
    if (st == IDFX_OK) st = alloc_n(c, &c->mlp_b, 64 + 64 + 16);
    c->net.w1t = c->mlp_wt;
    c->net.w2t = c->mlp_wt + 64 * 48;
    c->net.w3t = c->mlp_wt + 64 * 48 + 64 * 64;
    c->net.b1 = c->mlp_b;
    c->net.b2 = c->mlp_b + 64;
    c->net.b3 = c->mlp_b + 128;
    c->net.valid = 0;
  }
  if (st == IDFX_OK && d.mhd && d.dim == 2 && ops->fused2d_y && !needs_unfused(c) &&
      d.ntracer == 0) {
    bool user = false;
    for (int q = 0; q < 2; q++)
      user = user || d.lbound[q] == IDFX_BC_USERDEF || d.rbound[q] == IDFX_BC_USERDEF;
    // user-defined ghost zones are written by the caller into its own arrays, which the
    // ping-pong pair would bypass
    if (!user) {
      st = alloc_n(c, &c->Vc2, c->n3 * c->nvar);
      if (st == IDFX_OK) st = alloc_n(c, &c->Vs2, c->nvs);
      c->fused2d_ok = c->fused2d_on = st == IDFX_OK;
    }
  }
  if (st == IDFX_OK && d.mhd && d.dim == 3 && d.emf == IDFX_EMF_UCT_CONTACT && ops->ct3d) {
    c->fused3d_ok = c->fused3d_on = true;
    c->ct3d_use = (long long)d.nx[0] * d.nx[1] * d.nx[2] >= CT3D_MIN_CELLS;
  }
  if (st == IDFX_OK && c->fused3d_ok && ops->prs_from_faces && !d.isothermal &&
      d.ntracer == 0 && !d.shock_flattening) {
    bool user = false;
    for (int q = 0; q < 3; q++)
      user = user || d.lbound[q] == IDFX_BC_USERDEF || d.rbound[q] == IDFX_BC_USERDEF;
    if (!user) {
      st = alloc_n(c, &c->Vc2, c->n3 * c->nvar);
      c->fused3d_pp_ok = st == IDFX_OK;
    }
  }
  if (st == IDFX_OK) {
    void *h = nullptr;
    if (hipHostMalloc(&h, 64, hipHostMallocDefault) != hipSuccess) st = IDFX_E_ALLOC;
    c->host = static_cast<unsigned long long *>(h);
  }
  if (st == IDFX_OK && hipEventCreateWithFlags(&c->ev_dt, hipEventDisableTiming) != hipSuccess)
    st = IDFX_E_LAUNCH;
  if (st == IDFX_OK) {
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
      hipDeviceProp_t prop;
      if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
    }
    c->target_waves = ncu * 4 * 8;
  }
  if (st != IDFX_OK) {
    idfx_ctx_destroy(c);
    return st;
  }
  *out = c;
  return IDFX_OK;
}

int idfx_ctx_destroy(idfx_ctx *c) {
  if (!c) return IDFX_OK;
  hipDeviceSynchronize();
  for (void *p : c->allocs) hipFree(p);
  if (c->host) hipHostFree(c->host);
  if (c->ev_dt) hipEventDestroy(c->ev_dt);
  for (hipEvent_t e : c->ev_start) hipEventDestroy(e);
  for (hipEvent_t e : c->ev_stop) hipEventDestroy(e);
  delete c;
  return IDFX_OK;
}

int idfx_ctx_layout(const idfx_ctx *c, int *np_tot, int *nghost, int *nvar, int *vs_shape) {
  if (!c) return IDFX_E_ARGS;
  for (int q = 0; q < 3; q++) {
    if (np_tot) np_tot[q] = c->np_tot[q];
    if (nghost) nghost[q] = c->nghost[q];
    if (vs_shape) vs_shape[q] = c->vsn[q];
  }
  if (nvar) *nvar = c->nvar;
  return IDFX_OK;
}

int idfx_ctx_array(idfx_ctx *c, int which, double **ptr) {
  if (!c || !ptr || c->fp32) return IDFX_E_ARGS;
  if (which == 24) {
    *ptr = c->cMax;
    return IDFX_OK;
  }
  if (which >= 25 && which <= 27) {  // learned-limiter interface weights (float [16][n3])
    *ptr = reinterpret_cast<double *>(c->mlp ? c->mlpW[which - 25] : nullptr);
    return *ptr ? IDFX_OK : IDFX_E_ARGS;
  }
  // svx, svy, svz hold byte codes (twice the upwind weight), returned through the slot
  double *const arr[24] = {c->e64.ex,  c->e64.ey,  c->e64.ez,  c->e64.ezi, c->e64.ezj, c->e64.exj,
                           c->e64.exk, c->e64.eyi, c->e64.eyk,
                           reinterpret_cast<double *>(c->e64.svx),
                           reinterpret_cast<double *>(c->e64.svy),
                           reinterpret_cast<double *>(c->e64.svz),
                           c->e64.axL, c->e64.axR, c->e64.ayL, c->e64.ayR, c->e64.azL, c->e64.azR,
                           c->e64.dxL, c->e64.dxR, c->e64.dyL, c->e64.dyR, c->e64.dzL, c->e64.dzR};
  if (which < 0 || which >= 24) return IDFX_E_ARGS;
  *ptr = arr[which];
  return *ptr ? IDFX_OK : IDFX_E_ARGS;
}

int idfx_enforce_boundary_dir(idfx_ctx *c, int dir, double *Vc, double *Vs, void *stream) {
  if (!c || !Vc || (c->d.mhd && !Vs) || dir < 0 || dir >= c->d.dim) return IDFX_E_ARGS;
  hipStream_t s = (hipStream_t)stream;
  // built-in side types (user-defined sides are left alone by the kernel)
  IDFX_TRY(check(c->ops->boundary(c->g, Vc, Vs, dir, c->d.lbound[dir], c->d.rbound[dir], s)));
  for (int side = 0; side < 2; side++) {
    const int bc = side == 0 ? c->d.lbound[dir] : c->d.rbound[dir];
    if (bc != IDFX_BC_USERDEF) continue;
    // boundary.hpp:316-332: no enrolled function is a configuration error
    if (!c->user_fn) return IDFX_E_ARGS;
    c->user_fn(c->user_arg, dir, side, c->t_bound, Vc, Vs, stream);
  }
  return IDFX_OK;
}

int idfx_reconstruct_normal_field(idfx_ctx *c, int dir, double *Vs, void *stream) {
  if (!c || !c->d.mhd || !Vs || dir < 0 || dir >= c->d.dim) return IDFX_E_ARGS;
  // every side that is not periodic, user-defined ones included (boundary.hpp:409-410)
  const int recL = c->d.lbound[dir] != IDFX_BC_PERIODIC;
  const int recR = c->d.rbound[dir] != IDFX_BC_PERIODIC;
  if (!recL && !recR) return IDFX_OK;
  return check(c->ops->normal_field(c->g, Vs, dir, recL, recR, (hipStream_t)stream));
}

int idfx_set_boundaries(idfx_ctx *c, double *Vc, double *Vs, void *stream) {
  if (!c || !Vc || (c->d.mhd && !Vs)) return IDFX_E_ARGS;
  hipStream_t s = (hipStream_t)stream;
  for (int dir = 0; dir < c->d.dim; dir++) {
    IDFX_TRY(idfx_enforce_boundary_dir(c, dir, Vc, Vs, stream));
    if (c->d.mhd) IDFX_TRY(idfx_reconstruct_normal_field(c, dir, Vs, stream));
  }
  if (c->d.mhd) IDFX_TRY(check(c->ops->vc_field(c->g, Vs, Vc, s)));
  return IDFX_OK;
}

int idfx_ctx_set_userdef_boundary(idfx_ctx *c, idfx_userdef_boundary_fn fn, void *arg) {
  if (!c) return IDFX_E_ARGS;
  c->user_fn = fn;
  c->user_arg = arg;
  return IDFX_OK;
}

int idfx_ctx_set_time(idfx_ctx *c, double t) {
  if (!c) return IDFX_E_ARGS;
  c->t_bound = t;
  return IDFX_OK;
}

int idfx_prim_to_cons(idfx_ctx *c, const double *Vc, double *Uc, void *stream) {
  if (!c || !Vc || !Uc) return IDFX_E_ARGS;
  return check(c->ops->prim_to_cons(c->g, Vc, Uc, (hipStream_t)stream));
}

int idfx_cons_to_prim(idfx_ctx *c, double *Uc, double *Vc, const double *Vs, void *stream) {
  if (!c || !Uc || !Vc || (c->d.mhd && !Vs)) return IDFX_E_ARGS;
  const int lo[3] = {0, 0, 0};
  return check(c->ops->cons_to_prim(c->g, Uc, nullptr, Vc, Vs, lo, c->np_tot, 2, 0.0, 0.0,
                                    (hipStream_t)stream));
}

int idfx_reconstruct_vc_field(idfx_ctx *c, const double *Vs, double *X, void *stream) {
  if (!c || !Vs || !X) return IDFX_E_ARGS;
  return check(c->ops->vc_field(c->g, Vs, X, (hipStream_t)stream));
}

int idfx_reset_stage(idfx_ctx *c, double *InvDt, void *stream) {
  if (!c) return IDFX_E_ARGS;
  double *p = InvDt ? InvDt : c->InvDt;
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(c->n3)), dim3(256), 0, (hipStream_t)stream, p, c->n3, 0.0);
  return check(hipGetLastError());
}

int idfx_calc_flux(idfx_ctx *c, int dir, const double *Vc, const double *Vs, double *Flux,
                   void *stream) {
  if (!c || !Vc || !Flux || dir < 0 || dir >= c->d.dim || (c->d.mhd && !Vs) || c->fp32)
    return IDFX_E_ARGS;
  IDFX_TRY(shock_step(c, dir, Vc, (hipStream_t)stream));
  SweepArgs a = sweep_args(c, dir, 0.0, Vc, Vs, nullptr, nullptr, 0);
  a.Flux = Flux;
  IDFX_TRY(mlp_faces(c, dir, Vc, nullptr, (hipStream_t)stream));
  return check(c->ops->sweep(dir, false, c->g, a, c->emf(), c->target_waves, (hipStream_t)stream));
}

int idfx_calc_rhs(idfx_ctx *c, int dir, double dt, double *Flux, double *Uc, double *InvDt,
                  void *stream) {
  if (!c || !Flux || !Uc || dir < 0 || dir >= c->d.dim) return IDFX_E_ARGS;
  hipStream_t s = (hipStream_t)stream;
  IDFX_TRY(check(c->ops->correct_flux(dir, c->g, Flux, s)));
  return check(c->ops->rhs(dir, c->g, Flux, c->cMax, Uc, InvDt ? InvDt : c->InvDt, dt, s));
}

int idfx_sweep(idfx_ctx *c, int dir, double dt, const double *Vc, const double *Vs, double *Uc,
               double *InvDt, int flags, void *stream) {
  if (!c || !Vc || !Uc || dir < 0 || dir >= c->d.dim || (c->d.mhd && !Vs)) return IDFX_E_ARGS;
  if (needs_unfused(c)) return IDFX_E_ARGS;
  const int f = flags & (SW_INIT_UC | SW_SAVE_UC0 | SW_INVDT_SET | SW_INVDT_ADD);
  IDFX_TRY(shock_step(c, dir, Vc, (hipStream_t)stream));
  SweepArgs a = sweep_args(c, dir, dt, Vc, Vs, Uc, InvDt, f);
  IDFX_TRY(mlp_faces(c, dir, Vc, nullptr, (hipStream_t)stream));
  return check(c->ops->sweep(dir, true, c->g, a, c->emf(), c->target_waves, (hipStream_t)stream));
}

int idfx_corner_emf(idfx_ctx *c, const double *Vc, const double *Vs, void *stream) {
  if (!c || !Vc || !Vs || !c->d.mhd) return IDFX_E_ARGS;
  return check(c->ops->corner_emf(c->g, Vc, Vs, c->emf(), c->d.emf, (hipStream_t)stream));
}

int idfx_evolve_mag_field(idfx_ctx *c, double dt, double *Vs, void *stream) {
  if (!c || !Vs || !c->d.mhd) return IDFX_E_ARGS;
  return check(c->ops->evolve_b(c->g, Vs, nullptr, c->emf(), dt, 0, 0.0, 0.0, (hipStream_t)stream));
}

int idfx_evolve_stage(idfx_ctx *c, double t, double dt, const double *Vc, double *Vs, double *Uc,
                      double *InvDt, void *stream) {
  (void)t;
  if (!c || !Vc || !Uc || (c->d.mhd && !Vs)) return IDFX_E_ARGS;
  hipStream_t s = (hipStream_t)stream;
  for (int dir = 0; dir < c->d.dim; dir++) {
    if (needs_unfused(c)) {
      IDFX_TRY(do_sweep_unfused(c, dir, dt, Vc, Vs, Uc, InvDt, c->Flux, s));
    } else {
      IDFX_TRY(shock_step(c, dir, Vc, s));
      SweepArgs a = sweep_args(c, dir, dt, Vc, Vs, Uc, InvDt, SW_INVDT_ADD);
      IDFX_TRY(mlp_faces(c, dir, Vc, nullptr, s));
      IDFX_TRY(check(c->ops->sweep(dir, true, c->g, a, c->emf(), c->target_waves, s)));
    }
  }
  if (c->d.mhd && c->d.dim >= 2) {
    IDFX_TRY(check(c->ops->corner_emf(c->g, Vc, Vs, c->emf(), c->d.emf, s)));
    IDFX_TRY(check(c->ops->evolve_b(c->g, Vs, nullptr, c->emf(), dt, 0, 0.0, 0.0, s)));
    IDFX_TRY(check(c->ops->vc_field(c->g, Vs, Uc, s)));
  }
  return IDFX_OK;
}

int idfx_compute_timestep(idfx_ctx *c, const double *InvDt, double *dtmin, void *stream) {
  if (!c || !dtmin) return IDFX_E_ARGS;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, c->dtmax);
  hipLaunchKernelGGL(k_invdt_max, dim3(1024), dim3(256), 0, s, c->g, InvDt ? InvDt : c->InvDt,
                     c->dtmax);
  IDFX_TRY(check(hipGetLastError()));
  unsigned long long b;
  IDFX_TRY(read_u64(c, c->dtmax, &b, s));
  *dtmin = 1.0 / bits_to_double(b);  // min(1/x) == 1/max(x) under correct rounding
  return IDFX_OK;
}

int idfx_add_and_store(idfx_ctx *c, double wl, double wr, double *U, const double *U0,
                       long long n, void *stream) {
  if (!c || !U || !U0 || n < 0) return IDFX_E_ARGS;
  hipLaunchKernelGGL(k_add_and_store, dim3(blocks_for(n)), dim3(256), 0, (hipStream_t)stream, U,
                     U0, n, wl, wr);
  return check(hipGetLastError());
}

int idfx_check_divb(idfx_ctx *c, const double *Vs, double *divb, void *stream) {
  if (!c || !Vs || !divb || !c->d.mhd) return IDFX_E_ARGS;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, c->divb);
  IDFX_TRY(check(c->ops->divb(c->g, Vs, c->divb, s)));
  unsigned long long b;
  IDFX_TRY(read_u64(c, c->divb, &b, s));
  *divb = bits_to_double(b);
  return IDFX_OK;
}

// One RK stage on the fused path.  When newdt_host is given in stage 0, the InvDt
// maximum is copied to pinned memory asynchronously and ctx->ev_dt records it.
// Boundary step of a stage.  In a box periodic along every active direction whose
// interior cell field equals the face average (every stage after the first of a cycle:
// the previous stage's ConsToPrim wrote it), one kernel fills every ghost zone with the
// same values as the per-direction fill followed by ReconstructVcField.
static bool all_periodic(const idfx_ctx *c) {
  for (int d = 0; d < c->d.dim; d++)
    if (c->d.lbound[d] != IDFX_BC_PERIODIC || c->d.rbound[d] != IDFX_BC_PERIODIC) return false;
  return true;
}
static int stage_boundaries(idfx_ctx *c, double *Vc, double *Vs, bool cell_field_consistent,
                            hipStream_t s) {
  if (all_periodic(c) && (cell_field_consistent || !c->d.mhd))
    return check(c->ops->boundary_periodic(c->g, Vc, Vs, s));
  return idfx_set_boundaries(c, Vc, Vs, s);
}

// VcOut (3-D ping-pong schedule, fused3d_pp_ok): the z sweep writes the stage's new state
// into VcOut (rho, v, E - kinetic energy) and k_prs_from_faces completes it after the face
// update, in place of ConsToPrim; Vc is only read.
static int stage_enqueue(idfx_ctx *c, int stage, int nstages, double dt, double *Vc, double *Vs,
                         double *Uc, double *InvDt, int skip_boundaries, int write_uc,
                         bool want_dt, hipStream_t s, bool consistent = false,
                         double *VcOut = nullptr) {
  static const double WC[2][2] = {{0.5, 0.0}, {0.25, 2.0 / 3.0}};
  static const double W0[2][2] = {{0.5, 0.0}, {0.75, 1.0 / 3.0}};
  if (!skip_boundaries) {
    Region r(c, IDFX_REGION_BOUNDARIES, s);
    IDFX_TRY(stage_boundaries(c, Vc, Vs, consistent, s));
  }
  double wc = 0.0, w0 = 0.0;
  if (stage > 0) {
    wc = WC[nstages - 2][stage - 1];
    w0 = W0[nstages - 2][stage - 1];
  }
  const int dim = c->d.dim;
  if (want_dt) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, c->dtmax);
  for (int dir = 0; dir < dim; dir++) {
    int flags = 0;
    if (dir == 0) {
      flags |= SW_INIT_UC;
      if (stage == 0 && nstages > 1) flags |= SW_SAVE_UC0;
    }
    if (stage == 0) {
      flags |= (dir == 0) ? SW_INVDT_SET : SW_INVDT_ADD;
      if (dir == dim - 1 && want_dt) flags |= SW_DT_REDUCE;
    }
    // the learned limiter's weights first, outside the sweep's region
    if (!needs_unfused(c)) IDFX_TRY(mlp_faces(c, dir, Vc, nullptr, s));
    Region r(c, IDFX_REGION_SWEEP_X + dir, s);
    if (needs_unfused(c)) {
      // reference sequence: PrimToCons, (begin copy), CalcFlux, CalcRightHandSide
      if (dir == 0) {
        IDFX_TRY(check(c->ops->prim_to_cons(c->g, Vc, Uc, s)));
        if (stage == 0 && nstages > 1)
          IDFX_TRY(check(hipMemcpyAsync(c->Uc0, Uc, sizeof(double) * c->n3 * c->nvar,
                                        hipMemcpyDeviceToDevice, s)));
        if (stage == 0)
          hipLaunchKernelGGL(k_fill, dim3(blocks_for(c->n3)), dim3(256), 0, s,
                             InvDt ? InvDt : c->InvDt, c->n3, 0.0);
      }
      IDFX_TRY(do_sweep_unfused(c, dir, dt, Vc, Vs, Uc, InvDt, c->Flux, s));
      if (dir == dim - 1 && want_dt)
        hipLaunchKernelGGL(k_invdt_max, dim3(1024), dim3(256), 0, s, c->g,
                           InvDt ? InvDt : c->InvDt, c->dtmax);
    } else {
      IDFX_TRY(shock_step(c, dir, Vc, s));
      if (VcOut && dir == dim - 1)
        flags |= SW_FINAL | (stage > 0 ? SW_RK : 0) | (write_uc ? SW_WRITE_UC : 0);
      SweepArgs a = sweep_args(c, dir, dt, Vc, Vs, Uc, InvDt, flags);
      a.VcOut = VcOut;
      a.wc = wc;
      a.w0 = w0;
      IDFX_TRY(check(c->ops->sweep(dir, true, c->g, a, c->emf(), c->target_waves, s)));
    }
  }
  if (c->d.mhd && dim >= 2) {
    const int mode = nstages > 1 ? (stage == 0 ? 1 : 2) : 0;
    if (c->fused3d_on && c->ct3d_use) {
      // CalcCornerEMF + EvolveMagField in one kernel (the edge arrays are not written)
      Region r(c, IDFX_REGION_CT_FUSED, s);
      IDFX_TRY(check(c->ops->ct3d(c->g, Vc, Vs, c->Vs0, c->emf(), dt, mode, wc, w0, s)));
    } else {
      {
        Region r(c, IDFX_REGION_CORNER_EMF, s);
        IDFX_TRY(check(c->ops->corner_emf(c->g, Vc, Vs, c->emf(), c->d.emf, s)));
      }
      Region r(c, IDFX_REGION_EVOLVE_B, s);
      IDFX_TRY(check(c->ops->evolve_b(c->g, Vs, c->Vs0, c->emf(), dt, mode, wc, w0, s)));
    }
  } else if (c->d.mhd && stage > 0) {
    // 1-D MHD keeps Vs constant; AddAndStore still applies (a no-op on equal values)
  }
  if (want_dt) {
    IDFX_TRY(check(hipMemcpyAsync(c->host, c->dtmax, sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s)));
    IDFX_TRY(check(hipEventRecord(c->ev_dt, s)));
  }
  const int lo[3] = {c->g.beg[0], c->g.beg[1], c->g.beg[2]};
  const int hi[3] = {c->g.end[0], c->g.end[1], c->g.end[2]};
  const int flags = (stage > 0 ? 1 : 0) | (write_uc ? 2 : 0);
  Region r(c, IDFX_REGION_CONS_TO_PRIM, s);
  if (VcOut)
    IDFX_TRY(check(c->ops->prs_from_faces(c->g, VcOut, Vs, Uc, write_uc, s)));
  else
    IDFX_TRY(check(c->ops->cons_to_prim(c->g, Uc, c->Uc0, Vc, Vs, lo, hi, flags, wc, w0, s)));
  return IDFX_OK;
}

// One stage of a 2-D MHD run with the fused second half: boundaries on the input pair, the
// x sweep, then k_fused2d_y (y sweep, corner EMF, face-B update, AddAndStore, ConsToPrim),
// which writes the interior of the output pair.  Same arithmetic as stage_enqueue.
static int stage_fused2d(idfx_ctx *c, int stage, int nstages, double dt, double *Vc_in,
                         double *Vs_in, double *Vc_out, double *Vs_out, double *Uc,
                         double *InvDt, int write_uc, bool want_dt, hipStream_t s,
                         bool consistent) {
  static const double WC[2][2] = {{0.5, 0.0}, {0.25, 2.0 / 3.0}};
  static const double W0[2][2] = {{0.5, 0.0}, {0.75, 1.0 / 3.0}};
  {
    Region r(c, IDFX_REGION_BOUNDARIES, s);
    IDFX_TRY(stage_boundaries(c, Vc_in, Vs_in, consistent, s));
  }
  if (want_dt) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, c->dtmax);
  int flags = SW_INIT_UC;
  if (stage == 0 && nstages > 1) flags |= SW_SAVE_UC0;
  if (stage == 0) flags |= SW_INVDT_SET;
  IDFX_TRY(shock_step(c, 0, Vc_in, s));
  SweepArgs ax = sweep_args(c, 0, dt, Vc_in, Vs_in, Uc, InvDt, flags);
  // learned limiter: the interface weights of both directions come from the stage's
  // input state, before either sweep
  IDFX_TRY(mlp_faces(c, 0, Vc_in, nullptr, s));
  IDFX_TRY(mlp_faces(c, 1, Vc_in, nullptr, s));
  {
    Region r(c, IDFX_REGION_SWEEP_X, s);
    IDFX_TRY(check(c->ops->sweep(0, true, c->g, ax, c->emf(), c->target_waves, s)));
  }
  flags = 0;
  if (stage == 0) flags |= SW_INVDT_ADD | (want_dt ? SW_DT_REDUCE : 0);
  SweepArgs ay = sweep_args(c, 1, dt, Vc_in, Vs_in, Uc, InvDt, flags);
  Fused2dArgs f{};
  f.Vs_in = Vs_in;
  f.Vc_out = Vc_out;
  f.Vs_out = Vs_out;
  f.Vs0 = c->Vs0;
  f.rk = nstages > 1 ? (stage == 0 ? 1 : 2) : 0;
  f.wc = stage > 0 ? WC[nstages - 2][stage - 1] : 0.0;
  f.w0 = stage > 0 ? W0[nstages - 2][stage - 1] : 0.0;
  f.write_uc = write_uc;
  f.ext0 = ay.ext[0];
  {
    Region r(c, IDFX_REGION_FUSED2D_Y, s);
    IDFX_TRY(check(c->ops->fused2d_y(c->g, ay, c->emf(), f, c->target_waves, s)));
  }
  if (want_dt) {
    IDFX_TRY(check(hipMemcpyAsync(c->host, c->dtmax, sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, s)));
    IDFX_TRY(check(hipEventRecord(c->ev_dt, s)));
  }
  return IDFX_OK;
}

int idfx_ctx_set_profiling(idfx_ctx *c, int on) {
  if (!c) return IDFX_E_ARGS;
  if (!on) IDFX_TRY(resolve_regions(c));
  c->prof = on != 0;
  return IDFX_OK;
}

int idfx_ctx_region_times(idfx_ctx *c, double *ms, long *count, int n) {
  if (!c || n < 0 || n > IDFX_NREGIONS) return IDFX_E_ARGS;
  IDFX_TRY(resolve_regions(c));
  for (int q = 0; q < n; q++) {
    if (ms) ms[q] = c->region_ms[q];
    if (count) count[q] = c->region_n[q];
  }
  for (int q = 0; q < IDFX_NREGIONS; q++) {
    c->region_ms[q] = 0.0;
    c->region_n[q] = 0;
  }
  return IDFX_OK;
}

static unsigned short host_bf16(float f) {
  unsigned u;
  std::memcpy(&u, &f, sizeof(u));
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (unsigned short)((u >> 16) | 0x40);
  const unsigned rounding = ((u >> 16) & 1u) + 0x7fffu;  // round to nearest even
  return (unsigned short)((u + rounding) >> 16);
}

int idfx_ctx_set_iso_sound_speed(idfx_ctx *c, const double *cs) {
  if (!c || !c->d.isothermal) return IDFX_E_ARGS;
  c->g.csiso = cs;  // every launch takes GridDesc by value, so the next one sees it
  return IDFX_OK;
}

int idfx_ctx_set_mlp_params(idfx_ctx *c, const float *p) {
  if (!c || !c->mlp) return IDFX_E_ARGS;
  if (!p) {
    c->net.valid = 0;
    return IDFX_OK;
  }
  // packed [in][out] FP32 -> transposed [out][in] BF16
  std::vector<unsigned short> wt(64 * 48 + 64 * 64 + 16 * 64);
  std::vector<float> b(64 + 64 + 16);
  const int W1 = 0, B1 = 3072, W2 = 3136, B2 = 7232, W3 = 7296, B3 = 8320;
  for (int o = 0; o < 64; o++)
    for (int i = 0; i < 48; i++) wt[o * 48 + i] = host_bf16(p[W1 + i * 64 + o]);
  for (int o = 0; o < 64; o++)
    for (int i = 0; i < 64; i++) wt[3072 + o * 64 + i] = host_bf16(p[W2 + i * 64 + o]);
  for (int o = 0; o < 16; o++)
    for (int i = 0; i < 64; i++) wt[3072 + 4096 + o * 64 + i] = host_bf16(p[W3 + i * 16 + o]);
  for (int q = 0; q < 64; q++) b[q] = p[B1 + q];
  for (int q = 0; q < 64; q++) b[64 + q] = p[B2 + q];
  for (int q = 0; q < 16; q++) b[128 + q] = p[B3 + q];
  if (hipMemcpy(c->mlp_wt, wt.data(), wt.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->mlp_b, b.data(), b.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
    return IDFX_E_LAUNCH;
  c->net.valid = 1;
  return IDFX_OK;
}

int idfx_mlp_faces(idfx_ctx *c, int dir, const double *Vc, float *W, void *stream) {
  if (!c || !c->mlp || !Vc || dir < 0 || dir >= c->d.dim) return IDFX_E_ARGS;
  return mlp_faces(c, dir, Vc, W, (hipStream_t)stream);
}

int idfx_ctx_set_fusion(idfx_ctx *c, int level) {
  if (!c || level < 0 || level > 2) return IDFX_E_ARGS;
  if (level >= 1 && !c->fused2d_ok && !c->fused3d_ok) return IDFX_E_ARGS;
  c->fused2d_on = level >= 1 && c->fused2d_ok;
  c->fused3d_on = level >= 1 && c->fused3d_ok;
  c->ct3d_use = c->fused3d_on &&
                (level == 2 || (long long)c->d.nx[0] * c->d.nx[1] * c->d.nx[2] >= CT3D_MIN_CELLS);
  return IDFX_OK;
}

int idfx_stage(idfx_ctx *c, int stage, int nstages, double dt, double cfl, double *Vc, double *Vs,
               double *Uc, double *InvDt, int skip_boundaries, int write_uc, double *newdt_host,
               void *stream) {
  if (!c || !Vc || !Uc || (c->d.mhd && !Vs) || nstages < 1 || nstages > 3 || stage < 0 ||
      stage >= nstages)
    return IDFX_E_ARGS;
  hipStream_t s = (hipStream_t)stream;
  const bool want_dt = stage == 0 && newdt_host != nullptr;
  IDFX_TRY(stage_enqueue(c, stage, nstages, dt, Vc, Vs, Uc, InvDt, skip_boundaries, write_uc,
                         want_dt, s));
  if (want_dt) {
    IDFX_TRY(check(hipEventSynchronize(c->ev_dt)));
    *newdt_host = cfl * (1.0 / bits_to_double(*c->host));
  }
  return IDFX_OK;
}

int idfx_cycle(idfx_ctx *c, int nstages, double cfl, double cfl_max_var, long ncycles, double *t,
               double *dt, double *Vc, double *Vs, double *Uc, double *InvDt, void *stream) {
  if (!c || !t || !dt || nstages < 1 || nstages > 3) return IDFX_E_ARGS;
  hipStream_t s = (hipStream_t)stream;
  const double t0 = *t;
  // fused 2-D stages ping-pong between (Vc, Vs) and the context's pair; an odd stage count
  // runs its first stage in place so that the last stage writes the caller's arrays
  const bool fused = c->fused2d_on && nstages >= 2 && Vs;
  // 3-D: the same alternation between (Vc, Vs) and (Vc2, Vs), Vs updated in place
  const bool pp3 = c->fused3d_on && c->fused3d_pp_ok && nstages >= 2 && Vs;
  const int first_fused = (nstages % 2 == 1) ? 1 : 0;
  // 3-D ping-pong inputs and outputs per stage: RK2 Vc -> Vc2 -> Vc; RK3 Vc -> Vc3 -> Vc2
  // -> Vc (Vc3 allocated at the first such cycle; without it the first RK3 stage runs in
  // place on the per-kernel ConsToPrim and the others alternate Vc -> Vc2 -> Vc)
  if (pp3 && nstages == 3 && !c->Vc3 && !c->vc3_failed &&
      alloc_n(c, &c->Vc3, c->n3 * c->nvar) != IDFX_OK) {
    c->Vc3 = nullptr;
    c->vc3_failed = true;
    (void)hipGetLastError();  // the failed allocation must not surface as a launch error
  }
  double *pin[3] = {Vc, c->Vc2, Vc}, *pout[3] = {c->Vc2, Vc, nullptr};
  int pp_first = 0;
  if (nstages == 3) {
    if (c->Vc3) {
      pin[0] = Vc, pout[0] = c->Vc3;
      pin[1] = c->Vc3, pout[1] = c->Vc2;
      pin[2] = c->Vc2, pout[2] = Vc;
    } else {
      pp_first = 1;
      pin[1] = Vc, pout[1] = c->Vc2;
      pin[2] = c->Vc2, pout[2] = Vc;
    }
  }
  // stage times of the boundary calls (TimeIntegrator::Cycle advances t by dt after each
  // stage, then t = wc t + w0 t0): t0, t0 + dt, and t0 + dt/2 for the third RK3 stage
  const double tstage[3] = {t0, t0 + *dt, 0.25 * (t0 + *dt + *dt) + 0.75 * t0};
  for (int stage = 0; stage < nstages; stage++) {
    c->t_bound = tstage[stage];
    if (pp3 && stage >= pp_first) {
      IDFX_TRY(stage_enqueue(c, stage, nstages, *dt, pin[stage], Vs, Uc, InvDt, 0,
                             stage == nstages - 1, stage == 0, s, stage > 0, pout[stage]));
    } else if (!fused || stage < first_fused) {
      IDFX_TRY(stage_enqueue(c, stage, nstages, *dt, Vc, Vs, Uc, InvDt, 0,
                             stage == nstages - 1, stage == 0, s, stage > 0));
    } else {
      const bool fwd = (stage - first_fused) % 2 == 0;
      IDFX_TRY(stage_fused2d(c, stage, nstages, *dt, fwd ? Vc : c->Vc2, fwd ? Vs : c->Vs2,
                             fwd ? c->Vc2 : Vc, fwd ? c->Vs2 : Vs, Uc, InvDt,
                             stage == nstages - 1, stage == 0, s, stage > 0));
    }
  }
  // the dt readback of stage 0 completes while the later stages still run
  IDFX_TRY(check(hipEventSynchronize(c->ev_dt)));
  const double newdt = cfl * (1.0 / bits_to_double(*c->host));
  *t = t0 + *dt;
  int status = IDFX_OK;
  if (newdt > cfl_max_var * *dt) {
    *dt = cfl_max_var * *dt;
  } else {
    if (ncycles == 0 && newdt < 0.5 * *dt) status = IDFX_E_FIRST_DT;
    *dt = newdt;
  }
  if (status == IDFX_OK && *dt < 1e-15) status = IDFX_E_DT_SMALL;
  if (status == IDFX_OK && !(newdt == newdt)) status = IDFX_E_NAN;
  return status;
}

}  // extern "C"
