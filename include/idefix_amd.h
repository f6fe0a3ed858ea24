/*
 * idefix_amd: C ABI of the MI355X-native per-stage finite-volume update.
 *
 * Every entry point takes raw device pointers in the reference's LayoutRight layout
 * and returns an int status (IDFX_OK = 0).  Arrays (reference src/fluid/fluid.hpp:529-560):
 *   Vc, Uc, FluxRiemann : double [nvar][np_tot[2]][np_tot[1]][np_tot[0]]
 *   Vs                  : double [dim][np_tot[2]+KOFF][np_tot[1]+JOFF][np_tot[0]+IOFF]
 *   InvDt               : double [np_tot[2]][np_tot[1]][np_tot[0]]
 * with np_tot = nx + 2*nghost along active directions (nghost = 2 for PLM, 3 for PPM,
 * src/grid.cpp:60-65) and 1 along the others; IOFF/JOFF/KOFF are 1 along active
 * directions.  Variable order is src/idefix.hpp:74-108 (RHO, VX1..VXc, PRS, BX1..BXc).
 *
 * The library borrows these pointers and never frees them.  Scratch (face and corner
 * EMFs, cMax, the Runge-Kutta "begin" state, the dt reduction scalar) is owned by the
 * context.  All calls are asynchronous on the given HIP stream (a hipStream_t passed as
 * void*, NULL = default stream) except those documented as blocking; calls on one
 * context are not re-entrant.
 */
#ifndef IDEFIX_AMD_H_
#define IDEFIX_AMD_H_

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define IDFX_OK 0
#define IDFX_E_ARGS 1        /* invalid argument or unsupported configuration */
#define IDFX_E_LAUNCH 2      /* a HIP call failed */
#define IDFX_E_NAN 3         /* NaN found (TimeIntegrator's "Nan found" runtime_error) */
#define IDFX_E_DT_SMALL 4    /* dt < 1e-15 (timeIntegrator.cpp:401-405) */
#define IDFX_E_FIRST_DT 5    /* first_dt too large (timeIntegrator.cpp:393-398) */
#define IDFX_E_ALLOC 6       /* device allocation failed */

/* [Hydro] solver (riemannSolver.hpp:97-140) */
#define IDFX_SOLVER_HLLC 1
#define IDFX_SOLVER_HLLD 2
#define IDFX_SOLVER_ROE 3
#define IDFX_SOLVER_HLL 4    /* HllHD / HllMHD; with mhd, emf must not be UCT_HLLD */
#define IDFX_SOLVER_TVDLF 5  /* TvdlfHD / TvdlfMHD; with mhd, emf must not be UCT_HLLD */
/* reconstruction, the reference's ORDER (CMakeLists.txt:173-213) */
#define IDFX_RECON_CONSTANT 1
#define IDFX_RECON_PLM 2
#define IDFX_RECON_LIMO3 3  /* LimO3 (extrapolateToFaces.hpp:170-246), 2 ghost cells */
#define IDFX_RECON_PPM 4
/* No counterpart in the reference: the learned limiter (per-interface MLP weights inside
 * the flux form, PPM fallback), 3 ghost cells, three-component MHD only. */
#define IDFX_RECON_MLP 8
/* [Hydro] emf (constrainedTransport.hpp:156-176) */
#define IDFX_EMF_ARITHMETIC 0
#define IDFX_EMF_UCT0 1
#define IDFX_EMF_UCT_CONTACT 2
#define IDFX_EMF_UCT_HLL 3
#define IDFX_EMF_UCT_HLLD 4
/* [Boundary] types (boundary.hpp:286-385); USERDEF sides are filled by the function enrolled
 * with idfx_ctx_set_userdef_boundary */
#define IDFX_BC_PERIODIC 0
#define IDFX_BC_OUTFLOW 1
#define IDFX_BC_REFLECTIVE 2
#define IDFX_BC_USERDEF 3
/* precision of the stencil, flux and EMF arithmetic (state stays FP64) */
#define IDFX_PREC_FP64 0
#define IDFX_PREC_FP32 1

/* sweep flags (idfx_sweep) */
#define IDFX_SWEEP_INIT_UC 1     /* Uc = PrimToCons(Vc) + rhs instead of Uc += rhs */
#define IDFX_SWEEP_SAVE_UC0 2    /* also store PrimToCons(Vc) in the context's begin state */
#define IDFX_SWEEP_INVDT_SET 4   /* InvDt = term (first direction after ResetStage) */
#define IDFX_SWEEP_INVDT_ADD 8   /* InvDt += term */

typedef struct idfx_ctx idfx_ctx;

typedef struct {
  int dim;        /* DIMENSIONS */
  int comps;      /* COMPONENTS */
  int mhd;        /* 0: hydro, 1: ideal MHD with constrained transport */
  int nx[3];      /* interior cells per direction (1 for inactive directions) */
  int recon;      /* IDFX_RECON_* */
  int solver;     /* IDFX_SOLVER_* */
  int emf;        /* IDFX_EMF_* */
  int precision;  /* IDFX_PREC_* */
  int lbound[3];  /* IDFX_BC_* per direction, left side */
  int rbound[3];  /* IDFX_BC_* per direction, right side */
  double gamma;   /* adiabatic index (eos_adiabatic.hpp:21) */
  int ntracer;    /* passive tracers ([Hydro] tracer, fluid.hpp:315-325), 0..64: Vc/Uc hold
                     nvar = fluid variables + ntracer (Tracer, tracer.hpp:58-111).  Tracers are
                     advanced by idfx_cycle / idfx_stage / idfx_sweep; the flux-array entry
                     points (idfx_calc_flux, idfx_calc_rhs) cover the fluid variables only,
                     and the 1-D Roe HD configuration rejects tracers. */
  int shock_flattening;   /* [Hydro] shockFlattening given (riemannSolver.hpp:151-157):
                             cells flagged by ShockFlattening::FindShock take the first-order
                             slope; PLM and LimO3 only */
  double shock_smoothing; /* its value (the pressure-jump threshold) */
  int isothermal;         /* ISOTHERMAL build (idefix.hpp:63-92): no energy variable, so
                             nvar = 1 + comps (+ comps with mhd); gamma is ignored */
  double csiso;           /* [Hydro] csiso constant <value> (eos_isothermal.hpp:16-27), > 0 */
} idfx_desc;

/* Context: replaces the Fluid<Phys>/RiemannSolver/ConstrainedTransport constructors
 * (fluid.hpp:263-755, riemannSolver.hpp:90-169, constrainedTransport.hpp:152-263).
 * dx1..dx3 are host arrays of np_tot[d] cell widths (DataBlock::dx, gridHost.cpp:79-86).
 * A block holds fewer than 2^29 cells including ghosts and face padding
 * ((np_tot+1) product; e.g. up to 805^3 interior cells with PLM): the kernels address each
 * variable with 32-bit byte offsets.  Larger blocks return IDFX_E_ARGS. */
int idfx_ctx_create(const idfx_desc *desc, const double *dx1, const double *dx2,
                    const double *dx3, idfx_ctx **out);
int idfx_ctx_destroy(idfx_ctx *ctx);
/* np_tot[3], nghost[3], nvar, vs_shape[3] (i, j, k order) */
int idfx_ctx_layout(const idfx_ctx *ctx, int *np_tot, int *nghost, int *nvar, int *vs_shape);
const char *idfx_strerror(int status);
/* Device pointer of a context-owned array for user hooks: which = 0..23 in the order
 * ex ey ez ezi ezj exj exk eyi eyk svx svy svz axL axR ayL ayR azL azR dxL dxR dyL dyR
 * dzL dzR (constrainedTransport.hpp members), 24 = cMax (fluid.hpp cMax).  FP64 only.
 * svx, svy, svz (9..11) are unsigned char [np_tot cells] holding twice the uct_contact
 * upwind weight (0, 1, 2 for 0, 1/2, 1), returned through the double* slot.
 * 25..27: the learned limiter's interface weights of x, y, z (float [16][np_tot cells],
 * returned through the double* slot; IDFX_RECON_MLP contexts only). */
int idfx_ctx_array(idfx_ctx *ctx, int which, double **ptr);

/* Boundary::SetBoundaries (boundary.hpp:239-282): per direction EnforceBoundaryDir
 * (built-in sides, and the enrolled user function on IDFX_BC_USERDEF sides) then
 * ReconstructNormalField (:407-530) on every non-periodic side; last ReconstructVcField(Vc)
 * (:388-404).  IDFX_E_ARGS when a side is IDFX_BC_USERDEF and no function is enrolled
 * (the reference's "No function has been enrolled" error, :325-330). */
int idfx_set_boundaries(idfx_ctx *ctx, double *Vc, double *Vs, void *stream);

/* Boundary::EnforceBoundaryDir (boundary.hpp:285-385) for one direction: periodic, outflow
 * and reflective sides by kernel, IDFX_BC_USERDEF sides by the enrolled function. */
int idfx_enforce_boundary_dir(idfx_ctx *ctx, int dir, double *Vc, double *Vs, void *stream);

/* Boundary::ReconstructNormalField (boundary.hpp:407-530) for one direction: the normal
 * face field of the ghost faces of every non-periodic side from div B = 0. */
int idfx_reconstruct_normal_field(idfx_ctx *ctx, int dir, double *Vs, void *stream);

/* Boundary::EnrollUserDefBoundary (boundary.hpp:548-553) and the call in
 * EnforceBoundaryDir (:316-332, :362-377): fn(arg, dir, side, t, Vc, Vs, stream) fills the
 * ghost zones of side (0 left, 1 right) of direction dir, enqueuing its work on `stream`
 * (the stream of the calling entry point).  t is the stage time: idfx_cycle passes t0,
 * t0 + dt (and t0 + dt/2 in the third RK3 stage), as TimeIntegrator::Cycle does
 * (timeIntegrator.cpp:263-333); callers
 * of idfx_stage / idfx_set_boundaries set it with idfx_ctx_set_time.  fn == NULL
 * unenrolls.  Contexts with user-defined sides run the per-kernel schedule. */
typedef void (*idfx_userdef_boundary_fn)(void *arg, int dir, int side, double t, double *Vc,
                                         double *Vs, void *stream);
int idfx_ctx_set_userdef_boundary(idfx_ctx *ctx, idfx_userdef_boundary_fn fn, void *arg);
int idfx_ctx_set_time(idfx_ctx *ctx, double t);

/* Fluid::ConvertPrimToCons over all cells (convertConsToPrim.hpp:152-189). */
int idfx_prim_to_cons(idfx_ctx *ctx, const double *Vc, double *Uc, void *stream);

/* Fluid::ConvertConsToPrim over all cells, including ReconstructVcField(Uc)
 * (convertConsToPrim.hpp:105-149). */
int idfx_cons_to_prim(idfx_ctx *ctx, double *Uc, double *Vc, const double *Vs, void *stream);

/* Boundary::ReconstructVcField(X) (boundary.hpp:388-404). */
int idfx_reconstruct_vc_field(idfx_ctx *ctx, const double *Vs, double *X, void *stream);

/* Fluid::ResetStage (enroll.hpp:116-129): InvDt = 0. */
int idfx_reset_stage(idfx_ctx *ctx, double *InvDt, void *stream);

/* RiemannSolver::CalcFlux<dir> (calcFlux.hpp:27-85): reconstruction + Riemann flux into
 * FluxRiemann, cMax and the face EMFs of the context (unfused, for flux hooks). */
int idfx_calc_flux(idfx_ctx *ctx, int dir, const double *Vc, const double *Vs, double *Flux,
                   void *stream);

/* Fluid::CalcRightHandSide<dir> (calcRightHandSide.hpp:556-596): Flux *= A, then
 * Uc += -dt/dV (F(i+1)-F(i)) and InvDt += (cMax(i+1)+cMax(i))/(2 dl). */
int idfx_calc_rhs(idfx_ctx *ctx, int dir, double dt, double *Flux, double *Uc, double *InvDt,
                  void *stream);

/* Fused CalcFlux<dir> + CalcRightHandSide<dir>: the flux never reaches memory.  flags
 * are IDFX_SWEEP_*. */
int idfx_sweep(idfx_ctx *ctx, int dir, double dt, const double *Vc, const double *Vs,
               double *Uc, double *InvDt, int flags, void *stream);

/* ConstrainedTransport::CalcCornerEMF (calcCornerEmf.hpp:14-38, calcRiemannEmf.hpp). */
int idfx_corner_emf(idfx_ctx *ctx, const double *Vc, const double *Vs, void *stream);

/* ConstrainedTransport::EvolveMagField (evolveMagField.hpp:14-135). */
int idfx_evolve_mag_field(idfx_ctx *ctx, double dt, double *Vs, void *stream);

/* Fluid::EvolveStage (evolveStage.hpp:40-90): every sweep, CT and ReconstructVcField(Uc),
 * with Uc already holding the stage's initial conserved state. */
int idfx_evolve_stage(idfx_ctx *ctx, double t, double dt, const double *Vc, double *Vs,
                      double *Uc, double *InvDt, void *stream);

/* DataBlock::ComputeTimestep (dataBlock.cpp:333-364): *dtmin = min over the interior of
 * 1/InvDt.  Blocking (returns a host scalar). */
int idfx_compute_timestep(idfx_ctx *ctx, const double *InvDt, double *dtmin, void *stream);

/* No counterpart in the reference: idfx_compute_timestep for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same result as idfx_compute_timestep (blocking), but it calls no allocating routine:
 * its reduction scratch is a static device array of about 32 KiB per device (4096 per-block
 * partial maxima and one result word), so a block may hold at most 2^28 interior cells
 * (IDFX_E_ARGS beyond, with a message on stderr).  Every call on a device shares that
 * device's array: calls that overlap on one GPU (several host threads, or streams running
 * concurrently) race and must be serialised by the caller. */
int idfx_compute_timestep_noalloc(idfx_ctx *ctx, const double *InvDt, double *dtmin,
                                  void *stream);

/* StateContainer::AddAndStore (stateContainer.cpp:75-104): U = wl*U + wr*U0, n doubles. */
int idfx_add_and_store(idfx_ctx *ctx, double wl, double wr, double *U, const double *U0,
                       long long n, void *stream);

/* Fluid::CheckDivB (checkDivB.hpp:14-68).  Blocking. */
int idfx_check_divb(idfx_ctx *ctx, const double *Vs, double *divb, void *stream);

/* One Runge-Kutta stage of TimeIntegrator::Cycle (timeIntegrator.cpp:275-350) on the
 * fused path: built-in boundaries (unless skip_boundaries), fused sweeps with
 * PrimToCons, CT with the stage combination of Vs, and the combined ConvertConsToPrim.
 * After stage 0, *newdt_host = cfl * ComputeTimestep() (blocking in that stage only).
 * Uc receives the combined conserved state when write_uc is set. */
int idfx_stage(idfx_ctx *ctx, int stage, int nstages, double dt, double cfl, double *Vc,
               double *Vs, double *Uc, double *InvDt, int skip_boundaries, int write_uc,
               double *newdt_host, void *stream);

/* TimeIntegrator::Cycle (timeIntegrator.cpp:246-414) on the fused path: all stages and
 * the dt update (CFL_max_var, first-cycle check, dt < 1e-15).  *t and *dt are updated;
 * ncycles is the index of this cycle.  On return the interior cells and interior faces of
 * Vc/Vs (and Uc) hold the new state; with the fused schedules (idfx_ctx_set_fusion: 2-D
 * MHD, and 3-D adiabatic MHD with uct_contact, no tracers, no shock flattening, no
 * user-defined sides, whose stages alternate between Vc and context-owned state arrays)
 * the ghost zones hold the boundary values of an earlier stage, which the next stage's
 * boundary fill rewrites before any use, and the corner (2-D also y-face) EMF scratch
 * arrays are not written.  The first 3-D RK3 cycle allocates one more state array. */
int idfx_cycle(idfx_ctx *ctx, int nstages, double cfl, double cfl_max_var, long ncycles,
               double *t, double *dt, double *Vc, double *Vs, double *Uc, double *InvDt,
               void *stream);

/* No counterpart in the reference: selects the kernel schedule idfx_cycle uses.
 * level 0 runs each reference step as its own kernel (sweeps, CalcCornerEMF,
 * EvolveMagField, ConvertConsToPrim); level 1 (the default where supported) fuses, in 2-D
 * MHD without user-defined boundaries, the y sweep, corner EMF, face-field update,
 * AddAndStore and ConvertConsToPrim into one kernel, and in 3-D MHD with uct_contact the
 * corner EMF and the face-field update (with its AddAndStore) into one kernel, which then
 * leaves the edge EMF arrays unwritten and the face fields of ghost columns as they were;
 * for adiabatic 3-D MHD (no tracers, no shock flattening, no user-defined sides) the last
 * sweep also applies AddAndStore and the field-independent part of ConvertConsToPrim into
 * the next state array, and a pointwise kernel adds pressure and cell field.
 * In 3-D, level 1 takes the fused CT kernel on blocks of at least 2^26 interior cells
 * and k_corner_emf + k_evolve_b below (the faster choice on MI355X); level 2 takes the
 * fused CT kernel at every size.  All levels give identical interior cells, faces and Uc.
 * Returns IDFX_E_ARGS for levels 1 and 2 where the configuration supports neither fused
 * schedule, and for other levels. */
int idfx_ctx_set_fusion(idfx_ctx *ctx, int level);

/* Regions of idfx_cycle / idfx_stage timed by the region profiler. */
enum {
  IDFX_REGION_BOUNDARIES = 0,   /* SetBoundaries (ghost fill, normal field, cell B) */
  IDFX_REGION_SWEEP_X = 1,      /* CalcFlux + CalcRightHandSide along x (y: 2, z: 3) */
  IDFX_REGION_SWEEP_Y = 2,
  IDFX_REGION_SWEEP_Z = 3,
  IDFX_REGION_CORNER_EMF = 4,   /* CalcCornerEMF */
  IDFX_REGION_EVOLVE_B = 5,     /* EvolveMagField (+ AddAndStore of Vs) */
  IDFX_REGION_CONS_TO_PRIM = 6, /* AddAndStore of Uc + ConvertConsToPrim */
  IDFX_REGION_FUSED2D_Y = 7,    /* fused 2-D second half (fusion level 1) */
  IDFX_REGION_CT_FUSED = 8,     /* fused 3-D CalcCornerEMF + EvolveMagField (fusion level 1) */
  IDFX_REGION_MLP = 9,          /* learned limiter: interface weights of one direction */
  IDFX_NREGIONS = 10
};

/* idfx::pushRegion / popRegion with the profiler enabled (global.cpp:66-95, profiler.cpp):
 * when on, every region of idfx_cycle / idfx_stage is bracketed by a HIP event pair on the
 * launch stream.  idfx_ctx_region_times waits for the recorded regions, returns the total
 * milliseconds and launch count of the first n regions, and resets the totals. */
/* EquationOfState with "csiso userdef" (eos_isothermal.hpp:29-33 and GetWaveSpeed at
 * :70-75; filled by the function of Fluid::EnrollIsoSoundSpeed, enroll.hpp:13-20, at
 * EquationOfState::Refresh): cs is a device array of np_tot cells ([k][j][i], ghosts
 * included) holding every cell's sound speed.  It is borrowed, not copied: the sweeps
 * read it at each launch (face value: the average of the two cells', hllcHD.hpp:67-70,
 * hlldMHD.hpp:188-191; shock flags: each cell's own, shockFlattening.hpp:96-128), so a
 * caller refreshes it before a stage as the reference does.  NULL returns to the
 * constant csiso of the description.  IDFX_E_ARGS for a context that is not isothermal. */
int idfx_ctx_set_iso_sound_speed(idfx_ctx *ctx, const double *cs);

int idfx_ctx_set_profiling(idfx_ctx *ctx, int on);
int idfx_ctx_region_times(idfx_ctx *ctx, double *ms, long *count, int n);

/* ---- learned limiter (no counterpart in the reference) ------------------------------
 * An interface x_{i-1/2} sees the stencil i-3 .. i+2 of the 8 face-frame primitives
 * (rho, v_n, v_t, v_b, p, B_n, B_t, B_b).  Features per variable: (V_j - c) / s with
 * c = (V[i-1] + V[i]) / 2, s = max_j |V_j - c|.  Network 48 -> 64 -> 64 -> 16, hidden
 * activation z * sigmoid(1.702 z), linear outputs z3 that shift the PPM member of the
 * limiter family: wL_v = wL_PPM + z3[2v], wR_v = wR_PPM - z3[2v] (output 2v+1 unused),
 * w_PPM = (v_PPM - V[i-1]) / (V[i] - V[i-1]); states vL = V[i-1] + wL (V[i] - V[i-1]),
 * vR = V[i] - wR (V[i] - V[i-1]), so the jump vR - vL stays PPM's.  Weights outside the
 * TVD region fall back to PPM's.
 * Parameters are packed FP32, row-major [in][out]: W1 48x64, b1 64, W2 64x64, b2 64,
 * W3 64x16, b3 16 (idfx_mlp_param_count() floats).  Callers: setups that run
 * IDFX_RECON_MLP and the training pipeline that fits the network to block-averaged
 * high-resolution runs. */
typedef struct idfx_mlp idfx_mlp;
int idfx_mlp_param_count(void);
/* Training state (FP32 parameters, Adam moments, activations for batches up to
 * max_batch rows).  Allocates device memory. */
int idfx_mlp_create(int max_batch, idfx_mlp **out);
int idfx_mlp_destroy(idfx_mlp *h);
/* Host arrays of idfx_mlp_param_count() floats; set_params also resets Adam. */
int idfx_mlp_set_params(idfx_mlp *h, const float *params);
int idfx_mlp_get_params(idfx_mlp *h, float *params);
/* X: device [B][72] rows: the 48 features, the interface's 16 PPM weights and 8 loss
 * weights om_v; tau: device [B][8] normalised targets (t - c) / s.  Loss = mean over rows
 * of sum_v om_v^2 (eL_v^2 + eR_v^2), eL = vL/s-error of the left state, eR of the right.
 * loss_grad: blocking; mean loss and the gradient of every parameter (host arrays).
 * train_step: one Adam step (lr, beta 0.9 / 0.999, eps 1e-8), asynchronous;
 * idfx_mlp_loss returns the loss of the last step's batch (blocking). */
int idfx_mlp_loss_grad(idfx_mlp *h, const float *X, const float *tau, int B, double *loss,
                       float *grad, void *stream);
int idfx_mlp_train_step(idfx_mlp *h, const float *X, const float *tau, int B, float lr,
                        void *stream);
int idfx_mlp_loss(idfx_mlp *h, double *loss, void *stream);
/* Loads packed FP32 parameters (host) into an IDFX_RECON_MLP context as BF16 for MFMA
 * inference; params == NULL unloads them (every interface then takes the PPM weights). */
int idfx_ctx_set_mlp_params(idfx_ctx *ctx, const float *params);
/* Interface weights of direction dir for the faces beg..end along dir (every transverse
 * index) from Vc, into W (device float [16][np_tot cells], or the context's own array
 * when W is NULL).  The sweeps of an IDFX_RECON_MLP context call this themselves. */
int idfx_mlp_faces(idfx_ctx *ctx, int dir, const double *Vc, float *W, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* IDEFIX_AMD_H_ */
