/*
 * CPU oracle for the per-stage finite-volume update (TEST INFRASTRUCTURE ONLY).
 *
 * This is a plain CPU restatement of the reference's algorithm, used only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the reference
 * results the GPU path is compared with (and, in the bench, as the CPU baseline).
 * It is never linked into, loaded by or called from the product library.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * reference checkout, Idefix 2.2.02). See oracle.cpp for the pinning status of each
 * function.
 */
#ifndef IDFX_ORACLE_H_
#define IDFX_ORACLE_H_

#ifdef __cplusplus
extern "C" {
#endif

/* Reconstruction order, same values as the reference's ORDER macro. */
enum { ORC_RECON_CONSTANT = 1, ORC_RECON_PLM = 2, ORC_RECON_LIMO3 = 3, ORC_RECON_PPM = 4 };
/* Riemann solvers ([Hydro] solver). ROE, HLL and TVDLF are the HD or MHD variant
   depending on mhd. */
enum { ORC_SOLVER_HLLC = 1, ORC_SOLVER_HLLD = 2, ORC_SOLVER_ROE = 3, ORC_SOLVER_HLL = 4,
       ORC_SOLVER_TVDLF = 5 };
/* EMF averaging ([Hydro] emf). */
enum { ORC_EMF_ARITHMETIC = 0, ORC_EMF_UCT0 = 1, ORC_EMF_UCT_CONTACT = 2,
       ORC_EMF_UCT_HLL = 3, ORC_EMF_UCT_HLLD = 4 };
/* Boundary types ([Boundary] X?-beg / X?-end). */
enum { ORC_BC_PERIODIC = 0, ORC_BC_OUTFLOW = 1, ORC_BC_REFLECTIVE = 2, ORC_BC_USERDEF = 3 };

/* Arrays exposed for inspection. */
enum { ORC_VC = 0, ORC_VS = 1, ORC_UC = 2, ORC_INVDT = 3, ORC_CMAX = 4, ORC_FLUX = 5,
       ORC_EX = 6, ORC_EY = 7, ORC_EZ = 8,
       ORC_EZI = 9, ORC_EZJ = 10, ORC_EXJ = 11, ORC_EXK = 12, ORC_EYI = 13, ORC_EYK = 14,
       ORC_SVX = 15, ORC_SVY = 16, ORC_SVZ = 17,
       ORC_AXL = 18, ORC_AXR = 19, ORC_AYL = 20, ORC_AYR = 21, ORC_AZL = 22, ORC_AZR = 23,
       ORC_DXL = 24, ORC_DXR = 25, ORC_DYL = 26, ORC_DYR = 27, ORC_DZL = 28, ORC_DZR = 29,
       ORC_UC0 = 30, ORC_VS0 = 31, ORC_NARRAYS = 32 };

typedef struct {
  int dim;            /* DIMENSIONS 1..3 */
  int comps;          /* COMPONENTS, dim..3 */
  int mhd;            /* 0 = hydro, 1 = ideal MHD with CT */
  int order;          /* ORC_RECON_* */
  int solver;         /* ORC_SOLVER_* */
  int emf;            /* ORC_EMF_* */
  int nstages;        /* RK stages 1..3 */
  int nx[3];          /* interior cells per direction (1 for unused directions) */
  double xbeg[3];
  double xend[3];
  int lbound[3];
  int rbound[3];
  double gamma;
  double cfl;
  double cfl_max_var;
  double first_dt;
  int ntracer;        /* passive tracers ([Hydro] tracer, fluid.hpp:315-325) */
  int shock_flattening;   /* [Hydro] shockFlattening given (riemannSolver.hpp:151-157) */
  double shock_smoothing; /* its value */
  int isothermal;         /* ISOTHERMAL build: no energy variable (idefix.hpp:63-92) */
  double csiso;           /* [Hydro] csiso constant <value> (eos_isothermal.hpp:16-27) */
} orc_cfg;

void *orc_new(const orc_cfg *cfg);
void orc_free(void *o);
int orc_nthreads(int n);  /* set OpenMP threads (returns the previous value) */

/* Geometry / layout */
void orc_layout(void *o, int *np_tot, int *nghost, int *nvar, int *vs_shape);
double *orc_array(void *o, int which, long *nelem);
double *orc_coord(void *o, int dir, int kind); /* kind 0:x 1:xl 2:xr 3:dx */

/* User-defined sides (boundary.hpp:316-332): fn(arg, dir, side, t) fills the ghost zones of
 * that side in the arrays orc_array exposes; called inside SetBoundaries, before the normal
 * field of the direction is reconstructed. */
typedef void (*orc_userdef_fn)(void *arg, int dir, int side, double t);
void orc_set_userdef(void *o, orc_userdef_fn fn, void *arg);

/* Individual steps of the stage (same granularity as the reference kernels) */
void orc_set_boundaries(void *o);
/* per-cell isothermal sound speed (np_tot cells, copied), NULL: the constant csiso;
 * returns nonzero for a configuration that is not isothermal */
int orc_set_iso_sound_speed(void *o, const double *cs);
void orc_prim_to_cons(void *o);
void orc_cons_to_prim(void *o);
void orc_reset_stage(void *o);
void orc_calc_flux(void *o, int dir);
/* ShockFlattening::FindShock on the current Vc; copies the flags (np_tot cells) to out */
void orc_find_shock(void *o, unsigned char *out);
void orc_calc_rhs(void *o, int dir, double dt);
void orc_calc_corner_emf(void *o);
void orc_evolve_mag_field(void *o, double dt);
void orc_reconstruct_vc_field_uc(void *o);
void orc_evolve_stage(void *o, double t, double dt);
double orc_compute_timestep(void *o);
double orc_check_divb(void *o);
void orc_add_and_store(void *o, double wl, double wr);
void orc_copy_begin(void *o);

/* Driver: one full cycle / the main loop */
int orc_cycle(void *o);
int orc_run(void *o, double tstop, long maxcycles);
double orc_get_t(void *o);
double orc_get_dt(void *o);
void orc_set_t(void *o, double t);
void orc_set_dt(void *o, double dt);
long orc_get_ncycles(void *o);

/* Negative controls: switch on one planted mistake (0 = none) so that a test can show
 * that its check fails when a restated branch is wrong; returns the previous value. */
enum { ORC_PLANT_NONE = 0, ORC_PLANT_HLLD_SBX = 1, ORC_PLANT_CONTACT_SV = 2,
       ORC_PLANT_UCT0_INDEX = 3, ORC_PLANT_RIEMANN_PHI = 4, ORC_PLANT_PPM_EXTREMUM = 5,
       ORC_PLANT_ROE_ALFVEN = 6 };
int orc_plant(int which);

/* Pointwise kernels (for unit tests of the algebra) */
int orc_face_flux(const orc_cfg *cfg, int dir, const double *vL, const double *vR,
                  double *flux, double *cmax, double *emf7);
double orc_vanleer(double dvp, double dvm);
double orc_mclim(double dvp, double dvm);
double orc_limo3lim(double dvp, double dvm, double dx);
void orc_ppm_states(double vm2, double vm1, double v0, double vp1, double vp2,
                    double *vl, double *vr);
void orc_prim2cons_pt(const orc_cfg *cfg, const double *v, double *u);
void orc_cons2prim_pt(const orc_cfg *cfg, const double *u, double *v);

#ifdef __cplusplus
}
#endif
#endif  /* IDFX_ORACLE_H_ */
