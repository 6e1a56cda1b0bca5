/* suli.h - C ABI of the MI355X-native SuLi-IC time step (libsuli.so).
 *
 * Drop-in boundary for the solver loop of the reference Fortran driver
 * (SuLi-IC/1_principal.f90:64-108).  Each entry point replaces one
 * argument-less reference subroutine that works on module globals; here the
 * state lives on one GPU, allocated by suli_init and freed by suli_finalize,
 * and the host exchanges whole arrays in the reference's own extents and
 * Fortran order (0_modules.f90:87-135,177-201) through suli_upload /
 * suli_download.  The driver's mutable scalar dt is passed by value.
 *
 * Conventions: every function returns 0 on success and a nonzero status on
 * error; the message is available from suli_last_error.  One host thread,
 * one HIP stream; calls are serialised and the library is not re-entrant.
 * Calls that return scalars synchronise the stream; the others may return
 * before the GPU work has finished.
 */
#ifndef SULI_H
#define SULI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* precision of the fields and stencils */
#define SULI_PREC_FP64 0   /* everything in FP64 (parity with the reference)  */
#define SULI_PREC_MIXED 1  /* FP32 fields, stencils and Krylov sweeps; FP64
                              pressure, residuals, dot products, integrals    */

/* pressure solver of suli_graddin */
#define SULI_POISSON_REF 0 /* the reference CG: FP64, scaled residual sum of
                              squares <= tol_abs (default 1e-4/(nx*ny*nz)),
                              warm start, at most max_it iterations
                              (9_grad.f90:175)                                */
#define SULI_POISSON_FCG 1 /* flexible CG with FP64 iterative refinement,
                              stopped on the FP64 true relative residual
                              ||D^-1/2 (b - A p)|| / ||D^-1/2 b|| <= tol_rel  */

/* preconditioners of SULI_POISSON_FCG */
#define SULI_PRECOND_JACOBI 0  /* symmetric diagonal scaling (the reference's) */
#define SULI_PRECOND_GMG 1     /* geometric multigrid V-cycle                  */
#define SULI_PRECOND_LEARNED 2 /* 3-D convolutional U-Net, falls back to GMG   */

typedef struct {
  /* grid (0_modules.f90:7,15) */
  int32_t nx, ny, nz;
  double dx, dy, dz;
  /* wall codes (0_modules.f90:53-61): 0 periodic, 1 free slip, 2 no slip,
     3 prescribed velocity (ghost and wall-normal faces from the boundary
     planes SULI_BXX0 .. SULI_BZZF1), 4 (x walls only) outflow at xf and at
     x0 the inflow of 3_contorno.f90:143-146 (v from bxy0, w copied) */
  int32_t ccx0, ccxf, ccy0, ccyf, ccz0, cczf;
  /* model switches: m_turb 0 none, 1 Smagorinsky, 2 directional Smagorinsky (0_modules.f90:31) */
  int32_t m_turb;
  /* fluid and model constants (14_level_set.f90:25-30, 0_modules.f90:141-143,181) */
  double rho_f1, rho_f2, mi_f1, mi_f2, sigma, alpha1;
  double csmag, gx, gz, chezy;
  /* numerics */
  int32_t precision;  /* SULI_PREC_* */
  int32_t poisson;    /* SULI_POISSON_* */
  int32_t precond;    /* SULI_PRECOND_* */
  int32_t max_it;     /* Krylov iteration cap (0: 10000) */
  double tol_abs;     /* SULI_POISSON_REF threshold (0: 1e-4/(nx*ny*nz)) */
  double tol_rel;     /* SULI_POISSON_FCG threshold (0: 1e-8) */
  int32_t device;     /* HIP device ordinal */
} suli_params;

/* Field identifiers of suli_upload / suli_download; the host array has the
 * reference extents (name: Fortran declaration). */
enum {
  SULI_U = 0,      /* u(0:nx+2,0:ny+1,0:nz+1)  */
  SULI_V = 1,      /* v(0:nx+1,0:ny+2,0:nz+1)  */
  SULI_W = 2,      /* w(0:nx+1,0:ny+1,0:nz+2)  */
  SULI_PRD1 = 3,   /* prd1(0:nx+1,0:ny+1,0:nz+1) */
  SULI_RHO = 4,    /* rho(nx,ny,nz)            */
  SULI_LS_NU = 5,  /* ls_nu(nx,ny,nz)          */
  SULI_LS = 6,     /* ls(nx,ny,nz): level set, positive in water */
  SULI_MOD_LS = 7,
  SULI_KURV = 8,
  SULI_HS = 9,
  SULI_HSX = 10,
  SULI_HSY = 11,
  SULI_HSZ = 12,
  SULI_DDLSDX = 13,
  SULI_DDLSDY = 14,
  SULI_DDLSDZ = 15,
  SULI_NUT = 16,   /* nut(nx,ny,nz)            */
  SULI_XNUT = 17,  /* xnut(nx+1,ny,nz): download only, rebuilt from nut, ls_nu */
  SULI_YNUT = 18,
  SULI_ZNUT = 19,
  SULI_FU = 20,    /* Fu(nx+1,ny,nz): download only, after suli_convdiff_forces */
  SULI_FV = 21,
  SULI_FW = 22,
  SULI_U0 = 23,    /* u0(nx+1,ny,nz)           */
  SULI_V0 = 24,
  SULI_W0 = 25,
  SULI_TFU = 26,   /* tf_u(nx+1,ny,nz): momentum source of convdiff (6_convdiff.f90:156),
                      the manufactured-solution forcing; zero until uploaded */
  SULI_TFV = 27,
  SULI_TFW = 28,
  SULI_FU0 = 29,   /* fu0(nx+1,ny,nz): forces kept by the RK3 and AB2 schemes */
  SULI_FV0 = 30,
  SULI_FW0 = 31,
  SULI_FU1 = 32,   /* fu1(nx+1,ny,nz): stage-2 forces of the RK3 scheme */
  SULI_FV1 = 33,
  SULI_FW1 = 34,
  SULI_NFIELDS = 35,
  /* boundary planes of the prescribed-velocity walls (wall code 3; bxy0 also
   * feeds code 4 at x = 0), 0_modules.f90:96-120, each in its module extents
   * from index 0, e.g. bxx0(0:ny+1,0:nz+1), bxy0(0:ny+2,0:nz+1),
   * byx0(0:nx+2,0:nz+1), bzz0(0:nx+1,0:ny+1); zero until uploaded or set by
   * the manufactured solution (suli_mms_bc).  suli_field_len gives the count. */
  SULI_BXX0 = 64, SULI_BXX1, SULI_BXY0, SULI_BXZ0, SULI_BXXF, SULI_BXXF1, SULI_BXYF, SULI_BXZF,
  SULI_BYX0, SULI_BYY0, SULI_BYY1, SULI_BYZ0, SULI_BYXF, SULI_BYYF, SULI_BYYF1, SULI_BYZF,
  SULI_BZX0, SULI_BZY0, SULI_BZZ0, SULI_BZZ1, SULI_BZXF, SULI_BZYF, SULI_BZZF, SULI_BZZF1
};

/* Time-integration schemes of suli_tempo (t_tempo, 0_modules.f90:24). */
enum {
  SULI_TIME_EULER = 0, /* one stage, u += dt F                           */
  SULI_TIME_RK2 = 1,   /* stages dt/2, dt (default)                      */
  SULI_TIME_RK3 = 2,   /* stages dt/2, dt, dt                            */
  SULI_TIME_AB2 = 3    /* one stage, Adams-Bashforth 2 (Euler first)     */
};

typedef struct {
  int32_t iters;          /* Krylov iterations (inner iterations for FCG) */
  int32_t outer;          /* FP64 refinement passes (FCG)                 */
  int32_t fallbacks;      /* learned-preconditioner fallbacks to GMG      */
  int32_t converged;      /* 1 if the stopping rule was met               */
  double res0;            /* initial residual (REF: sum of squares; FCG: relative 2-norm) */
  double res;             /* final residual, same measure                 */
  double ms;              /* GPU time of the solve: stream events at its
                             start and after its last kernel             */
} suli_poisson_stats;

typedef struct {
  double vol_ini, vol_ins;  /* level-set volume, initial and current     */
  int32_t reinit_iters;     /* reinitialisation iterations (1..3)        */
  int32_t mass_corr_iters;  /* volume-correction passes                  */
} suli_ls_stats;

/* lifetime */
int suli_init(const suli_params *p);
int suli_finalize(void);
const char *suli_version(void);
int suli_last_error(char *buf, int len);

/* host <-> device, arrays in reference extents and Fortran order (FP64) */
int suli_upload(int field, const double *host, int64_t n);
int suli_download(int field, double *host, int64_t n);
int64_t suli_field_len(int field);

/* Tail of level_set_ini (14_level_set.f90:205-217): heaviside, initial
 * volume, mod_ls1, for a level set uploaded with suli_upload(SULI_LS). */
int suli_level_set_ini(double *vol_ini);
/* Reference volume of the level set (ls_param vol_ini, 0_modules.f90:194),
 * for a state restored from the host instead of built by suli_level_set_ini. */
int suli_set_vol_ini(double vol_ini);
int suli_get_volumes(double *vol_ini, double *vol_ins);
/* Dam-break initial level set: signed distance (positive in water) to a
 * column x <= width, z <= height spanning y; then suli_level_set_ini. */
int suli_init_dambreak(double width, double height, double *vol_ini);

/* the time step, one call per reference subroutine (1_principal.f90:74-92) */
int suli_level_set(double dt, suli_ls_stats *st);  /* level_set, 14_level_set.f90:221 */
int suli_contorno(int nlock);                      /* contorno, 3_contorno.f90:10      */
int suli_visco(void);                              /* visco, 5_viscosidade.f90:12      */
int suli_convdiff(void);                           /* convdiff, 6_convdiff.f90:14      */
int suli_tempo(int tt, double dt);                 /* tempo, 11_diversos.f90:10        */
int suli_graddin(double dt, suli_poisson_stats *st);/* graddin, 9_grad.f90:10          */
int suli_posdin(double dt);                        /* posdin, 10_pos.f90:1             */
int suli_est_div(double *maxdiv);                  /* est, 4_plot.f90:500              */

/* No counterpart in the reference: a variant of suli_est_div for callers
 * that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  max |div u| of est (4_plot.f90:500-525)
 * over the caller's DEVICE arrays in the reference extents and Fortran
 * order, FP64: u(0:nx+2,0:ny+1,0:nz+1), v(0:nx+1,0:ny+2,0:nz+1),
 * w(0:nx+1,0:ny+1,0:nz+2); the result is written to the device memory
 * *div_max.  Needs no suli_init; enqueued on `stream` (null: the null
 * stream) and returns without synchronising.  Allocates nothing: the block
 * partials go to a static device buffer of 2048 doubles per GPU (16 KiB,
 * held until the process ends), shared by every call on that GPU, so calls
 * that overlap on one GPU race.  The first call on a device (or suli_init
 * on it) looks up the buffer's address: make one of them before capturing
 * a graph.  Grids: each extent >= 1, ny*nz <= 2^30, velocity arrays of at
 * most 2^36 elements; others are rejected with an error. */
int suli_est_div_noalloc(const double *u, const double *v, const double *w, int nx, int ny, int nz,
                         double dx, double dy, double dz, double *div_max, void *stream);


/* Fu, Fv, Fw of convdiff materialised for inspection (suli_convdiff leaves
 * them implicit: the forces are fused into suli_tempo's update). */
int suli_convdiff_forces(void);

/* One stage (visco .. contorno(1)) and one whole time step
 * (1_principal.f90:74-95) of the selected time scheme, stage steps a_dt(tt)
 * of 2_iniciais.f90:85-102 (RK2: dt/2, dt; RK3: dt/2, dt, dt; Euler and
 * AB2: dt).  suli_step fills ps[0 .. stages-1], one entry per stage
 * (suli_time_stages: at most 3). */
int suli_stage(int tt, double dt_base, suli_poisson_stats *st);
int suli_step(double dt, suli_ls_stats *ls, suli_poisson_stats *ps);

/* Select the time scheme of later suli_tempo / suli_stage / suli_step calls
 * (t_tempo of 0_modules.f90:24, a compile-time parameter in the reference;
 * SULI_TIME_RK2 until set).  tempo (11_diversos.f90:10-228) per scheme:
 *   EULER  u = u + dt F
 *   RK2    tt 1: u0 = u, u = u + dt F; tt 2: u = u0 + dt F
 *   RK3    tt 1: u0 = u, fu0 = F, u = u + dt F;
 *          tt 2: fu1 = F, u = u0 - dt fu0 + 2 dt F;
 *          tt 3: u = u0 + dt/6 (fu0 + 4 fu1 + F)
 *   AB2    u = u + dt/2 (3 F - fu0), fu0 = F; the first suli_tempo after
 *          this call is the Euler update (the reference's it = 1) and keeps F
 *          in fu0.
 * Selecting a scheme restarts the AB2 history. */
int suli_set_time_scheme(int t_tempo);
int suli_time_stages(void); /* ntt of the selected scheme (1, 2 or 3), -1 without a context */

/* Select the advective term of later convdiff calls (adv_type and der of
 * 0_modules.f90:26-27, compile-time parameters in the reference; 1, 1 until
 * set): adv_type 1 classico with der 1 upwind (7_advectivo.f90:33-113),
 * 2 central (:116-173) or 3 second-order upwind (:176-291); adv_type 2
 * rotacional (:296-440) and 3 antissim (:443-653), der 1 only (the
 * reference stops on the other pairs).  The turbulence model is
 * suli_params.m_turb: 0 none, 1 Smagorinsky, 2 directional Smagorinsky
 * (5_viscosidade.f90:104-170: face viscosities xnut, ynut, znut stored and
 * downloadable, nut left unchanged). */
int suli_set_advection(int adv_type, int der);

/* Bottom obstacle (obst_t != 0 in the reference, 0_modules.f90:29,76-79):
 * the columns ku(0:nx+2,0:ny+1), kv(0:nx+1,0:ny+2), kw(0:nx+1,0:ny+1), Fortran
 * order, as obstaculo leaves them (3_contorno.f90:741-760: -1 <= ku, kv <= nz,
 * -1 <= kw <= nz+1).  contorno(1), contorno(2) zero the velocities of each
 * column up to k = ku(i,j) and take the friction plane of the column at
 * ku(i,j)+1 instead of the wall planes (3_contorno.f90:266-291); contorno(3)
 * fills the level set of the cells below kw (:445-490) and recomputes the
 * heaviside fields; the field output marks those cells in kaux; the restart
 * file carries the columns.  All three null removes the obstacle.  The
 * reference reads the level set outside its array where an obstacle column
 * on the first or last row or column stands beside a lower ghost column;
 * the library reads the ghost cell there. */
int suli_set_obstacle(const int32_t *ku, const int32_t *kv, const int32_t *kw);

/* obstaculo (3_contorno.f90:494-760) on the host, no context needed: the
 * columns ku, kv, kw (extents as suli_set_obstacle) of the reference's
 * obstacle geometry obst_t = 1 (sinusoidal dunes), 2 (Yue et al. dune),
 * 3 (Gaussian bump), 4 (Beji and Battjes bar), 5, 6 (Delft 1980 channels),
 * 7 (hole and hump) or 10 (Koshizuka / Kleefsman box) on an nx*ny*nz grid of
 * spacing dx, dy, dz.  8 and 9 read unset heights in the reference and are
 * refused.  Entries may be -1 (a column cut below the floor). */
int suli_obstacle_columns(int obst_t, int nx, int ny, int nz, double dx, double dy, double dz, int32_t *ku,
                          int32_t *kv, int32_t *kw);

/* Time-dependent horizontal body force (sloshing tanks): convdiff uses
 * gx(t) = gx + amplitude*sin(2 pi t/period + phase) in place of the constant
 * gx of 6_convdiff.f90:155, with t the time of the current step.
 * suli_set_time sets t (the driver's t = it*dt, 1_principal.f90:65);
 * suli_step advances it by dt before the step. */
int suli_set_forcing(double amplitude, double period, double phase);
int suli_set_time(double t);

/* The manufactured solution of Weng (2009) (13_mms.f90), computed on the
 * device at the time of suli_set_time / suli_step (the reference's t).
 * suli_mms_set: mms_t of 0_modules.f90:37 (0 off, 1 steady, 2 unsteady) and
 * the constants a, h0, coef of module mms_m (0_modules.f90:208-212); needs
 * h0 > |a|.  In this mode suli_stage runs no projection (1_principal.f90:89-93),
 * suli_step computes the sources first (steady: at its first step,
 * unsteady: every step; 1_principal.f90:70-71), the level set of the
 * unsteady case corrects no volume (14_level_set.f90:257), and contorno(1|2)
 * set the boundary planes first (3_contorno.f90:30).
 *   suli_mms_i         mms_i (13_mms.f90:12-212): level set h - z, the
 *                      interior-face velocities and (unsteady) the pressure;
 *                      follow with suli_level_set_ini (level_set_ini's tail)
 *   suli_mms_bc        mms_bc (:468-816): the 24 boundary planes (steady:
 *                      only at t = 0); in the unsteady case bzy0 and bzyf
 *                      take the formula of w, as in the reference
 *   suli_termo_fonte1  termo_fonte1 (:819-1064) and termo_fonte2
 *   suli_termo_fonte2  (:1070-1322): the momentum sources tf_u, tf_v, tf_w,
 *                      exact derivatives of the manufactured fields; the
 *                      reference's hand-derived ones agree for a = 0 and
 *                      differ for a != 0
 *   suli_mms           mms (:218-465): err[0..3] = RMS error of u, v, w over
 *                      the faces whose interpolated level set is positive and
 *                      of the pressure over the cells with positive level set
 *                      (its squared errors kept between calls as the
 *                      reference's erro_p); the steady case has no exact
 *                      pressure: err[3] = NaN */
int suli_mms_set(int mms_t, double a, double h0, double coef);
int suli_mms_i(void);
int suli_mms_bc(void);
int suli_termo_fonte1(void);
int suli_termo_fonte2(void);
int suli_mms(double *err);

/* Wave inflow through the prescribed wall x = 0 (ccx0 = 3; wave_t of
 * 0_modules.f90:35).  suli_waves_coef: waves_coef (3_contorno.f90:1070-1205):
 * wave_t 0 (off), 1 Stokes I, 2 Stokes II, 5 Stokes V (Fenton's fifth-order
 * theory) for amplitude a_w, period p_w and still depth h0_f (the
 * reference's waves_coef fixes 0.03 m, 2.5 s, 1 m) and the params' gravity
 * gz; coef (may be null, 14 doubles) receives the wavelength, wave number,
 * angular frequency, celerity and the Stokes V coefficients aeta1..5,
 * avel1..5 (zero otherwise).  Stokes V scans 10000 wavelengths as the
 * reference does (host threads, about a second).  suli_boundary_waves:
 * boundary_waves (:952-1068) at the library's time: the planes bxx0, bxx1,
 * bxz0 of the wave's u and w at x = -dx/2 .. 0 (zero above its surface;
 * bxy0 zero), the level set of the first cell column from the wave's
 * surface and, as the reference's single-precision copy does, the rest of
 * the level set rounded to single precision.  suli_stage calls it after
 * tempo when a wave is set (1_principal.f90:84). */
int suli_waves_coef(int wave_t, double a_w, double p_w, double h0_f, double *coef);
int suli_boundary_waves(void);

/* Switch the pressure solver of later suli_graddin calls (the fields stay
 * as they are).  tol_rel <= 0 or max_it <= 0 keeps the current value.  No
 * counterpart in the reference: its solver is fixed at compile time. */
int suli_set_solver(int poisson, int precond, double tol_rel, int max_it);

/* Multigrid tuning, no counterpart in the reference.  Levels below the
 * finest with at most fuse_max_cells cells run all the half sweeps of one
 * side of the V-cycle in one launch (LDS tiles with a halo; the same cell
 * updates as one launch per half sweep; 0 turns it off; default 65536).
 * coarse_sweeps: red-black Gauss-Seidel sweeps on the coarsest level
 * (default 24).  fuse_fine 1: the operator set-up runs in two passes (cells,
 * faces) and the input and output scaling of the V-cycle, the prolongations
 * and the z.q sum of FCG run inside the neighbouring half sweeps; 2 (default):
 * as 1, with the finest level's iterate and right-hand side held in FP16 in
 * the V-cycles of the FCG iteration (their input scaled to a unit RMS
 * residual; the preconditioner changes by about 2^-11 relative, the solve
 * still meets its FP64 tolerance); 0: separate launches.  A negative value
 * keeps the current setting; a value above 2 is an error. */
int suli_set_mg(int64_t fuse_max_cells, int coarse_sweeps, int fuse_fine);

/* Stencil kernel form, no counterpart in the reference.  on = 1 (default):
 * the momentum predictor (convdiff + tempo, 6_convdiff.f90:147-186) of the
 * default advection form (adv_type 1, der 1, m_turb 1) runs its inner box
 * (1 <= i <= nx, 2 <= j <= ny-1, 2 <= k <= nz-1) in LDS tiles that read each
 * input once per tile; 0: every stencil value read from memory per cell.
 * Both forms compute the same values in the same order.  A negative value
 * keeps the setting. */
int suli_set_fast_stencils(int on);

/* z = M r with the preconditioner of the current solver setting and the
 * operator of the last FCG suli_graddin (its dt and density): r and z are
 * scaled residuals in cell order (i fastest, nx*ny*nz values), computed in
 * FP32.  For tests of the preconditioners; no counterpart in the
 * reference. */
int suli_precond_apply(const double *r, double *z, int64_t n);

/* Reads of the FCG loop's control block by the host since the solver was
 * set up: published = those that found the copy the finish kernel wrote
 * into pinned host memory, copied = those that fell back to a device-to-host
 * copy.  Diagnostics, no counterpart in the reference. */
int suli_fcg_readbacks(int64_t *published, int64_t *copied);

/* Learned preconditioner (SULI_PRECOND_LEARNED).  No counterpart in the
 * reference (9_grad.f90 is Jacobi-scaled CG).  An L-level 3-D convolutional
 * U-Net (channels 16/32/64/64.., 3x3x3 kernels, ReLU, average pooling, nearest
 * upsampling with additive skips) maps (r_s/rms(r_s), normalised log rho) to
 * the search direction z; FCG is invariant to the scale of z, and each new
 * direction is A-orthogonalised against the last four (FCG(4)), which keeps
 * the iteration convergent for a nonlinear preconditioner.  Weights are
 * FP32, convolutions run in BF16 with FP32 accumulation.  The grid
 * dimensions must be multiples of 2^(L-1).  A solve that stalls (64 inner
 * iterations in one refinement pass without reaching its target), breaks
 * down (p.Ap <= 0) or produces a non-finite residual finishes on multigrid
 * and counts one fallback in suli_poisson_stats.fallbacks. */
/* Number of levels L (2..6, default 3): channels min(16*2^l, 64) at level
 * l; the grid dimensions must be multiples of 2^(L-1).  Discards the
 * weights (suli_unet_init or suli_unet_set_weights follows). */
int suli_unet_configure(int levels);
int suli_unet_levels(void);
/* Where the network enters the multigrid hierarchy (default 0).  0: the
 * network is the whole preconditioner, z = M(r_s) on the fine grid.  e >= 1:
 * two-grid form: the V-cycle of SULI_PRECOND_GMG smooths levels 0..e-1 as
 * usual and the network replaces the coarse solve of level e (the grid
 * halved e times): on that level's Jacobi-scaled problem it maps
 * (b~/rms(b~), normalised log of the block-mean density) to a correction,
 * scaled by the energy-optimal step.  Training samples are then that
 * level's problems (right-hand side, density and face conductances),
 * captured in multigrid and learned solves alike.  The grid dimensions must
 * be multiples of 2^(e+L-1).  Changing the level empties the training set. */
int suli_unet_set_entry(int level);
int suli_unet_entry(void);
/* Directions FCG keeps with the learned preconditioner: 4 (default: each new
 * direction A-orthogonalised against the last four, FCG(4)) or 1 (FCG(1),
 * the multigrid iteration; used only with entry levels >= 1, where the
 * network is one part of a V-cycle). */
int suli_unet_set_fcg_m(int m);
int64_t suli_unet_nparams(void);
/* He-uniform initial weights from a seed; clears the Adam moments. */
int suli_unet_init(uint64_t seed);
/* Weights as one FP32 vector: for each layer in the order encoder (two per
 * level 0..L-2), bottom (two), decoder (levels L-2 down to 0), output, the
 * kernel W[27][c_out][c_in] (tap = (dz+1)*9+(dy+1)*3+(dx+1)) then the
 * bias[c_out].  L = 3: c_in->c_out = 2->16, 16->16, 16->32, 32->32, 32->64,
 * 64->32, 32->16, 16->16, 16->1. */
int suli_unet_set_weights(const float *w, int64_t n);
int suli_unet_get_weights(float *w, int64_t n);
/* max_samples > 0: arm the capture of training samples: the following FCG
 * solves append the scaled residual and the density (FP32, on the device)
 * at the start of every refinement pass and, when every > 0, at every
 * every-th inner iteration, until the set holds max_samples.  max_samples
 * 0: stop capturing.  suli_unet_clear empties the set and stops. */
int suli_unet_capture(int max_samples, int every);
int suli_unet_clear(void);
int suli_unet_dataset_size(void);
/* iters Adam steps (lr, beta 0.9/0.999) on samples drawn with the seed from
 * the captured set; loss = 1 - (r.Az)^2/(|r|^2 |Az|^2), the squared residual
 * left by one optimally scaled step along z.  *loss = mean of the last
 * min(iters, 32) steps. */
int suli_unet_train(int iters, double lr, uint64_t seed, double *loss);
/* Loss, weight gradient (n = suli_unet_nparams, may be null) and network
 * output z (may be null) for one sample given on the host in cell order
 * (i fastest, nx*ny*nz values): r is the scaled residual, rho the density. */
int suli_unet_loss_grad(const double *r, const double *rho, double *loss, float *grad, float *z);
/* The held training set, for inspection and tests: its grid (the network's
 * grid at the entry level it was captured at) and sample idx: r (the scaled
 * right-hand side), rho (density) and, for entry >= 1, the level's face
 * conductances (x faces (nx+1)*ny*nz in order i + (nx+1)*(j + ny*k), then y
 * and z faces alike; zero on the walls).  Any pointer may be null. */
int suli_unet_dataset_dims(int *nx, int *ny, int *nz, int *entry);
int suli_unet_dataset_get(int idx, float *r, float *rho, float *faces);
/* Loss, weight gradient and network output of held sample idx, through the
 * operator that training uses for it. */
int suli_unet_sample_loss_grad(int idx, double *loss, float *grad, float *z);

/* Initial guess of the FCG projection (SULI_POISSON_FCG; the reference CG
 * always starts from the current pressure, 9_grad.f90:121).  1 (default):
 * a projection of the first or last stage of a step (the stage set by the
 * last suli_tempo) starts from the current pressure plus its stage's offset
 * extrapolated linearly in time, p + 2 d(n-1) - d(n-2) (p + d(n-1) with one
 * offset known), where d(n) = g(n) - h(n-1) for the first stage and
 * h(n) - g(n) for the last (g, h: the first- and last-stage pressures of
 * step n, same step size); 0: from the current pressure, as the reference.
 * Uploading prd1, a restart file or another time scheme restarts the
 * history.  No counterpart in the reference. */
int suli_set_guess(int mode);

/* Field output (plot_f, 4_plot.f90:260-363): the one FP32 record of an
 * arquivos/campos_NNNNN file, built on the device from the resident fields:
 * x11, y11, z11, uaux, vaux, waux ((nx+1)(ny+1)(nz+1) node values each:
 * coordinates and node-averaged velocities), ls and kaux (nx*ny*nz; kaux = 0,
 * no obstacles) and, for t_plot = 1, prdaux, vorti, vortj, vortk (nx*ny*nz),
 * xnuta ((nx+1)*ny*nz), ynuta (nx*(ny+1)*nz), znuta (nx*ny*(nz+1)), each in
 * Fortran order.  suli_frame_values gives the number of values,
 * suli_frame_record copies them to host memory, suli_write_frame writes
 * the file as the reference does (Fortran unformatted sequential: a 4-byte
 * little-endian byte count before and after the values; records beyond
 * 2^31-1 bytes are refused).  mms_t = 0 output only. */
int suli_frame_values(int t_plot, int64_t *n);
int suli_frame_record(int t_plot, float *buf, int64_t n);
int suli_write_frame(const char *path, int t_plot);
/* The ParaView index of visu (12_paraview_3D.f90:1-179) for the records
 * campos_<first>..campos_<last> of an nx*ny*nz grid: the same grids,
 * attributes, dimensions and byte offsets, as well-formed XML.  Needs no
 * library context. */
int suli_write_xdmf(const char *path, int nx, int ny, int nz, int first, int last, int t_plot, double dt_frame);

/* Checkpoint and restart (restart_salva / restart_ini,
 * 11_diversos.f90:232-329).  suli_checkpoint_save/load: binary, the device
 * state of the run as it is (fields with their ghosts in the run's
 * precision, volumes, time, forcing, time scheme, the projection's guess
 * history, the bottom-obstacle columns, the boundary planes and the
 * manufactured-solution mode with its pressure errors), so that a run resumed from it
 * continues bit for bit; loads into a context of the same grid and
 * precision only, and the file's obstacle (or its absence) replaces the
 * context's.  suli_restart_write/read: the reference's own restart file
 * dados/arquivorestart, list-directed text in restart_salva's order and
 * extents.  The obstacle columns ku, kv, kw are written and read back:
 * reading a file replaces the context's obstacle with the file's columns
 * (a context with an obstacle that reads all-zero columns keeps the
 * obstacle switched on, every column at k = 0, its friction planes at
 * k = 1 as the reference's).  The boundary planes and the momentum sources
 * tf_u, tf_v, tf_w go both ways (sources read only when nonzero).  Items
 * outside this library's model (d_max, d_min, ls_m) are written as zeros
 * and skipped on reading, as are items it rebuilds (prd0, prd, ub, vb, wb,
 * xnut, ynut, znut). */
int suli_checkpoint_save(const char *path);
int suli_checkpoint_load(const char *path);
int suli_restart_write(const char *path);
int suli_restart_read(const char *path);

/* Per-kernel GPU timing: HIP events recorded on the library's stream around
 * every kernel launch while enabled.  suli_prof_enable(1) clears the totals;
 * suli_prof_get returns kernel idx's name, launch count and summed time. */
int suli_prof_enable(int on);
/* Time only the scope of this name (null or "": all), and of it only every
 * every-th launch (events cost a few microseconds per launch), so that a
 * timed run is not slowed; keeps the totals. */
int suli_prof_select(const char *name, int every);
int suli_prof_count(void);
int suli_prof_get(int idx, char *name, int len, int64_t *launches, double *total_ms);

/* wait for all queued GPU work */
int suli_sync(void);

#ifdef __cplusplus
}
#endif
#endif
