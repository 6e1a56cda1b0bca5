/* CPU oracle for the SuLi-IC time step: a plain-C, FP64 restatement of the
 * reference routines, used only as a checker by the tests, by the smoke check
 * and by the CPU-baseline leg of the benchmark.  It is never linked into the
 * product library.
 *
 * Every array is held in the reference's own extents and Fortran order
 * (i fastest), exactly as declared in 0_modules.f90:87-135,177-201,71-73.
 * The state is one flat float64 buffer holding the arrays in the order of
 * OR_ARRAYS below; or_bind() points the struct members into it.
 */
#ifndef SULI_ORACLE_H
#define SULI_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  /* grid and model switches (0_modules.f90) */
  int nx, ny, nz;
  double dx, dy, dz;
  int ccx0, ccxf, ccy0, ccyf, ccz0, cczf;
  int m_turb;          /* 0 none, 1 Smagorinsky, 2 directional Smagorinsky */
  double csmag, gx, gz, chezy;
  /* fluid properties (14_level_set.f90:25-30) */
  double rho_f1, rho_f2, mi_f1, mi_f2, sigma, alpha1;
  /* scalars mutated by the driver */
  double dt, t, vol_ini, vol_ins;
  double dt1;          /* pseudo time step of the reinitialisation, 0.1*max(h) */
  double a_dt[3];
  int tt;
  int t_tempo;         /* time scheme, 0_modules.f90:24: 0 Euler, 1 RK2, 2 RK3, 3 AB2 */
  int it;              /* step counter of 1_principal.f90:64 (AB2 takes Euler at it = 1) */
  int der;             /* advection differences, 0_modules.f90:26: 1 upwind, 2 central, 3 upwind 2nd order */
  int adv_type;        /* advection form, 0_modules.f90:27: 1 classic, 2 rotational, 3 antisymmetric */
  int obst_t;          /* 0: no bottom obstacle; else ku, kv, kw below hold its columns (0_modules.f90:29,76-79) */
  /* diagnostics written by the oracle */
  int cg_iters;        /* iterations of the last graddin call */
  double cg_res2;      /* final sum of squared scaled residuals */
  int reinit_iters;    /* iterations of the last reinitialisation */
  int mass_corr_iters; /* volume-correction passes of the last level_set call */
  int cg_max_it;       /* iteration cap of graddin (0: the reference's 10000, 9_grad.f90:175) */
  /* arrays (pointers into the bound buffer) */
  double *u, *v, *w, *prd1, *rho, *ls_nu, *ls, *mod_ls, *kurv, *hs, *hsx, *hsy, *hsz;
  double *ddlsdx, *ddlsdy, *ddlsdz, *nut, *xnut, *ynut, *znut, *Fu, *Fv, *Fw;
  double *u0, *v0, *w0, *ub, *vb, *wb;
  double *fu0, *fv0, *fw0, *fu1, *fv1, *fw1;  /* kept forces of RK3 / AB2 */
  /* obstacle columns ku(0:nx+2,0:ny+1), kv(0:nx+1,0:ny+2), kw(0:nx+1,0:ny+1)
   * (integer values held as float64): velocities zeroed up to index k = ku */
  double *ku, *kv, *kw;
  /* optional momentum sources (manufactured solutions); NULL means zero */
  double *tf_u, *tf_v, *tf_w;
  /* manufactured-solution mode (0_modules.f90:37): 2 skips the volume
   * correction of level_set (14_level_set.f90:257) */
  int mms_t;
  /* boundary planes of the prescribed-velocity walls (wall code 3), in the
   * declaration order of 0_modules.f90:96-120 (bxx0, bxx1, bxy0, bxz0, bxxf,
   * bxxf1, bxyf, bxzf, byx0, ..., bzzf1), each in Fortran order from index 0
   * with its module extents; NULL when no wall has code 3 */
  double *bp[24];
  /* wave inflow at x = 0 (0_modules.f90:35, module wave_c :164-174): 0 none,
   * 1 Stokes I, 2 Stokes II, 5 Stokes V; amplitude, period, still depth
   * (set by waves_coef, 3_contorno.f90:1086-1088) and the coefficients
   * or_waves_coef computes: l_w, n_w, f_w, c_w, aeta1..5, avel1..5 */
  int wave_t;
  double wave_a, wave_p, wave_h0;
  double wc[14];
} or_state;

/* number of float64 values in the flat state buffer for an nx*ny*nz grid */
int64_t or_state_len(int nx, int ny, int nz);
void or_bind(or_state *s, double *buf);

/* reference routines (argument conventions follow the Fortran ones) */
void or_weno1(double *dphidxp, double *dphidxn, int n, double h, const double *phi0,
              int64_t stride, int ihs);
int  or_der_weno(const or_state *s, const double *ls, double *ta1, double *tb1, double *tc1,
                 double *td1, double *te1, double *tf1);
void or_heaviside(or_state *s);
void or_mod_ls1(or_state *s);
void or_reinic_weno(or_state *s);
void or_level_set(or_state *s);
void or_level_set_ini_tail(or_state *s);
void or_visco(or_state *s);
void or_classico(const or_state *s, double *uint_, double *vint_, double *wint_);
void or_convdiff(or_state *s);
void or_tempo(or_state *s);
void or_contorno(or_state *s, int nlock);
void or_graddin(or_state *s);
void or_posdin(or_state *s);
double or_est_div(const or_state *s);
/* one RK stage (visco..contorno(1)) and one full step, 1_principal.f90:74-95 */
void or_stage(or_state *s, int tt, double dt_base);
void or_stage_noproj(or_state *s, int tt, double dt_base);
/* waves_coef and boundary_waves (3_contorno.f90:1070-1205, :952-1068) */
void or_waves_coef(or_state *s);
void or_boundary_waves(or_state *s);
void or_step(or_state *s, double dt_base);

#ifdef __cplusplus
}
#endif
#endif
