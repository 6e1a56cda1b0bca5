/* CPU oracle for the SuLi-IC time step (test infrastructure only).
 *
 * A plain-C FP64 restatement of the reference Fortran routines, written to
 * evaluate every expression in the reference's own order (Fortran evaluates
 * a+b+c as (a+b)+c and a*b*c as (a*b)*c), so that with IEEE double
 * arithmetic and no FMA contraction it reproduces the reference bit for bit.
 * Each routine cites the reference lines it follows (paths relative to
 * SuLi-IC/).  Only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark use this file; the product library never does.
 */
#include "suli_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.141592653589793115997963468544185161590576171875 /* acos(-1.), 0_modules.f90:4 */

/* index maps of the reference extents (0_modules.f90:92-133) */
#define DIMS const int nx = s->nx, ny = s->ny, nz = s->nz; (void)nx; (void)ny; (void)nz
#define IU(i, j, k) ((int64_t)(i) + (int64_t)(nx + 3) * ((int64_t)(j) + (int64_t)(ny + 2) * (int64_t)(k)))
#define IV(i, j, k) ((int64_t)(i) + (int64_t)(nx + 2) * ((int64_t)(j) + (int64_t)(ny + 3) * (int64_t)(k)))
#define IW(i, j, k) ((int64_t)(i) + (int64_t)(nx + 2) * ((int64_t)(j) + (int64_t)(ny + 2) * (int64_t)(k)))
#define IP(i, j, k) IW(i, j, k)
#define IC(i, j, k) ((int64_t)((i)-1) + (int64_t)nx * ((int64_t)((j)-1) + (int64_t)ny * (int64_t)((k)-1)))
#define IFX(i, j, k) ((int64_t)((i)-1) + (int64_t)(nx + 1) * ((int64_t)((j)-1) + (int64_t)ny * (int64_t)((k)-1)))
#define IFY(i, j, k) ((int64_t)((i)-1) + (int64_t)nx * ((int64_t)((j)-1) + (int64_t)(ny + 1) * (int64_t)((k)-1)))
#define IFZ(i, j, k) IC(i, j, k)
/* generic 1-based (dimx,dimy,dimz) array */
#define IA(i, j, k, dx_, dy_) ((int64_t)((i)-1) + (int64_t)(dx_) * ((int64_t)((j)-1) + (int64_t)(dy_) * (int64_t)((k)-1)))

static double fmax_f(double a, double b) { return a > b ? a : b; }
static double fmin_f(double a, double b) { return a < b ? a : b; }

int64_t or_state_len(int nx, int ny, int nz) {
  int64_t U = (int64_t)(nx + 3) * (ny + 2) * (nz + 2);
  int64_t V = (int64_t)(nx + 2) * (ny + 3) * (nz + 2);
  int64_t W = (int64_t)(nx + 2) * (ny + 2) * (nz + 3);
  int64_t P = (int64_t)(nx + 2) * (ny + 2) * (nz + 2);
  int64_t C = (int64_t)nx * ny * nz;
  int64_t FX = (int64_t)(nx + 1) * ny * nz, FY = (int64_t)nx * (ny + 1) * nz, FZ = (int64_t)nx * ny * (nz + 1);
  /* u v w prd1 | 12 cell arrays | nut | xnut ynut znut | Fu Fv Fw | u0 v0 w0 | ub vb wb |
     fu0 fv0 fw0 | fu1 fv1 fw1 | ku kv kw */
  return U + V + W + P + 12 * C + C + FX + FY + FZ + FX + FY + FZ + FX + FY + FZ + U + V + W +
         2 * (FX + FY + FZ) + (int64_t)(nx + 3) * (ny + 2) + (int64_t)(nx + 2) * (ny + 3) +
         (int64_t)(nx + 2) * (ny + 2);
}

void or_bind(or_state *s, double *b) {
  DIMS;
  int64_t U = (int64_t)(nx + 3) * (ny + 2) * (nz + 2);
  int64_t V = (int64_t)(nx + 2) * (ny + 3) * (nz + 2);
  int64_t W = (int64_t)(nx + 2) * (ny + 2) * (nz + 3);
  int64_t P = (int64_t)(nx + 2) * (ny + 2) * (nz + 2);
  int64_t C = (int64_t)nx * ny * nz;
  int64_t FX = (int64_t)(nx + 1) * ny * nz, FY = (int64_t)nx * (ny + 1) * nz, FZ = (int64_t)nx * ny * (nz + 1);
  s->u = b; b += U; s->v = b; b += V; s->w = b; b += W; s->prd1 = b; b += P;
  s->rho = b; b += C; s->ls_nu = b; b += C; s->ls = b; b += C; s->mod_ls = b; b += C;
  s->kurv = b; b += C; s->hs = b; b += C; s->hsx = b; b += C; s->hsy = b; b += C; s->hsz = b; b += C;
  s->ddlsdx = b; b += C; s->ddlsdy = b; b += C; s->ddlsdz = b; b += C;
  s->nut = b; b += C; s->xnut = b; b += FX; s->ynut = b; b += FY; s->znut = b; b += FZ;
  s->Fu = b; b += FX; s->Fv = b; b += FY; s->Fw = b; b += FZ;
  s->u0 = b; b += FX; s->v0 = b; b += FY; s->w0 = b; b += FZ;
  s->ub = b; b += U; s->vb = b; b += V; s->wb = b; b += W;
  s->fu0 = b; b += FX; s->fv0 = b; b += FY; s->fw0 = b; b += FZ;
  s->fu1 = b; b += FX; s->fv1 = b; b += FY; s->fw1 = b; b += FZ;
  s->ku = b; b += (int64_t)(nx + 3) * (ny + 2);
  s->kv = b; b += (int64_t)(nx + 2) * (ny + 3);
  s->kw = b; b += (int64_t)(nx + 2) * (ny + 2);
}

/* ------------------------------------------------------------------------ */
/* WENO5 one-sided derivatives of a pencil, 14_level_set.f90:511-610.        */
void or_weno1(double *dphidxp, double *dphidxn, int n, double h, const double *phi0,
              int64_t stride, int ihs) {
  const double aux1 = 13. / 12., aux2 = 1. / 4., aux3 = 1. / 6.;
  const double auxx[3] = {0.1, 0.6, 0.3};
  const double aux6 = 0.00000001;
  double *phi1_ = (double *)malloc(sizeof(double) * (size_t)(n + 6));
  double *up_ = (double *)malloc(sizeof(double) * (size_t)(n + 4));
  double *un_ = (double *)malloc(sizeof(double) * (size_t)(n + 4));
#define PHI1(i) phi1_[(i) + 2]   /* phi1(-2:n+3) */
#define UP(i) up_[(i) + 3]       /* up(-3:n)     */
#define UN(i) un_[(i)-1]         /* un(1:n+4)    */
  for (int i = 1; i <= n; i++) PHI1(i) = phi0[(int64_t)(i - 1) * stride];
  if (ihs == 0) { /* periodic, :532-540 */
    PHI1(0) = PHI1(n - 1);
    PHI1(-1) = PHI1(n - 2);
    PHI1(-2) = PHI1(n - 3);
    PHI1(n + 1) = PHI1(2);
    PHI1(n + 2) = PHI1(3);
    PHI1(n + 3) = PHI1(4);
  } else if (ihs == 1) { /* extrapolation, asymmetric 1/5 and 1/11, :541-556 */
    PHI1(0) = 1. / 5. * (12. * PHI1(1) - 9. * PHI1(2) + 2. * PHI1(3));
    PHI1(-1) = 1. / 5. * (12. * PHI1(0) - 9. * PHI1(1) + 2. * PHI1(2));
    PHI1(-2) = 1. / 5. * (12. * PHI1(-1) - 9. * PHI1(0) + 2. * PHI1(1));
    PHI1(n + 1) = 1. / 11. * (18. * PHI1(n) - 9. * PHI1(n - 1) + 2. * PHI1(n - 2));
    PHI1(n + 2) = 1. / 11. * (18. * PHI1(n + 1) - 9. * PHI1(n) + 2. * PHI1(n - 1));
    PHI1(n + 3) = 1. / 11. * (18. * PHI1(n + 2) - 9. * PHI1(n + 1) + 2. * PHI1(n));
  } else if (ihs == 2) { /* zero derivative, 1/11 both sides, :557-569 */
    PHI1(0) = 1. / 11. * (18. * PHI1(1) - 9. * PHI1(2) + 2. * PHI1(3));
    PHI1(-1) = 1. / 11. * (18. * PHI1(0) - 9. * PHI1(1) + 2. * PHI1(2));
    PHI1(-2) = 1. / 11. * (18. * PHI1(-1) - 9. * PHI1(0) + 2. * PHI1(1));
    PHI1(n + 1) = 1. / 11. * (18. * PHI1(n) - 9. * PHI1(n - 1) + 2. * PHI1(n - 2));
    PHI1(n + 2) = 1. / 11. * (18. * PHI1(n + 1) - 9. * PHI1(n) + 2. * PHI1(n - 1));
    PHI1(n + 3) = 1. / 11. * (18. * PHI1(n + 2) - 9. * PHI1(n + 1) + 2. * PHI1(n));
  }
  for (int i = -3; i <= n; i++) UP(i) = (PHI1(i + 3) - PHI1(i + 2)) / h;
  for (int i = 1; i <= n + 4; i++) UN(i) = (PHI1(i - 2) - PHI1(i - 3)) / h;
  for (int i = 1; i <= n; i++) {
    double isup[3], isun[3], alpup[3], alpun[3], omgup[3], omgun[3], a, b;
    a = UP(i) - 2 * UP(i - 1) + UP(i - 2);
    b = UP(i) - 4 * UP(i - 1) + 3 * UP(i - 2);
    isup[0] = aux1 * a * a + aux2 * b * b;
    a = UN(i) - 2 * UN(i + 1) + UN(i + 2);
    b = UN(i) - 4 * UN(i + 1) + 3 * UN(i + 2);
    isun[0] = aux1 * a * a + aux2 * b * b;
    a = UP(i - 1) - 2 * UP(i - 2) + UP(i - 3);
    b = UP(i - 1) - UP(i - 3);
    isup[1] = aux1 * a * a + aux2 * b * b;
    a = UN(i + 1) - 2 * UN(i + 2) + UN(i + 3);
    b = UN(i + 1) - UN(i + 3);
    isun[1] = aux1 * a * a + aux2 * b * b;
    a = UP(i - 2) - 2 * UP(i - 3) + UP(i - 4);
    b = 3 * UP(i - 2) - 4 * UP(i - 3) + UP(i - 4);
    isup[2] = aux1 * a * a + aux2 * b * b;
    a = UN(i + 2) - 2 * UN(i + 3) + UN(i + 4);
    b = 3 * UN(i + 2) - 4 * UN(i + 3) + UN(i + 4);
    isun[2] = aux1 * a * a + aux2 * b * b;
    for (int q = 0; q < 3; q++) {
      alpup[q] = auxx[q] / ((aux6 + isup[q]) * (aux6 + isup[q]));
      alpun[q] = auxx[q] / ((aux6 + isun[q]) * (aux6 + isun[q]));
    }
    for (int q = 0; q < 3; q++) {
      omgup[q] = alpup[q] / (alpup[0] + alpup[1] + alpup[2]);
      omgun[q] = alpun[q] / (alpun[0] + alpun[1] + alpun[2]);
    }
    dphidxp[(int64_t)(i - 1) * stride] =
        aux3 * (omgup[0] * (2 * UP(i) - 7 * UP(i - 1) + 11 * UP(i - 2)) +
                omgup[1] * (-UP(i - 1) + 5 * UP(i - 2) + 2 * UP(i - 3)) +
                omgup[2] * (2 * UP(i - 2) + 5 * UP(i - 3) - UP(i - 4)));
    dphidxn[(int64_t)(i - 1) * stride] =
        aux3 * (omgun[0] * (2 * UN(i) - 7 * UN(i + 1) + 11 * UN(i + 2)) +
                omgun[1] * (-UN(i + 1) + 5 * UN(i + 2) + 2 * UN(i + 3)) +
                omgun[2] * (2 * UN(i + 2) + 5 * UN(i + 3) - UN(i + 4)));
  }
#undef PHI1
#undef UP
#undef UN
  free(phi1_);
  free(up_);
  free(un_);
}

/* der_weno, 14_level_set.f90:343-384: x and y pencils use ihs 0 (periodic)
 * or 2, z pencils ihs 1; returns the ihs left in the caller (always 1). */
int or_der_weno(const or_state *s, const double *ls, double *ta1, double *tb1, double *tc1,
                double *td1, double *te1, double *tf1) {
  DIMS;
  int ihs = (s->ccx0 == 0) ? 0 : 2;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      or_weno1(ta1 + IC(1, j, k), td1 + IC(1, j, k), nx, s->dx, ls + IC(1, j, k), 1, ihs);
  ihs = (s->ccy0 == 0) ? 0 : 2;
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++)
      or_weno1(tb1 + IC(i, 1, k), te1 + IC(i, 1, k), ny, s->dy, ls + IC(i, 1, k), nx, ihs);
  ihs = 1;
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++)
      or_weno1(tc1 + IC(i, j, 1), tf1 + IC(i, j, 1), nz, s->dz, ls + IC(i, j, 1), (int64_t)nx * ny, ihs);
  return ihs;
}

static double *cell_alloc(const or_state *s) {
  return (double *)calloc((size_t)s->nx * s->ny * s->nz, sizeof(double));
}

/* SSP-RK3 stage update intt_ls, 14_level_set.f90:287-306, with the
 * coefficients of 14_level_set.f90:34-44. */
static const double ADTL[3] = {1., 3. / 4., 1. / 3.};
static const double BDTL[3] = {0., 1. / 4., 2. / 3.};
static const double GDTL[3] = {1., 1. / 4., 2. / 3.};
static void intt_ls(const or_state *s, const double *hx, double *gx, double *ta1, int itrl, double *ls1,
                    double dt1) {
  int64_t n = (int64_t)s->nx * s->ny * s->nz;
  double a = ADTL[itrl - 1], b = BDTL[itrl - 1], g = GDTL[itrl - 1];
  if (a == 1.) memcpy(ta1, ls1, sizeof(double) * (size_t)n);
  memcpy(gx, ls1, sizeof(double) * (size_t)n);
  for (int64_t q = 0; q < n; q++) ls1[q] = a * ta1[q] + b * gx[q] + g * dt1 * hx[q];
}

/* conv_weno, 14_level_set.f90:308-341 */
static void conv_weno(const or_state *s, double *sy7) {
  DIMS;
  double *ta1 = cell_alloc(s), *tb1 = cell_alloc(s), *tc1 = cell_alloc(s);
  double *td1 = cell_alloc(s), *te1 = cell_alloc(s), *tf1 = cell_alloc(s);
  or_der_weno(s, s->ls, ta1, tb1, tc1, td1, te1, tf1);
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double apos = fmax_f((s->u[IU(i, j, k)] + s->u[IU(i + 1, j, k)]) * 0.5, 0.);
        double aneg = fmin_f((s->u[IU(i, j, k)] + s->u[IU(i + 1, j, k)]) * 0.5, 0.);
        double bpos = fmax_f((s->v[IV(i, j, k)] + s->v[IV(i, j + 1, k)]) * 0.5, 0.);
        double bneg = fmin_f((s->v[IV(i, j, k)] + s->v[IV(i, j + 1, k)]) * 0.5, 0.);
        double cpos = fmax_f((s->w[IW(i, j, k)] + s->w[IW(i, j, k + 1)]) * 0.5, 0.);
        double cneg = fmin_f((s->w[IW(i, j, k)] + s->w[IW(i, j, k + 1)]) * 0.5, 0.);
        int64_t c = IC(i, j, k);
        sy7[c] = -(apos * td1[c] + aneg * ta1[c] + bpos * te1[c] + bneg * tb1[c] + cpos * tf1[c] +
                   cneg * tc1[c]);
      }
  free(ta1); free(tb1); free(tc1); free(td1); free(te1); free(tf1);
}

/* Godunov upwinding with the 1e-8 floors, 14_level_set.f90:406-430 */
static void godunov(double lsv, double *ta, double *tb, double *tc, double td, double te, double tf) {
  double aux1, aux2;
  if (lsv > 0.) {
    aux1 = fmax_f(td, 0.00000001); aux2 = -fmin_f(*ta, 0.00000001); *ta = fmax_f(aux1, aux2);
    aux1 = fmax_f(te, 0.00000001); aux2 = -fmin_f(*tb, 0.00000001); *tb = fmax_f(aux1, aux2);
    aux1 = fmax_f(tf, 0.00000001); aux2 = -fmin_f(*tc, 0.00000001); *tc = fmax_f(aux1, aux2);
  } else {
    aux1 = fmax_f(*ta, 0.00000001); aux2 = -fmin_f(td, 0.00000001); *ta = fmax_f(aux1, aux2);
    aux1 = fmax_f(*tb, 0.00000001); aux2 = -fmin_f(te, 0.00000001); *tb = fmax_f(aux1, aux2);
    aux1 = fmax_f(*tc, 0.00000001); aux2 = -fmin_f(tf, 0.00000001); *tc = fmax_f(aux1, aux2);
  }
}

/* reinic_weno, 14_level_set.f90:386-509 */
void or_reinic_weno(or_state *s) {
  DIMS;
  const int64_t n = (int64_t)nx * ny * nz;
  const double dx1 = fmax_f(fmax_f(s->dx, s->dy), s->dz);
  double *ta1 = cell_alloc(s), *tb1 = cell_alloc(s), *tc1 = cell_alloc(s);
  double *td1 = cell_alloc(s), *te1 = cell_alloc(s), *tf1 = cell_alloc(s);
  double *func_s = cell_alloc(s), *sy7 = cell_alloc(s), *gx = cell_alloc(s), *ta1s = cell_alloc(s);
  double *lsaux = cell_alloc(s);
  const int l = 3;
  or_der_weno(s, s->ls, ta1, tb1, tc1, td1, te1, tf1);
  for (int64_t c = 0; c < n; c++) {
    godunov(s->ls[c], &ta1[c], &tb1[c], &tc1[c], td1[c], te1[c], tf1[c]);
    double m = sqrt(ta1[c] * ta1[c] + tb1[c] * tb1[c] + tc1[c] * tc1[c]);
    func_s[c] = s->ls[c] / sqrt(s->ls[c] * s->ls[c] + m * m * dx1 * dx1);
  }
  int il = 0;
  double error = 999.;
  while (error > s->dt1 * dx1 * dx1) {
    il = il + 1;
    for (int itrl = 1; itrl <= 3; itrl++) {
      or_der_weno(s, s->ls, ta1, tb1, tc1, td1, te1, tf1);
      for (int64_t c = 0; c < n; c++) {
        godunov(s->ls[c], &ta1[c], &tb1[c], &tc1[c], td1[c], te1[c], tf1[c]);
        double m = sqrt(ta1[c] * ta1[c] + tb1[c] * tb1[c] + tc1[c] * tc1[c]);
        sy7[c] = func_s[c] * (1. - m);
        lsaux[c] = s->ls[c];
      }
      intt_ls(s, sy7, gx, ta1s, itrl, s->ls, s->dt1);
    }
    if (il == l) {
      error = 0.;
    } else if (il < 1) {
      error = 999.;
    } else {
      error = 0.;
      int nr = 0;
      for (int64_t c = 0; c < n; c++)
        if (fabs(s->ls[c]) <= 2. * s->alpha1 * dx1) {
          error = error + s->ls[c] - lsaux[c];
          nr = nr + 1;
        }
      error = error / nr;
    }
  }
  s->reinit_iters = il;
  free(ta1); free(tb1); free(tc1); free(td1); free(te1); free(tf1);
  free(func_s); free(sy7); free(gx); free(ta1s); free(lsaux);
}

/* heaviside, 14_level_set.f90:612-707 (t_hs = 0) */
void or_heaviside(or_state *s) {
  DIMS;
  const int64_t n = (int64_t)nx * ny * nz;
  const double dx1 = fmax_f(fmax_f(s->dx, s->dy), s->dz);
  double *ta1 = cell_alloc(s), *tb1 = cell_alloc(s), *tc1 = cell_alloc(s);
  double *td1 = cell_alloc(s), *te1 = cell_alloc(s), *tf1 = cell_alloc(s);
  or_der_weno(s, s->ls, ta1, tb1, tc1, td1, te1, tf1);
  for (int64_t c = 0; c < n; c++) {
    if (!(fabs(ta1[c]) > fabs(td1[c]))) ta1[c] = td1[c];
    if (!(fabs(tb1[c]) > fabs(te1[c]))) tb1[c] = te1[c];
    if (!(fabs(tc1[c]) > fabs(tf1[c]))) tc1[c] = tf1[c];
  }
  const double a = s->alpha1;
  for (int64_t c = 0; c < n; c++) {
    double lsv = s->ls[c];
    if (lsv < -a * dx1) {
      s->hs[c] = 0.; s->hsx[c] = 0.; s->hsy[c] = 0.; s->hsz[c] = 0.;
    } else if (lsv > a * dx1) {
      s->hs[c] = 1.; s->hsx[c] = 0.; s->hsy[c] = 0.; s->hsz[c] = 0.;
    } else {
      s->hs[c] = 0.5 * (1. + lsv / (a * dx1) + 1. / PI * sin(PI * lsv / (a * dx1)));
      s->hsx[c] = 0.5 * ta1[c] * (1. + cos(PI * lsv / (a * dx1))) / (a * dx1);
      s->hsy[c] = 0.5 * tb1[c] * (1. + cos(PI * lsv / (a * dx1))) / (a * dx1);
      s->hsz[c] = 0.5 * tc1[c] * (1. + cos(PI * lsv / (a * dx1))) / (a * dx1);
    }
    s->rho[c] = s->rho_f1 * (1. - s->hs[c]) + s->rho_f2 * s->hs[c];
    s->ls_nu[c] = s->mi_f1 * (1. - s->hs[c]) + s->mi_f2 * s->hs[c];
  }
  free(ta1); free(tb1); free(tc1); free(td1); free(te1); free(tf1);
}

/* mod_ls1, 14_level_set.f90:709-805: normals from the larger one-sided
 * derivative, then WENO (ihs = 1 in every direction) of the normals and the
 * curvature as the sum of the larger one-sided derivatives. */
void or_mod_ls1(or_state *s) {
  DIMS;
  const int64_t n = (int64_t)nx * ny * nz;
  double *ta1 = cell_alloc(s), *tb1 = cell_alloc(s), *tc1 = cell_alloc(s);
  double *td1 = cell_alloc(s), *te1 = cell_alloc(s), *tf1 = cell_alloc(s);
  double *nxa = cell_alloc(s), *nya = cell_alloc(s), *nza = cell_alloc(s);
  int ihs = or_der_weno(s, s->ls, ta1, tb1, tc1, td1, te1, tf1);
  for (int64_t c = 0; c < n; c++) {
    if (!(fabs(ta1[c]) > fabs(td1[c]))) ta1[c] = td1[c];
    if (!(fabs(tb1[c]) > fabs(te1[c]))) tb1[c] = te1[c];
    if (!(fabs(tc1[c]) > fabs(tf1[c]))) tc1[c] = tf1[c];
  }
  for (int64_t c = 0; c < n; c++) {
    s->mod_ls[c] = sqrt(ta1[c] * ta1[c] + tb1[c] * tb1[c] + tc1[c] * tc1[c]);
    if (s->mod_ls[c] == 0) {
      nxa[c] = 0.; nya[c] = 0.; nza[c] = 0.;
    } else {
      nxa[c] = ta1[c] / s->mod_ls[c];
      nya[c] = tb1[c] / s->mod_ls[c];
      nza[c] = tc1[c] / s->mod_ls[c];
    }
  }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      or_weno1(ta1 + IC(1, j, k), td1 + IC(1, j, k), nx, s->dx, nxa + IC(1, j, k), 1, ihs);
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++)
      or_weno1(tb1 + IC(i, 1, k), te1 + IC(i, 1, k), ny, s->dy, nya + IC(i, 1, k), nx, ihs);
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++)
      or_weno1(tc1 + IC(i, j, 1), tf1 + IC(i, j, 1), nz, s->dz, nza + IC(i, j, 1), (int64_t)nx * ny, ihs);
  for (int64_t c = 0; c < n; c++) {
    s->ddlsdx[c] = (fabs(td1[c]) > fabs(ta1[c])) ? td1[c] : ta1[c];
    s->ddlsdy[c] = (fabs(te1[c]) > fabs(tb1[c])) ? te1[c] : tb1[c];
    s->ddlsdz[c] = (fabs(tf1[c]) > fabs(tc1[c])) ? tf1[c] : tc1[c];
    s->kurv[c] = s->ddlsdx[c] + s->ddlsdy[c] + s->ddlsdz[c];
  }
  free(ta1); free(tb1); free(tc1); free(td1); free(te1); free(tf1);
  free(nxa); free(nya); free(nza);
}

static double volume(const or_state *s) {
  DIMS;
  double vol = 0.;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) vol = vol + (1. - (1. - s->hs[IC(i, j, k)])) * s->dx * s->dy * s->dz;
  return vol;
}

/* Tail of level_set_ini (14_level_set.f90:205-217): heaviside, initial
 * volume, mod_ls1. */
void or_level_set_ini_tail(or_state *s) {
  or_heaviside(s);
  s->vol_ini = volume(s);
  or_mod_ls1(s);
}

/* level_set, 14_level_set.f90:221-285 */
void or_level_set(or_state *s) {
  DIMS;
  const int64_t n = (int64_t)nx * ny * nz;
  double *sy7 = cell_alloc(s), *gx = cell_alloc(s), *ta1 = cell_alloc(s);
  for (int itrl = 1; itrl <= 3; itrl++) {
    conv_weno(s, sy7);
    intt_ls(s, sy7, gx, ta1, itrl, s->ls, s->dt);  /* dt1 := dt, :231-232 */
  }
  or_reinic_weno(s);
  or_mod_ls1(s);
  or_heaviside(s);
  s->vol_ins = volume(s);
  int nc = 0;
  while (fabs(s->vol_ins - s->vol_ini) / s->vol_ins > 0.1 && s->mms_t != 2) {
    for (int64_t c = 0; c < n; c++)
      s->ls[c] = s->ls[c] - s->dt1 * (s->vol_ins - s->vol_ini) * s->mod_ls[c] / s->vol_ini;
    or_mod_ls1(s);
    or_heaviside(s);
    s->vol_ins = volume(s);
    nc++;
  }
  s->mass_corr_iters = nc;
  or_mod_ls1(s);
  or_heaviside(s);
  free(sy7); free(gx); free(ta1);
}

/* ------------------------------------------------------------------------ */
/* finite-difference primitives, 8_derivadas.f90:1-68,304-455 */

/* centre -> face along x with end faces copied, (dimx,dimy,dimz)->(dimx+1,...) */
static void interpx_cf(const double *a, int dx_, int dy_, int dz_, double *o) {
  for (int k = 1; k <= dz_; k++)
    for (int j = 1; j <= dy_; j++) {
      for (int i = 2; i <= dx_; i++)
        o[IA(i, j, k, dx_ + 1, dy_)] = (a[IA(i, j, k, dx_, dy_)] + a[IA(i - 1, j, k, dx_, dy_)]) * 0.5;
      o[IA(1, j, k, dx_ + 1, dy_)] = o[IA(2, j, k, dx_ + 1, dy_)];
      o[IA(dx_ + 1, j, k, dx_ + 1, dy_)] = o[IA(dx_, j, k, dx_ + 1, dy_)];
    }
}
static void interpy_cf(const double *a, int dx_, int dy_, int dz_, double *o) {
  for (int k = 1; k <= dz_; k++)
    for (int j = 2; j <= dy_; j++)
      for (int i = 1; i <= dx_; i++)
        o[IA(i, j, k, dx_, dy_ + 1)] = (a[IA(i, j, k, dx_, dy_)] + a[IA(i, j - 1, k, dx_, dy_)]) * 0.5;
  for (int k = 1; k <= dz_; k++)
    for (int i = 1; i <= dx_; i++) {
      o[IA(i, 1, k, dx_, dy_ + 1)] = o[IA(i, 2, k, dx_, dy_ + 1)];
      o[IA(i, dy_ + 1, k, dx_, dy_ + 1)] = o[IA(i, dy_, k, dx_, dy_ + 1)];
    }
}
static void interpz_cf(const double *a, int dx_, int dy_, int dz_, double *o) {
  for (int k = 2; k <= dz_; k++)
    for (int j = 1; j <= dy_; j++)
      for (int i = 1; i <= dx_; i++)
        o[IA(i, j, k, dx_, dy_)] = (a[IA(i, j, k, dx_, dy_)] + a[IA(i, j, k - 1, dx_, dy_)]) * 0.5;
  for (int j = 1; j <= dy_; j++)
    for (int i = 1; i <= dx_; i++) {
      o[IA(i, j, 1, dx_, dy_)] = o[IA(i, j, 2, dx_, dy_)];
      o[IA(i, j, dz_ + 1, dx_, dy_)] = o[IA(i, j, dz_, dx_, dy_)];
    }
}
/* face -> centre, (dimx,...) -> (dimx-1,...) etc. */
static void interpx_fc(const double *a, int dx_, int dy_, int dz_, double *o) {
  for (int k = 1; k <= dz_; k++)
    for (int j = 1; j <= dy_; j++)
      for (int i = 1; i <= dx_ - 1; i++)
        o[IA(i, j, k, dx_ - 1, dy_)] = (a[IA(i, j, k, dx_, dy_)] + a[IA(i + 1, j, k, dx_, dy_)]) * 0.5;
}
static void interpy_fc(const double *a, int dx_, int dy_, int dz_, double *o) {
  for (int k = 1; k <= dz_; k++)
    for (int j = 1; j <= dy_ - 1; j++)
      for (int i = 1; i <= dx_; i++)
        o[IA(i, j, k, dx_, dy_ - 1)] = (a[IA(i, j, k, dx_, dy_)] + a[IA(i, j + 1, k, dx_, dy_)]) * 0.5;
}
static void interpz_fc(const double *a, int dx_, int dy_, int dz_, double *o) {
  for (int k = 1; k <= dz_ - 1; k++)
    for (int j = 1; j <= dy_; j++)
      for (int i = 1; i <= dx_; i++)
        o[IA(i, j, k, dx_, dy_)] = (a[IA(i, j, k, dx_, dy_)] + a[IA(i, j, k + 1, dx_, dy_)]) * 0.5;
}
/* central derivatives of a ghosted (0:dimx+1,0:dimy+1,0:dimz+1) array */
#define IG(i, j, k, dx_, dy_) ((int64_t)(i) + (int64_t)((dx_) + 2) * ((int64_t)(j) + (int64_t)((dy_) + 2) * (int64_t)(k)))
static void derivax(const double *a, int dx_, int dy_, int dz_, double h, double *o) {
  for (int k = 1; k <= dz_; k++)
    for (int j = 1; j <= dy_; j++)
      for (int i = 1; i <= dx_; i++)
        o[IA(i, j, k, dx_, dy_)] = (a[IG(i + 1, j, k, dx_, dy_)] - a[IG(i - 1, j, k, dx_, dy_)]) / (2. * h);
}
static void derivay(const double *a, int dx_, int dy_, int dz_, double h, double *o) {
  for (int k = 1; k <= dz_; k++)
    for (int j = 1; j <= dy_; j++)
      for (int i = 1; i <= dx_; i++)
        o[IA(i, j, k, dx_, dy_)] = (a[IG(i, j + 1, k, dx_, dy_)] - a[IG(i, j - 1, k, dx_, dy_)]) / (2. * h);
}
static void derivaz(const double *a, int dx_, int dy_, int dz_, double h, double *o) {
  for (int k = 1; k <= dz_; k++)
    for (int j = 1; j <= dy_; j++)
      for (int i = 1; i <= dx_; i++)
        o[IA(i, j, k, dx_, dy_)] = (a[IG(i, j, k + 1, dx_, dy_)] - a[IG(i, j, k - 1, dx_, dy_)]) / (2. * h);
}

static double *alloc_n(int64_t n) { return (double *)calloc((size_t)n, sizeof(double)); }

/* visco, 5_viscosidade.f90:12-219 (m_turb 0 or 1) */
void or_visco(or_state *s) {
  DIMS;
  const int nx1 = nx + 1, ny1 = ny + 1, nz1 = nz + 1;
  const int64_t C = (int64_t)nx * ny * nz, FX = (int64_t)nx1 * ny * nz, FY = (int64_t)nx * ny1 * nz,
                FZ = (int64_t)nx * ny * nz1;
  double *xnut_a = alloc_n(FX), *ynut_a = alloc_n(FY), *znut_a = alloc_n(FZ);
  if (s->m_turb == 1) {
    double *dudx_i = alloc_n(FX), *dudy_i = alloc_n(FX), *dudz_i = alloc_n(FX);
    double *dvdx_i = alloc_n(FY), *dvdy_i = alloc_n(FY), *dvdz_i = alloc_n(FY);
    double *dwdx_i = alloc_n(FZ), *dwdy_i = alloc_n(FZ), *dwdz_i = alloc_n(FZ);
    derivax(s->u, nx1, ny, nz, s->dx, dudx_i); derivax(s->v, nx, ny1, nz, s->dx, dvdx_i); derivax(s->w, nx, ny, nz1, s->dx, dwdx_i);
    derivay(s->u, nx1, ny, nz, s->dy, dudy_i); derivay(s->v, nx, ny1, nz, s->dy, dvdy_i); derivay(s->w, nx, ny, nz1, s->dy, dwdy_i);
    derivaz(s->u, nx1, ny, nz, s->dz, dudz_i); derivaz(s->v, nx, ny1, nz, s->dz, dvdz_i); derivaz(s->w, nx, ny, nz1, s->dz, dwdz_i);
    double *dudx = alloc_n(C), *dudy = alloc_n(C), *dudz = alloc_n(C);
    double *dvdx = alloc_n(C), *dvdy = alloc_n(C), *dvdz = alloc_n(C);
    double *dwdx = alloc_n(C), *dwdy = alloc_n(C), *dwdz = alloc_n(C);
    interpx_fc(dudx_i, nx1, ny, nz, dudx); interpx_fc(dudy_i, nx1, ny, nz, dudy); interpx_fc(dudz_i, nx1, ny, nz, dudz);
    interpy_fc(dvdx_i, nx, ny1, nz, dvdx); interpy_fc(dvdy_i, nx, ny1, nz, dvdy); interpy_fc(dvdz_i, nx, ny1, nz, dvdz);
    interpz_fc(dwdx_i, nx, ny, nz1, dwdx); interpz_fc(dwdy_i, nx, ny, nz1, dwdy); interpz_fc(dwdz_i, nx, ny, nz1, dwdz);
    const double c23 = pow(s->dx * s->dy * s->dz, 2. / 3.);
    for (int64_t c = 0; c < C; c++) {
      double p1 = dudx[c] + 0.5 * (dudy[c] + dvdx[c] + dudz[c] + dwdx[c]);
      double p2 = dvdy[c] + 0.5 * (dvdx[c] + dudy[c] + dvdz[c] + dwdy[c]);
      double p3 = dwdz[c] + 0.5 * (dwdy[c] + dvdz[c] + dudz[c] + dwdx[c]);
      s->nut[c] = c23 * s->csmag * s->csmag * sqrt(2. * (p1 + p2 + p3) * (p1 + p2 + p3)) * s->rho[c];
    }
    interpx_cf(s->nut, nx, ny, nz, xnut_a);
    interpy_cf(s->nut, nx, ny, nz, ynut_a);
    interpz_cf(s->nut, nx, ny, nz, znut_a);
    free(dudx_i); free(dudy_i); free(dudz_i); free(dvdx_i); free(dvdy_i); free(dvdz_i);
    free(dwdx_i); free(dwdy_i); free(dwdz_i);
    free(dudx); free(dudy); free(dudz); free(dvdx); free(dvdy); free(dvdz); free(dwdx); free(dwdy); free(dwdz);
  } else if (s->m_turb == 2) {
    /* directional Smagorinsky, 5_viscosidade.f90:104-170: face viscosities
     * from the strain at each face; nut is left as it is */
    double *dudx_i = alloc_n(FX), *dudy_i = alloc_n(FX), *dudz_i = alloc_n(FX);
    double *dvdx_i = alloc_n(FY), *dvdy_i = alloc_n(FY), *dvdz_i = alloc_n(FY);
    double *dwdx_i = alloc_n(FZ), *dwdy_i = alloc_n(FZ), *dwdz_i = alloc_n(FZ);
    derivax(s->u, nx1, ny, nz, s->dx, dudx_i); derivax(s->v, nx, ny1, nz, s->dx, dvdx_i); derivax(s->w, nx, ny, nz1, s->dx, dwdx_i);
    derivay(s->u, nx1, ny, nz, s->dy, dudy_i); derivay(s->v, nx, ny1, nz, s->dy, dvdy_i); derivay(s->w, nx, ny, nz1, s->dy, dwdy_i);
    derivaz(s->u, nx1, ny, nz, s->dz, dudz_i); derivaz(s->v, nx, ny1, nz, s->dz, dvdz_i); derivaz(s->w, nx, ny, nz1, s->dz, dwdz_i);
    const int64_t EXY = (int64_t)nx1 * ny1 * nz, EXZ = (int64_t)nx1 * ny * nz1, EYZ = (int64_t)nx * ny1 * nz1;
    double *ea = alloc_n(EXY), *eb = alloc_n(EXZ), *ec = alloc_n(EYZ);
    double *dvdx_x = alloc_n(FX), *dwdx_x = alloc_n(FX), *dudy_y = alloc_n(FY), *dwdy_y = alloc_n(FY);
    double *dudz_z = alloc_n(FZ), *dvdz_z = alloc_n(FZ);
    interpx_cf(dvdx_i, nx, ny1, nz, ea); interpy_fc(ea, nx1, ny1, nz, dvdx_x);
    interpx_cf(dwdx_i, nx, ny, nz1, eb); interpz_fc(eb, nx1, ny, nz1, dwdx_x);
    interpy_cf(dudy_i, nx1, ny, nz, ea); interpx_fc(ea, nx1, ny1, nz, dudy_y);
    interpy_cf(dwdy_i, nx, ny, nz1, ec); interpz_fc(ec, nx, ny1, nz1, dwdy_y);
    interpz_cf(dudz_i, nx1, ny, nz, eb); interpx_fc(eb, nx1, ny, nz1, dudz_z);
    interpz_cf(dvdz_i, nx, ny1, nz, ec); interpy_fc(ec, nx, ny1, nz1, dvdz_z);
    const double cs = s->csmag;
    for (int64_t c = 0; c < FX; c++) {
      const double p1 = dudx_i[c] + 0.5 * (dudy_i[c] + dvdx_x[c] + dudz_i[c] + dwdx_x[c]);
      xnut_a[c] = s->dx * s->dx * cs * cs * sqrt(2. * p1 * p1);
    }
    for (int64_t c = 0; c < FY; c++) {
      const double p1 = dvdy_i[c] + 0.5 * (dvdx_i[c] + dudy_y[c] + dvdz_i[c] + dwdy_y[c]);
      ynut_a[c] = s->dy * s->dy * cs * cs * sqrt(2. * p1 * p1);
    }
    for (int64_t c = 0; c < FZ; c++) {
      const double p1 = dwdz_i[c] + 0.5 * (dwdy_i[c] + dvdz_z[c] + dudz_z[c] + dwdx_i[c]);
      znut_a[c] = s->dz * s->dz * cs * cs * sqrt(2. * p1 * p1);
    }
    free(dudx_i); free(dudy_i); free(dudz_i); free(dvdx_i); free(dvdy_i); free(dvdz_i);
    free(dwdx_i); free(dwdy_i); free(dwdz_i); free(ea); free(eb); free(ec);
    free(dvdx_x); free(dwdx_x); free(dudy_y); free(dwdy_y); free(dudz_z); free(dvdz_z);
  }
  double *lx = alloc_n(FX), *ly = alloc_n(FY), *lz = alloc_n(FZ);
  interpx_cf(s->ls_nu, nx, ny, nz, lx);
  interpy_cf(s->ls_nu, nx, ny, nz, ly);
  interpz_cf(s->ls_nu, nx, ny, nz, lz);
  for (int64_t c = 0; c < FX; c++) s->xnut[c] = lx[c] + xnut_a[c];
  for (int64_t c = 0; c < FY; c++) s->ynut[c] = ly[c] + ynut_a[c];
  for (int64_t c = 0; c < FZ; c++) s->znut[c] = lz[c] + znut_a[c];
  free(lx); free(ly); free(lz); free(xnut_a); free(ynut_a); free(znut_a);
}

/* classico, der = 1 (first-order upwind), 7_advectivo.f90:33-113 */
/* second-order one-sided derivatives of a ghosted array, 8_derivadas.f90:83-290:
 * u2n backward (central at the first face), u2p forward (central at the last) */
static double d_u2n(const double *a, int64_t c, int64_t st, int f, double h) {
  if (f >= 2) return 0.5 * (a[c - 2 * st] - 4 * a[c - st] + 3 * a[c]) / h;
  return 0.5 * (a[c + st] - a[c - st]) / h;
}
static double d_u2p(const double *a, int64_t c, int64_t st, int f, int n, double h) {
  if (f <= n - 1) return 0.5 * (-a[c + 2 * st] + 4 * a[c + st] - 3 * a[c]) / h;
  return 0.5 * (a[c + st] - a[c - st]) / h;
}

/* transverse advecting velocities of der = 2, 3 (7_advectivo.f90:129-145,
 * 176-192): face -> centre -> face interpolation, end faces copied */
static void adv_interp(const or_state *s, double *bma, double *dma, double *amb, double *dmb, double *amd,
                       double *bmd) {
  DIMS;
  const int nx1 = nx + 1, ny1 = ny + 1, nz1 = nz + 1;
  double *aux = alloc_n((int64_t)nx * ny * nz);
  double *us = alloc_n((int64_t)nx1 * ny * nz), *vs = alloc_n((int64_t)nx * ny1 * nz),
         *ws = alloc_n((int64_t)nx * ny * nz1);
  for (int k = 1; k <= nz; k++) /* the sections u(1:nx1,1:ny,1:nz), v(1:nx,1:ny1,1:nz), w(1:nx,1:ny,1:nz1) */
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx1; i++) us[IFX(i, j, k)] = s->u[IU(i, j, k)];
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny1; j++)
      for (int i = 1; i <= nx; i++) vs[IFY(i, j, k)] = s->v[IV(i, j, k)];
  for (int k = 1; k <= nz1; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) ws[IFZ(i, j, k)] = s->w[IW(i, j, k)];
  interpy_fc(vs, nx, ny1, nz, aux); interpx_cf(aux, nx, ny, nz, bma);
  interpz_fc(ws, nx, ny, nz1, aux); interpx_cf(aux, nx, ny, nz, dma);
  interpx_fc(us, nx1, ny, nz, aux); interpy_cf(aux, nx, ny, nz, amb);
  interpz_fc(ws, nx, ny, nz1, aux); interpy_cf(aux, nx, ny, nz, dmb);
  interpx_fc(us, nx1, ny, nz, aux); interpz_cf(aux, nx, ny, nz, amd);
  interpy_fc(vs, nx, ny1, nz, aux); interpz_cf(aux, nx, ny, nz, bmd);
  free(aux); free(us); free(vs); free(ws);
}

/* classico der = 2 (central, 7_advectivo.f90:116-173) and der = 3 (upwind
 * second order, :176-291) */
static void classico_der23(const or_state *s, double *uint_, double *vint_, double *wint_) {
  DIMS;
  const double *u = s->u, *v = s->v, *w = s->w;
  const double dx = s->dx, dy = s->dy, dz = s->dz;
  const int nx1 = nx + 1, ny1 = ny + 1, nz1 = nz + 1;
  const int64_t FX = (int64_t)nx1 * ny * nz, FY = (int64_t)nx * ny1 * nz, FZ = (int64_t)nx * ny * nz1;
  double *bma = alloc_n(FX), *dma = alloc_n(FX), *amb = alloc_n(FY), *dmb = alloc_n(FY), *amd = alloc_n(FZ),
         *bmd = alloc_n(FZ);
  adv_interp(s, bma, dma, amb, dmb, amd, bmd);
  const int64_t uX = 1, uY = nx + 3, uZ = (int64_t)(nx + 3) * (ny + 2);
  const int64_t vX = 1, vY = nx + 2, vZ = (int64_t)(nx + 2) * (ny + 3);
  const int64_t wX = 1, wY = nx + 2, wZ = (int64_t)(nx + 2) * (ny + 2);
  if (s->der == 2) {
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx1; i++) {
          const int64_t c = IU(i, j, k);
          const double dudx = (u[c + uX] - u[c - uX]) / (2. * dx), dudy = (u[c + uY] - u[c - uY]) / (2. * dy),
                       dudz = (u[c + uZ] - u[c - uZ]) / (2. * dz);
          uint_[IFX(i, j, k)] = u[c] * dudx + bma[IFX(i, j, k)] * dudy + dma[IFX(i, j, k)] * dudz;
        }
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny1; j++)
        for (int i = 1; i <= nx; i++) {
          const int64_t c = IV(i, j, k);
          const double dvdx = (v[c + vX] - v[c - vX]) / (2. * dx), dvdy = (v[c + vY] - v[c - vY]) / (2. * dy),
                       dvdz = (v[c + vZ] - v[c - vZ]) / (2. * dz);
          vint_[IFY(i, j, k)] = amb[IFY(i, j, k)] * dvdx + v[c] * dvdy + dmb[IFY(i, j, k)] * dvdz;
        }
    for (int k = 1; k <= nz1; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          const int64_t c = IW(i, j, k);
          const double dwdx = (w[c + wX] - w[c - wX]) / (2. * dx), dwdy = (w[c + wY] - w[c - wY]) / (2. * dy),
                       dwdz = (w[c + wZ] - w[c - wZ]) / (2. * dz);
          wint_[IFZ(i, j, k)] = amd[IFZ(i, j, k)] * dwdx + bmd[IFZ(i, j, k)] * dwdy + w[c] * dwdz;
        }
  } else {
#define UPW2(q, c, st, f, n, h, a) (fmax_f(a, 0.) * d_u2n(q, c, st, f, h) + fmin_f(a, 0.) * d_u2p(q, c, st, f, n, h))
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx1; i++) {
          const int64_t c = IU(i, j, k);
          const double dudx = UPW2(u, c, uX, i, nx1, dx, u[c]);
          const double dudy = UPW2(u, c, uY, j, ny, dy, bma[IFX(i, j, k)]);
          const double dudz = UPW2(u, c, uZ, k, nz, dz, dma[IFX(i, j, k)]);
          uint_[IFX(i, j, k)] = dudx + dudy + dudz;
        }
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny1; j++)
        for (int i = 1; i <= nx; i++) {
          const int64_t c = IV(i, j, k);
          const double dvdx = UPW2(v, c, vX, i, nx, dx, amb[IFY(i, j, k)]);
          const double dvdy = UPW2(v, c, vY, j, ny1, dy, v[c]);
          const double dvdz = UPW2(v, c, vZ, k, nz, dz, dmb[IFY(i, j, k)]);
          vint_[IFY(i, j, k)] = dvdx + dvdy + dvdz;
        }
    for (int k = 1; k <= nz1; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          const int64_t c = IW(i, j, k);
          const double dwdx = UPW2(w, c, wX, i, nx, dx, amd[IFZ(i, j, k)]);
          const double dwdy = UPW2(w, c, wY, j, ny, dy, bmd[IFZ(i, j, k)]);
          const double dwdz = UPW2(w, c, wZ, k, nz1, dz, w[c]);
          wint_[IFZ(i, j, k)] = dwdx + dwdy + dwdz;
        }
#undef UPW2
  }
  free(bma); free(dma); free(amb); free(dmb); free(amd); free(bmd);
}

/* The advecting velocities of rotacional / antissim (7_advectivo.f90:334-423,
 * 471-600) at a face of each component: the component itself along its own
 * direction, four-face averages of the others, each in the reference's order. */
static double av_bma(const or_state *s, int i, int j, int k) {
  DIMS; const double *v = s->v;
  return (v[IV(i, j, k)] + v[IV(i - 1, j, k)] + v[IV(i, j + 1, k)] + v[IV(i - 1, j + 1, k)]) * 0.25;
}
static double av_dma(const or_state *s, int i, int j, int k) {
  DIMS; const double *w = s->w;
  return (w[IW(i, j, k)] + w[IW(i - 1, j, k)] + w[IW(i, j, k + 1)] + w[IW(i - 1, j, k + 1)]) * 0.25;
}
static double av_amb(const or_state *s, int i, int j, int k) {
  DIMS; const double *u = s->u;
  return (u[IU(i, j, k)] + u[IU(i + 1, j, k)] + u[IU(i, j - 1, k)] + u[IU(i + 1, j - 1, k)]) * 0.25;
}
static double av_dmb(const or_state *s, int i, int j, int k) {
  DIMS; const double *w = s->w;
  return (w[IW(i, j, k)] + w[IW(i, j - 1, k)] + w[IW(i, j, k + 1)] + w[IW(i, j - 1, k + 1)]) * 0.25;
}
static double av_amd(const or_state *s, int i, int j, int k) {
  DIMS; const double *u = s->u;
  return (u[IU(i, j, k)] + u[IU(i + 1, j, k)] + u[IU(i, j, k - 1)] + u[IU(i + 1, j, k - 1)]) * 0.25;
}
static double av_bmd(const or_state *s, int i, int j, int k) {
  DIMS; const double *v = s->v;
  return (v[IV(i, j, k)] + v[IV(i, j + 1, k)] + v[IV(i, j, k - 1)] + v[IV(i, j + 1, k - 1)]) * 0.25;
}

/* One direction of the flux-form upwind term for face (i,j,k) of component q
 * (ghosted array, stride st): rotacional (:356-360,...) when anti = 0, the
 * antisymmetric average with the advective form (:512-523,...) when anti = 1.
 * Ap, Am, Ac are the advecting velocities along this direction at the faces
 * behind, ahead and here; for antissim the caller has already copied the
 * end entries (ap(0) = ap(1), an(n+2) = an(n+1)). */
static double flux_term(const double *q, int64_t c, int64_t st, double Am, double Ac, double Ap, double h,
                        int anti) {
  const double apc = fmax_f(Ac, 0.), apm = fmax_f(Am, 0.), anp = fmin_f(Ap, 0.), anc = fmin_f(Ac, 0.);
  if (!anti) return (apc * q[c] - apm * q[c - st]) / h + (anp * q[c + st] - anc * q[c]) / h;
  return ((apc * q[c] - apm * q[c - st]) / h + (anp * q[c + st] - anc * q[c]) / h + apc * (q[c] - q[c - st]) / h +
          anc * (q[c + st] - q[c]) / h) *
         0.5;
}

/* rotacional (adv_type = 2, der = 1), 7_advectivo.f90:296-440, and antissim
 * (adv_type = 3, der = 1), :443-653.  rotacional's w loop reads dn at
 * k = nz+2, an entry the reference never sets (it assigns dp there,
 * :419): this restatement uses min(w(nz+2), 0), the evident intent, so
 * Fw on the plane k = nz+1 is parity unpinned for adv_type = 2. */
static void flux_form(const or_state *s, double *uint_, double *vint_, double *wint_, int anti) {
  DIMS;
  const double *u = s->u, *v = s->v, *w = s->w;
  const double dx = s->dx, dy = s->dy, dz = s->dz;
  const int64_t uX = 1, uY = nx + 3, uZ = (int64_t)(nx + 3) * (ny + 2);
  const int64_t vX = 1, vY = nx + 2, vZ = (int64_t)(nx + 2) * (ny + 3);
  const int64_t wX = 1, wY = nx + 2, wZ = (int64_t)(nx + 2) * (ny + 2);
  /* antissim copies the first entry of ap, bp, dp from the second and the
   * last entry of the normal direction's an from the one before */
#define LO(f) (anti && (f) < 1 ? 1 : (f))
#define HI(f, n) (anti && (f) > (n) ? (n) : (f))
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx + 1; i++) {
        const int64_t c = IU(i, j, k);
        const double tx = flux_term(u, c, uX, u[IU(LO(i - 1), j, k)], u[c], u[IU(HI(i + 1, nx + 1), j, k)], dx, anti);
        const double ty = flux_term(u, c, uY, av_bma(s, i, LO(j - 1), k), av_bma(s, i, j, k), av_bma(s, i, j + 1, k),
                                    dy, anti);
        const double tz = flux_term(u, c, uZ, av_dma(s, i, j, LO(k - 1)), av_dma(s, i, j, k), av_dma(s, i, j, k + 1),
                                    dz, anti);
        uint_[IFX(i, j, k)] = tx + ty + tz;
      }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny + 1; j++)
      for (int i = 1; i <= nx; i++) {
        const int64_t c = IV(i, j, k);
        const double tx = flux_term(v, c, vX, av_amb(s, LO(i - 1), j, k), av_amb(s, i, j, k), av_amb(s, i + 1, j, k),
                                    dx, anti);
        const double ty = flux_term(v, c, vY, v[IV(i, LO(j - 1), k)], v[c], v[IV(i, HI(j + 1, ny + 1), k)], dy, anti);
        const double tz = flux_term(v, c, vZ, av_dmb(s, i, j, LO(k - 1)), av_dmb(s, i, j, k), av_dmb(s, i, j, k + 1),
                                    dz, anti);
        vint_[IFY(i, j, k)] = tx + ty + tz;
      }
  for (int k = 1; k <= nz + 1; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const int64_t c = IW(i, j, k);
        const double tx = flux_term(w, c, wX, av_amd(s, LO(i - 1), j, k), av_amd(s, i, j, k), av_amd(s, i + 1, j, k),
                                    dx, anti);
        const double ty = flux_term(w, c, wY, av_bmd(s, i, LO(j - 1), k), av_bmd(s, i, j, k), av_bmd(s, i, j + 1, k),
                                    dy, anti);
        const double tz = flux_term(w, c, wZ, w[IW(i, j, LO(k - 1))], w[c], w[IW(i, j, HI(k + 1, nz + 1))], dz, anti);
        wint_[IFZ(i, j, k)] = tx + ty + tz;
      }
#undef LO
#undef HI
}

void or_classico(const or_state *s, double *uint_, double *vint_, double *wint_) {
  DIMS;
  const double *u = s->u, *v = s->v, *w = s->w;
  const double dx = s->dx, dy = s->dy, dz = s->dz;
  if (s->der == 2 || s->der == 3) {
    classico_der23(s, uint_, vint_, wint_);
    return;
  }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx + 1; i++) {
        double bma = (v[IV(i, j, k)] + v[IV(i - 1, j, k)] + v[IV(i, j + 1, k)] + v[IV(i - 1, j + 1, k)]) * 0.25;
        double dma = (w[IW(i, j, k)] + w[IW(i - 1, j, k)] + w[IW(i, j, k + 1)] + w[IW(i - 1, j, k + 1)]) * 0.25;
        double a1 = fmax_f(u[IU(i, j, k)], 0.), a2 = fmin_f(u[IU(i, j, k)], 0.);
        double dudx = a1 * (u[IU(i, j, k)] - u[IU(i - 1, j, k)]) / dx + a2 * (u[IU(i + 1, j, k)] - u[IU(i, j, k)]) / dx;
        a1 = fmax_f(bma, 0.); a2 = fmin_f(bma, 0.);
        double dudy = a1 * (u[IU(i, j, k)] - u[IU(i, j - 1, k)]) / dy + a2 * (u[IU(i, j + 1, k)] - u[IU(i, j, k)]) / dy;
        a1 = fmax_f(dma, 0.); a2 = fmin_f(dma, 0.);
        double dudz = a1 * (u[IU(i, j, k)] - u[IU(i, j, k - 1)]) / dz + a2 * (u[IU(i, j, k + 1)] - u[IU(i, j, k)]) / dz;
        uint_[IFX(i, j, k)] = dudx + dudy + dudz;
      }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny + 1; j++)
      for (int i = 1; i <= nx; i++) {
        double amb = (u[IU(i, j, k)] + u[IU(i + 1, j, k)] + u[IU(i, j - 1, k)] + u[IU(i + 1, j - 1, k)]) * 0.25;
        double dmb = (w[IW(i, j, k)] + w[IW(i, j - 1, k)] + w[IW(i, j, k + 1)] + w[IW(i, j - 1, k + 1)]) * 0.25;
        double a1 = fmax_f(amb, 0.), a2 = fmin_f(amb, 0.);
        double dvdx = a1 * (v[IV(i, j, k)] - v[IV(i - 1, j, k)]) / dx + a2 * (v[IV(i + 1, j, k)] - v[IV(i, j, k)]) / dx;
        a1 = fmax_f(v[IV(i, j, k)], 0.); a2 = fmin_f(v[IV(i, j, k)], 0.);
        double dvdy = a1 * (v[IV(i, j, k)] - v[IV(i, j - 1, k)]) / dy + a2 * (v[IV(i, j + 1, k)] - v[IV(i, j, k)]) / dy;
        a1 = fmax_f(dmb, 0.); a2 = fmin_f(dmb, 0.);
        double dvdz = a1 * (v[IV(i, j, k)] - v[IV(i, j, k - 1)]) / dz + a2 * (v[IV(i, j, k + 1)] - v[IV(i, j, k)]) / dz;
        vint_[IFY(i, j, k)] = dvdx + dvdy + dvdz;
      }
  for (int k = 1; k <= nz + 1; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double amd = (u[IU(i, j, k)] + u[IU(i + 1, j, k)] + u[IU(i, j, k - 1)] + u[IU(i + 1, j, k - 1)]) * 0.25;
        double bmd = (v[IV(i, j, k)] + v[IV(i, j + 1, k)] + v[IV(i, j, k - 1)] + v[IV(i, j + 1, k - 1)]) * 0.25;
        double a1 = fmax_f(amd, 0.), a2 = fmin_f(amd, 0.);
        double dwdx = a1 * (w[IW(i, j, k)] - w[IW(i - 1, j, k)]) / dx + a2 * (w[IW(i + 1, j, k)] - w[IW(i, j, k)]) / dx;
        a1 = fmax_f(bmd, 0.); a2 = fmin_f(bmd, 0.);
        double dwdy = a1 * (w[IW(i, j, k)] - w[IW(i, j - 1, k)]) / dy + a2 * (w[IW(i, j + 1, k)] - w[IW(i, j, k)]) / dy;
        a1 = fmax_f(w[IW(i, j, k)], 0.); a2 = fmin_f(w[IW(i, j, k)], 0.);
        double dwdz = a1 * (w[IW(i, j, k)] - w[IW(i, j, k - 1)]) / dz + a2 * (w[IW(i, j, k + 1)] - w[IW(i, j, k)]) / dz;
        wint_[IFZ(i, j, k)] = dwdx + dwdy + dwdz;
      }
}

/* convdiff, 6_convdiff.f90:14-189 (adv_type 1, 2, 3 at :70-82; no sponge layer) */
void or_convdiff(or_state *s) {
  DIMS;
  const int nx1 = nx + 1, ny1 = ny + 1, nz1 = nz + 1;
  const int64_t C = (int64_t)nx * ny * nz, FX = (int64_t)nx1 * ny * nz, FY = (int64_t)nx * ny1 * nz,
                FZ = (int64_t)nx * ny * nz1;
  const double dx = s->dx, dy = s->dy, dz = s->dz;
  const double *u = s->u, *v = s->v, *w = s->w;
  double *uint_ = alloc_n(FX), *vint_ = alloc_n(FY), *wint_ = alloc_n(FZ);
  if (s->adv_type == 2 || s->adv_type == 3) flux_form(s, uint_, vint_, wint_, s->adv_type == 3);
  else or_classico(s, uint_, vint_, wint_);
  double *rhox = alloc_n(FX), *rhoy = alloc_n(FY), *rhoz = alloc_n(FZ);
  interpx_cf(s->rho, nx, ny, nz, rhox);
  interpy_cf(s->rho, nx, ny, nz, rhoy);
  interpz_cf(s->rho, nx, ny, nz, rhoz);
  double *tmp = alloc_n(C);
  double *dhsdx = alloc_n(FX), *dhsdy = alloc_n(FY), *dhsdz = alloc_n(FZ);
  for (int64_t c = 0; c < C; c++) tmp[c] = s->hsx[c] * s->kurv[c] / s->rho[c];
  interpx_cf(tmp, nx, ny, nz, dhsdx);
  for (int64_t c = 0; c < C; c++) tmp[c] = s->hsy[c] * s->kurv[c] / s->rho[c];
  interpy_cf(tmp, nx, ny, nz, dhsdy);
  for (int64_t c = 0; c < C; c++) tmp[c] = s->hsz[c] * s->kurv[c] / s->rho[c];
  interpz_cf(tmp, nx, ny, nz, dhsdz);
  /* edge viscosities, :118-145 */
  int64_t EXY = (int64_t)nx1 * ny1 * nz, EXZ = (int64_t)nx1 * ny * nz1, EYZ = (int64_t)nx * ny1 * nz1;
  double *bma = alloc_n(EXY), *dma = alloc_n(EXZ), *amb = alloc_n(EXY), *dmb = alloc_n(EYZ);
  double *amd = alloc_n(EXZ), *bmd = alloc_n(EYZ);
  interpy_cf(s->xnut, nx1, ny, nz, bma);
  interpz_cf(s->xnut, nx1, ny, nz, dma);
  interpx_cf(s->ynut, nx, ny1, nz, amb);
  interpz_cf(s->ynut, nx, ny1, nz, dmb);
  interpx_cf(s->znut, nx, ny, nz1, amd);
  interpy_cf(s->znut, nx, ny, nz1, bmd);
  for (int64_t c = 0; c < EXY; c++) { bma[c] = bma[c] / dy; amb[c] = amb[c] / dx; }
  for (int64_t c = 0; c < EXZ; c++) { dma[c] = dma[c] / dz; amd[c] = amd[c] / dx; }
  for (int64_t c = 0; c < EYZ; c++) { dmb[c] = dmb[c] / dz; bmd[c] = bmd[c] / dy; }
  /* ama(0:nx1,ny,nz), bmb(nx,0:ny1,nz), dmd(nx,ny,0:nz1) */
#define AMA(i, j, k) ama[(int64_t)(i) + (int64_t)(nx + 2) * ((int64_t)((j)-1) + (int64_t)ny * (int64_t)((k)-1))]
#define BMB(i, j, k) bmb[(int64_t)((i)-1) + (int64_t)nx * ((int64_t)(j) + (int64_t)(ny + 2) * (int64_t)((k)-1))]
#define DMD(i, j, k) dmd[(int64_t)((i)-1) + (int64_t)nx * ((int64_t)((j)-1) + (int64_t)ny * (int64_t)(k))]
  double *ama = alloc_n((int64_t)(nx + 2) * ny * nz), *bmb = alloc_n((int64_t)nx * (ny + 2) * nz),
         *dmd = alloc_n((int64_t)nx * ny * (nz + 2));
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        AMA(i, j, k) = s->nut[IC(i, j, k)] / dx;
        BMB(i, j, k) = s->nut[IC(i, j, k)] / dy;
        DMD(i, j, k) = s->nut[IC(i, j, k)] / dz;
      }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++) { AMA(0, j, k) = AMA(1, j, k); AMA(nx1, j, k) = AMA(nx, j, k); }
  for (int k = 1; k <= nz; k++)
    for (int i = 1; i <= nx; i++) { BMB(i, 0, k) = BMB(i, 1, k); BMB(i, ny1, k) = BMB(i, ny, k); }
  for (int j = 1; j <= ny; j++)
    for (int i = 1; i <= nx; i++) { DMD(i, j, 0) = DMD(i, j, 1); DMD(i, j, nz1) = DMD(i, j, nz); }
#define EXYI(i, j, k) IA(i, j, k, nx1, ny1)
#define EXZI(i, j, k) IA(i, j, k, nx1, ny)
#define EYZI(i, j, k) IA(i, j, k, nx, ny1)
  const double fr = s->gz / (s->chezy * s->chezy);
  for (int k = 1; k <= nz; k++) {
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx1; i++) {
        double tfu = s->tf_u ? s->tf_u[IFX(i, j, k)] : 0.;
        double ubv = s->ub[IU(i, j, k)];
        s->Fu[IFX(i, j, k)] =
            -uint_[IFX(i, j, k)] +
            ((AMA(i, j, k) * 2. * (u[IU(i + 1, j, k)] - u[IU(i, j, k)]) -
              AMA(i - 1, j, k) * 2. * (u[IU(i, j, k)] - u[IU(i - 1, j, k)])) / dx +
             bma[EXYI(i, j + 1, k)] * ((u[IU(i, j + 1, k)] - u[IU(i, j, k)]) / dy + (v[IV(i, j + 1, k)] - v[IV(i - 1, j + 1, k)]) / dx) -
             bma[EXYI(i, j, k)] * ((u[IU(i, j, k)] - u[IU(i, j - 1, k)]) / dy + (v[IV(i, j, k)] - v[IV(i - 1, j, k)]) / dx) +
             dma[EXZI(i, j, k + 1)] * ((u[IU(i, j, k + 1)] - u[IU(i, j, k)]) / dz + (w[IW(i, j, k + 1)] - w[IW(i - 1, j, k + 1)]) / dx) -
             dma[EXZI(i, j, k)] * ((u[IU(i, j, k)] - u[IU(i, j, k - 1)]) / dz + (w[IW(i, j, k)] - w[IW(i - 1, j, k)]) / dx)) /
                rhox[IFX(i, j, k)] +
            s->gx - s->sigma * dhsdx[IFX(i, j, k)] + tfu - fr * sqrt(ubv * ubv) * u[IU(i, j, k)];
      }
    for (int j = 1; j <= ny1; j++)
      for (int i = 1; i <= nx; i++) {
        double tfv = s->tf_v ? s->tf_v[IFY(i, j, k)] : 0.;
        double vbv = s->vb[IV(i, j, k)];
        s->Fv[IFY(i, j, k)] =
            -vint_[IFY(i, j, k)] +
            (amb[EXYI(i + 1, j, k)] * ((v[IV(i + 1, j, k)] - v[IV(i, j, k)]) / dx + (u[IU(i + 1, j, k)] - u[IU(i + 1, j - 1, k)]) / dy) -
             amb[EXYI(i, j, k)] * ((v[IV(i, j, k)] - v[IV(i - 1, j, k)]) / dx + (u[IU(i, j, k)] - u[IU(i, j - 1, k)]) / dy) +
             (BMB(i, j, k) * 2. * (v[IV(i, j + 1, k)] - v[IV(i, j, k)]) -
              BMB(i, j - 1, k) * 2. * (v[IV(i, j, k)] - v[IV(i, j - 1, k)])) / dy +
             dmb[EYZI(i, j, k + 1)] * ((v[IV(i, j, k + 1)] - v[IV(i, j, k)]) / dz + (w[IW(i, j, k + 1)] - w[IW(i, j - 1, k + 1)]) / dy) -
             dmb[EYZI(i, j, k)] * ((v[IV(i, j, k)] - v[IV(i, j, k - 1)]) / dz + (w[IW(i, j, k)] - w[IW(i, j - 1, k)]) / dy)) /
                rhoy[IFY(i, j, k)] -
            s->sigma * dhsdy[IFY(i, j, k)] + tfv - fr * sqrt(vbv * vbv) * v[IV(i, j, k)];
      }
  }
  for (int k = 1; k <= nz1; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double tfw = s->tf_w ? s->tf_w[IFZ(i, j, k)] : 0.;
        double wbv = s->wb[IW(i, j, k)];
        const double epis_z = 0.;
        s->Fw[IFZ(i, j, k)] =
            -wint_[IFZ(i, j, k)] +
            (amd[EXZI(i + 1, j, k)] * ((w[IW(i + 1, j, k)] - w[IW(i, j, k)]) / dx + (u[IU(i + 1, j, k)] - u[IU(i + 1, j, k - 1)]) / dz) -
             amd[EXZI(i, j, k)] * ((w[IW(i, j, k)] - w[IW(i - 1, j, k)]) / dx + (u[IU(i, j, k)] - u[IU(i, j, k - 1)]) / dz) +
             bmd[EYZI(i, j + 1, k)] * ((w[IW(i, j + 1, k)] - w[IW(i, j, k)]) / dy + (v[IV(i, j + 1, k)] - v[IV(i, j + 1, k - 1)]) / dz) -
             bmd[EYZI(i, j, k)] * ((w[IW(i, j, k)] - w[IW(i, j - 1, k)]) / dy + (v[IV(i, j, k)] - v[IV(i, j, k - 1)]) / dz) +
             (DMD(i, j, k) * 2. * (w[IW(i, j, k + 1)] - w[IW(i, j, k)]) -
              DMD(i, j, k - 1) * 2. * (w[IW(i, j, k)] - w[IW(i, j, k - 1)])) / dz) /
                rhoz[IFZ(i, j, k)] -
            (w[IW(i, j, k)] - 0.) * epis_z - s->gz - s->sigma * dhsdz[IFZ(i, j, k)] + tfw -
            fr * sqrt(wbv * wbv) * w[IW(i, j, k)];
      }
#undef AMA
#undef BMB
#undef DMD
  free(uint_); free(vint_); free(wint_); free(rhox); free(rhoy); free(rhoz); free(tmp);
  free(dhsdx); free(dhsdy); free(dhsdz); free(bma); free(dma); free(amb); free(dmb); free(amd); free(bmd);
  free(ama); free(bmb); free(dmd);
}

/* tempo, 11_diversos.f90:10-228: the update of one face of one component
 * (q the velocity, q0 the start-of-step copy, F the forces, f0/f1 the kept
 * forces fu0/fu1), for every t_tempo scheme, in the reference's order of
 * evaluation (left to right). */
static void tempo_face(const or_state *s, double *q, double *q0, const double *F, double *f0, double *f1) {
  const double dt = s->dt;
  if (s->t_tempo == 0 || (s->t_tempo == 3 && s->it == 1)) { /* :19-52 */
    *q = *q + dt * *F;
    if (s->t_tempo == 3) *f0 = *F;
  } else if (s->t_tempo == 1) { /* :53-111 */
    const double aux1 = dt;
    if (s->tt == 1) {
      *q0 = *q;
      *q = *q + aux1 * *F;
    } else if (s->tt == 2) {
      *q = *q0 + aux1 * *F;
    }
  } else if (s->t_tempo == 2) { /* :112-198 */
    if (s->tt == 1) {
      const double aux1 = dt;
      *q0 = *q;
      *f0 = *F;
      *q = *q + aux1 * *F;
    } else if (s->tt == 2) {
      const double aux1 = 2. * dt;
      *f1 = *F;
      *q = *q0 - dt * *f0 + aux1 * *F;
    } else if (s->tt == 3) {
      const double aux1 = dt / 6.;
      *q = *q0 + aux1 * (*f0 + 4. * *f1 + *F);
    }
  } else if (s->t_tempo == 3) { /* it /= 1, :199-226 */
    *q = *q + dt * 0.5 * (3 * *F - *f0);
    *f0 = *F;
  }
}

void or_tempo(or_state *s) {
  DIMS;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx + 1; i++)
        tempo_face(s, &s->u[IU(i, j, k)], &s->u0[IFX(i, j, k)], &s->Fu[IFX(i, j, k)], &s->fu0[IFX(i, j, k)],
                   &s->fu1[IFX(i, j, k)]);
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny + 1; j++)
      for (int i = 1; i <= nx; i++)
        tempo_face(s, &s->v[IV(i, j, k)], &s->v0[IFY(i, j, k)], &s->Fv[IFY(i, j, k)], &s->fv0[IFY(i, j, k)],
                   &s->fv1[IFY(i, j, k)]);
  for (int k = 1; k <= nz + 1; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        tempo_face(s, &s->w[IW(i, j, k)], &s->w0[IFZ(i, j, k)], &s->Fw[IFZ(i, j, k)], &s->fw0[IFZ(i, j, k)],
                   &s->fw1[IFZ(i, j, k)]);
}

/* number of stages and their steps a_dt, 2_iniciais.f90:85-102 */
static int or_ntt(int t_tempo) { return t_tempo == 2 ? 3 : t_tempo == 1 ? 2 : 1; }
static double or_a_dt(int t_tempo, int tt, double dt) {
  return (or_ntt(t_tempo) > 1 && tt == 1) ? 0.5 * dt : 1.0 * dt;
}

/* plane n of the prescribed-velocity walls at (a, b), first extent n1 */
#define BP(n, a, b, n1) (s->bp[n][(a) + (int64_t)(n1) * (b)])
enum { BXX0, BXX1, BXY0, BXZ0, BXXF, BXXF1, BXYF, BXZF, BYX0, BYY0, BYY1, BYZ0, BYXF, BYYF, BYYF1, BYZF,
       BZX0, BZY0, BZZ0, BZZ1, BZXF, BZYF, BZZF, BZZF1 };

/* contorno, 3_contorno.f90:29-443 with prd_corr (:872-950, every dp ≡ 0);
 * wall codes 0 (periodic), 1 (free slip), 2 (no slip), 3 (prescribed: the
 * ghost planes from the boundary planes bxx0..bzzf1, the wall-normal faces
 * from bxx1, bxxf, byy1, byyf, bzz1, bzzf), 4 (outflow at xf; at x = 0 the
 * ghost v from bxy0 and w copied from the first cells);
 * a bottom obstacle (obst_t != 0) zeroes the velocities of its columns and
 * moves the friction planes on top of it (:266-291); nlock = 3 fills the
 * level set inside it (:445-490). */
void or_contorno(or_state *s, int nlock) {
  DIMS;
  const int nx1 = nx + 1, ny1 = ny + 1, nz1 = nz + 1;
  double *u = s->u, *v = s->v, *w = s->w, *p = s->prd1;
  int i, j, k;
  if (nlock == 1 || nlock == 2) {
    /* pressure ghosts, :37-54 */
    for (k = 0; k <= nz + 1; k++)
      for (j = 0; j <= ny + 1; j++) {
        if (s->ccx0 == 0 && s->ccxf == 0) { p[IP(0, j, k)] = p[IP(nx, j, k)]; p[IP(nx + 1, j, k)] = p[IP(1, j, k)]; }
        else { p[IP(0, j, k)] = p[IP(1, j, k)]; p[IP(nx + 1, j, k)] = p[IP(nx, j, k)]; }
      }
    for (k = 0; k <= nz + 1; k++)
      for (i = 0; i <= nx + 1; i++) {
        if (s->ccy0 == 0 && s->ccyf == 0) { p[IP(i, 0, k)] = p[IP(i, ny, k)]; p[IP(i, ny + 1, k)] = p[IP(i, 1, k)]; }
        else { p[IP(i, 0, k)] = p[IP(i, 1, k)]; p[IP(i, ny + 1, k)] = p[IP(i, ny, k)]; }
      }
    for (j = 0; j <= ny + 1; j++)
      for (i = 0; i <= nx + 1; i++) { p[IP(i, j, 0)] = p[IP(i, j, 1)]; p[IP(i, j, nz + 1)] = p[IP(i, j, nz)]; }
    /* y = 0 wall, :58-85 */
    if (s->ccy0 == 0 && s->ccyf == 0) {
      for (k = 0; k <= nz + 1; k++) {
        for (i = 0; i <= nx1 + 1; i++) u[IU(i, 0, k)] = u[IU(i, ny, k)];
        for (i = 0; i <= nx + 1; i++) { v[IV(i, 0, k)] = v[IV(i, ny1 - 1, k)]; v[IV(i, 1, k)] = v[IV(i, ny1, k)]; }
      }
      for (k = 0; k <= nz1 + 1; k++) for (i = 0; i <= nx + 1; i++) w[IW(i, 0, k)] = w[IW(i, ny, k)];
    }
    if (s->ccy0 == 1 || s->ccy0 == 2) {
      double sg = (s->ccy0 == 1) ? 1. : -1.;
      for (k = 0; k <= nz + 1; k++) {
        for (i = 0; i <= nx1 + 1; i++) u[IU(i, 0, k)] = (s->ccy0 == 1) ? u[IU(i, 1, k)] : -u[IU(i, 1, k)];
        for (i = 0; i <= nx + 1; i++) v[IV(i, 0, k)] = (s->ccy0 == 1) ? -v[IV(i, 2, k)] : v[IV(i, 2, k)];
      }
      for (k = 0; k <= nz1 + 1; k++) for (i = 0; i <= nx + 1; i++) w[IW(i, 0, k)] = (s->ccy0 == 1) ? w[IW(i, 1, k)] : -w[IW(i, 1, k)];
      (void)sg;
    }
    if (s->ccy0 == 3) { /* :80-84 */
      for (k = 0; k <= nz + 1; k++) {
        for (i = 0; i <= nx1 + 1; i++) u[IU(i, 0, k)] = BP(BYX0, i, k, nx1 + 2);
        for (i = 0; i <= nx + 1; i++) v[IV(i, 0, k)] = BP(BYY0, i, k, nx + 2);
      }
      for (k = 0; k <= nz1 + 1; k++) for (i = 0; i <= nx + 1; i++) w[IW(i, 0, k)] = BP(BYZ0, i, k, nx + 2);
    }
    /* y = yf wall, :87-114 */
    if (s->ccy0 == 0 && s->ccyf == 0) {
      for (k = 0; k <= nz + 1; k++) {
        for (i = 0; i <= nx1 + 1; i++) u[IU(i, ny + 1, k)] = u[IU(i, 1, k)];
        for (i = 0; i <= nx + 1; i++) v[IV(i, ny1 + 1, k)] = v[IV(i, 2, k)];
      }
      for (k = 0; k <= nz1 + 1; k++) for (i = 0; i <= nx + 1; i++) w[IW(i, ny + 1, k)] = w[IW(i, 1, k)];
    }
    if (s->ccyf == 1 || s->ccyf == 2) {
      for (k = 0; k <= nz + 1; k++) {
        for (i = 0; i <= nx1 + 1; i++) u[IU(i, ny + 1, k)] = (s->ccyf == 1) ? u[IU(i, ny, k)] : -u[IU(i, ny, k)];
        for (i = 0; i <= nx + 1; i++) v[IV(i, ny1 + 1, k)] = (s->ccyf == 1) ? -v[IV(i, ny1 - 1, k)] : v[IV(i, ny1 - 1, k)];
      }
      for (k = 0; k <= nz1 + 1; k++) for (i = 0; i <= nx + 1; i++) w[IW(i, ny + 1, k)] = (s->ccyf == 1) ? w[IW(i, ny, k)] : -w[IW(i, ny, k)];
    }
    if (s->ccyf == 3) { /* :109-113 */
      for (k = 0; k <= nz + 1; k++) {
        for (i = 0; i <= nx1 + 1; i++) u[IU(i, ny + 1, k)] = BP(BYXF, i, k, nx1 + 2);
        for (i = 0; i <= nx + 1; i++) v[IV(i, ny1 + 1, k)] = BP(BYYF1, i, k, nx + 2);
      }
      for (k = 0; k <= nz1 + 1; k++) for (i = 0; i <= nx + 1; i++) w[IW(i, ny + 1, k)] = BP(BYZF, i, k, nx + 2);
    }
    /* x = 0 wall, :116-161 */
    if (s->ccx0 == 0 && s->ccxf == 0) {
      for (k = 0; k <= nz + 1; k++) {
        for (j = 0; j <= ny + 1; j++) u[IU(0, j, k)] = u[IU(nx1 - 1, j, k)];
        for (j = 0; j <= ny1 + 1; j++) v[IV(0, j, k)] = v[IV(nx, j, k)];
      }
      for (k = 0; k <= nz1 + 1; k++) for (j = 0; j <= ny + 1; j++) w[IW(0, j, k)] = w[IW(nx, j, k)];
    }
    if (s->ccx0 == 1 || s->ccx0 == 2) {
      for (k = 0; k <= nz + 1; k++) {
        for (j = 0; j <= ny + 1; j++) u[IU(0, j, k)] = (s->ccx0 == 1) ? -u[IU(2, j, k)] : u[IU(2, j, k)];
        for (j = 0; j <= ny1 + 1; j++) v[IV(0, j, k)] = (s->ccx0 == 1) ? v[IV(1, j, k)] : -v[IV(1, j, k)];
      }
      for (k = 0; k <= nz1 + 1; k++) for (j = 0; j <= ny + 1; j++) w[IW(0, j, k)] = (s->ccx0 == 1) ? w[IW(1, j, k)] : -w[IW(1, j, k)];
    }
    if (s->ccx0 == 3) { /* :136-140 */
      for (k = 0; k <= nz + 1; k++) {
        for (j = 0; j <= ny + 1; j++) u[IU(0, j, k)] = BP(BXX0, j, k, ny + 2);
        for (j = 0; j <= ny1 + 1; j++) v[IV(0, j, k)] = BP(BXY0, j, k, ny1 + 2);
      }
      for (k = 0; k <= nz1 + 1; k++) for (j = 0; j <= ny + 1; j++) w[IW(0, j, k)] = BP(BXZ0, j, k, ny + 2);
    }
    if (s->ccx0 == 4) { /* :143-146 */
      for (k = 0; k <= nz + 1; k++) for (j = 0; j <= ny1 + 1; j++) v[IV(0, j, k)] = BP(BXY0, j, k, ny1 + 2);
      for (k = 0; k <= nz1 + 1; k++) for (j = 0; j <= ny + 1; j++) w[IW(0, j, k)] = w[IW(1, j, k)];
    }
    /* x = xf wall, :163-197 */
    if (s->ccx0 == 0 && s->ccxf == 0) {
      for (k = 0; k <= nz + 1; k++) {
        for (j = 0; j <= ny + 1; j++) u[IU(nx1 + 1, j, k)] = u[IU(2, j, k)];
        for (j = 0; j <= ny1 + 1; j++) v[IV(nx + 1, j, k)] = v[IV(1, j, k)];
      }
      for (k = 0; k <= nz1 + 1; k++) for (j = 0; j <= ny + 1; j++) w[IW(nx + 1, j, k)] = w[IW(1, j, k)];
    }
    if (s->ccxf == 1 || s->ccxf == 2 || s->ccxf == 4) {
      for (k = 0; k <= nz + 1; k++) {
        for (j = 0; j <= ny + 1; j++)
          u[IU(nx1 + 1, j, k)] = (s->ccxf == 1) ? -u[IU(nx1 - 1, j, k)] : (s->ccxf == 2) ? u[IU(nx1 - 1, j, k)] : u[IU(nx1, j, k)];
        for (j = 0; j <= ny1 + 1; j++) v[IV(nx + 1, j, k)] = (s->ccxf == 2) ? -v[IV(nx, j, k)] : v[IV(nx, j, k)];
      }
      for (k = 0; k <= nz1 + 1; k++)
        for (j = 0; j <= ny + 1; j++) w[IW(nx + 1, j, k)] = (s->ccxf == 2) ? -w[IW(nx, j, k)] : w[IW(nx, j, k)];
    }
    if (s->ccxf == 3) { /* :185-189 */
      for (k = 0; k <= nz + 1; k++) {
        for (j = 0; j <= ny + 1; j++) u[IU(nx1 + 1, j, k)] = BP(BXXF1, j, k, ny + 2);
        for (j = 0; j <= ny1 + 1; j++) v[IV(nx + 1, j, k)] = BP(BXYF, j, k, ny1 + 2);
      }
      for (k = 0; k <= nz1 + 1; k++) for (j = 0; j <= ny + 1; j++) w[IW(nx + 1, j, k)] = BP(BXZF, j, k, ny + 2);
    }
    /* bottom, :199-219 */
    if (s->ccz0 == 1 || s->ccz0 == 2) {
      for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx1 + 1; i++) u[IU(i, j, 0)] = (s->ccz0 == 1) ? u[IU(i, j, 1)] : -u[IU(i, j, 1)];
      for (j = 0; j <= ny1 + 1; j++) for (i = 0; i <= nx + 1; i++) v[IV(i, j, 0)] = (s->ccz0 == 1) ? v[IV(i, j, 1)] : -v[IV(i, j, 1)];
      for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx + 1; i++) w[IW(i, j, 0)] = (s->ccz0 == 1) ? -w[IW(i, j, 2)] : w[IW(i, j, 2)];
    }
    if (s->ccz0 == 3) { /* :214-218 */
      for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx1 + 1; i++) u[IU(i, j, 0)] = BP(BZX0, i, j, nx1 + 2);
      for (j = 0; j <= ny1 + 1; j++) for (i = 0; i <= nx + 1; i++) v[IV(i, j, 0)] = BP(BZY0, i, j, nx + 2);
      for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx + 1; i++) w[IW(i, j, 0)] = BP(BZZ0, i, j, nx + 2);
    }
    /* top, :221-241 */
    if (s->cczf == 1 || s->cczf == 2) {
      for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx1 + 1; i++) u[IU(i, j, nz + 1)] = (s->cczf == 1) ? u[IU(i, j, nz)] : -u[IU(i, j, nz)];
      for (j = 0; j <= ny1 + 1; j++) for (i = 0; i <= nx + 1; i++) v[IV(i, j, nz + 1)] = (s->cczf == 1) ? v[IV(i, j, nz)] : -v[IV(i, j, nz)];
      for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx + 1; i++) w[IW(i, j, nz1 + 1)] = (s->cczf == 1) ? -w[IW(i, j, nz1 - 1)] : w[IW(i, j, nz1 - 1)];
    }
    if (s->cczf == 3) { /* :236-240 */
      for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx1 + 1; i++) u[IU(i, j, nz + 1)] = BP(BZXF, i, j, nx1 + 2);
      for (j = 0; j <= ny1 + 1; j++) for (i = 0; i <= nx + 1; i++) v[IV(i, j, nz + 1)] = BP(BZYF, i, j, nx + 2);
      for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx + 1; i++) w[IW(i, j, nz1 + 1)] = BP(BZZF1, i, j, nx + 2);
    }
    if (s->obst_t != 0) { /* bottom obstacle, :266-291 */
#define KU(i, j) ((int)s->ku[(i) + (int64_t)(nx + 3) * (j)])
#define KV(i, j) ((int)s->kv[(i) + (int64_t)(nx + 2) * (j)])
#define KW(i, j) ((int)s->kw[(i) + (int64_t)(nx + 2) * (j)])
      for (j = 0; j <= ny + 1; j++)
        for (i = 0; i <= nx + 1; i++) {
          for (k = 0; k <= KU(i, j); k++) u[IU(i, j, k)] = 0.;
          for (k = 0; k <= KV(i, j); k++) v[IV(i, j, k)] = 0.;
          for (k = 0; k <= KW(i, j); k++) w[IW(i, j, k)] = 0.;
          s->ub[IU(i, j, KU(i, j) + 1)] = u[IU(i, j, KU(i, j) + 1)];
          s->vb[IV(i, j, KV(i, j) + 1)] = v[IV(i, j, KV(i, j) + 1)];
          s->wb[IW(i, j, KW(i, j) + 1)] = w[IW(i, j, KW(i, j) + 1)];
        }
      for (j = 0; j <= ny + 1; j++)
        for (k = 0; k <= KU(nx1 + 1, j); k++) u[IU(nx1 + 1, j, k)] = 0.;
      for (i = 0; i <= nx + 1; i++)
        for (k = 0; k <= KV(i, ny1 + 1); k++) v[IV(i, ny1 + 1, k)] = 0.;
    } else if (s->ccz0 == 2) { /* friction planes, obst_t = 0, :247-264 */
      for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx1 + 1; i++) s->ub[IU(i, j, 2)] = u[IU(i, j, 2)];
      for (j = 0; j <= ny1 + 1; j++) for (i = 0; i <= nx + 1; i++) s->vb[IV(i, j, 2)] = v[IV(i, j, 2)];
      for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx + 1; i++) s->wb[IW(i, j, 2)] = w[IW(i, j, 2)];
    }
    if (s->ccy0 == 2 && s->obst_t == 0) {
      for (k = 0; k <= nz + 1; k++) for (i = 0; i <= nx1 + 1; i++) s->ub[IU(i, 2, k)] = u[IU(i, 2, k)];
      for (k = 0; k <= nz + 1; k++) for (i = 0; i <= nx + 1; i++) s->vb[IV(i, 2, k)] = v[IV(i, 2, k)];
      for (k = 0; k <= nz1 + 1; k++) for (i = 0; i <= nx + 1; i++) s->wb[IW(i, 2, k)] = w[IW(i, 2, k)];
    }
    if (s->ccyf == 2 && s->obst_t == 0) {
      for (k = 0; k <= nz + 1; k++) for (i = 0; i <= nx1 + 1; i++) s->ub[IU(i, ny1 - 1, k)] = u[IU(i, ny1 - 1, k)];
      for (k = 0; k <= nz + 1; k++) for (i = 0; i <= nx + 1; i++) s->vb[IV(i, ny1 - 1, k)] = v[IV(i, ny1 - 1, k)];
      for (k = 0; k <= nz1 + 1; k++) for (i = 0; i <= nx + 1; i++) s->wb[IW(i, ny1 - 1, k)] = w[IW(i, ny1 - 1, k)];
    }
  }
  if (nlock == 2) { /* wall-normal velocities, :297-443 */
    if (s->ccy0 == 0 && s->ccyf == 0) {
      for (k = 0; k <= nz + 1; k++) for (i = 0; i <= nx + 1; i++) v[IV(i, 1, k)] = v[IV(i, ny1, k)];
    }
    if (s->ccy0 == 1 || s->ccy0 == 2) for (k = 0; k <= nz + 1; k++) for (i = 0; i <= nx + 1; i++) v[IV(i, 1, k)] = 0. + 0.;
    if (s->ccy0 == 3) for (k = 0; k <= nz + 1; k++) for (i = 0; i <= nx + 1; i++) v[IV(i, 1, k)] = BP(BYY1, i, k, nx + 2) + 0.;
    if (s->ccy0 == 0 && s->ccyf == 0) {
      for (k = 0; k <= nz + 1; k++) for (i = 0; i <= nx + 1; i++) v[IV(i, ny1, k)] = v[IV(i, 1, k)];
    }
    if (s->ccyf == 1 || s->ccyf == 2) for (k = 0; k <= nz + 1; k++) for (i = 0; i <= nx + 1; i++) v[IV(i, ny1, k)] = 0. + 0.;
    if (s->ccyf == 3) for (k = 0; k <= nz + 1; k++) for (i = 0; i <= nx + 1; i++) v[IV(i, ny1, k)] = BP(BYYF, i, k, nx + 2) + 0.;
    if (s->ccx0 == 0 && s->ccxf == 0) {
      for (k = 0; k <= nz + 1; k++) for (j = 0; j <= ny + 1; j++) u[IU(1, j, k)] = u[IU(nx1, j, k)];
    }
    if (s->ccx0 == 1 || s->ccx0 == 2) for (k = 0; k <= nz + 1; k++) for (j = 0; j <= ny + 1; j++) u[IU(1, j, k)] = 0. + 0.;
    if (s->ccx0 == 3) for (k = 0; k <= nz + 1; k++) for (j = 0; j <= ny + 1; j++) u[IU(1, j, k)] = BP(BXX1, j, k, ny + 2) + 0.;
    if (s->ccx0 == 0 && s->ccxf == 0) {
      for (k = 0; k <= nz + 1; k++) for (j = 0; j <= ny + 1; j++) u[IU(nx1, j, k)] = u[IU(1, j, k)];
    }
    if (s->ccxf == 1) for (k = 0; k <= nz + 1; k++) for (j = 0; j <= ny + 1; j++) u[IU(nx1, j, k)] = 0. + 0.;
    if (s->ccxf == 2) {
      for (k = 0; k <= nz + 1; k++) for (j = 0; j <= ny + 1; j++) {
        u[IU(nx1, j, k)] = 0. + 0.;
        u[IU(nx1 + 1, j, k)] = u[IU(nx1 - 1, j, k)];
      }
      for (k = 0; k <= nz + 1; k++) for (j = 0; j <= ny1 + 1; j++) v[IV(nx + 1, j, k)] = -v[IV(nx, j, k)];
      for (k = 0; k <= nz1 + 1; k++) for (j = 0; j <= ny + 1; j++) w[IW(nx + 1, j, k)] = -w[IW(nx, j, k)];
    }
    if (s->ccxf == 3) for (k = 0; k <= nz + 1; k++) for (j = 0; j <= ny + 1; j++) u[IU(nx1, j, k)] = BP(BXXF, j, k, ny + 2) + 0.;
    if (s->ccxf == 4) for (k = 0; k <= nz + 1; k++) for (j = 0; j <= ny + 1; j++) u[IU(nx1, j, k)] = u[IU(nx1 - 1, j, k)];
    if (s->ccz0 == 1 || s->ccz0 == 2) for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx + 1; i++) w[IW(i, j, 1)] = 0. + 0.;
    if (s->ccz0 == 3) for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx + 1; i++) w[IW(i, j, 1)] = BP(BZZ1, i, j, nx + 2) + 0.;
    if (s->cczf == 1 || s->cczf == 2) for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx + 1; i++) w[IW(i, j, nz1)] = 0. + 0.;
    if (s->cczf == 3) for (j = 0; j <= ny + 1; j++) for (i = 0; i <= nx + 1; i++) w[IW(i, j, nz1)] = BP(BZZF, i, j, nx + 2) + 0.;
  }
  if (nlock == 3 && s->obst_t != 0) {
    /* level set inside the obstacle from its neighbours, :445-490; every
     * value read lies outside the obstacle, so the loop order is immaterial */
    double *ls = s->ls;
    for (j = 1; j <= ny; j++)
      for (i = 1; i <= nx; i++)
        for (k = 1; k <= KW(i, j) - 1; k++) {
          if (KW(i - 1, j) < k && KW(i, j - 1) < k) ls[IC(i, j, k)] = (ls[IC(i - 1, j, k)] + ls[IC(i, j - 1, k)]) * 0.5;
          else if (KW(i + 1, j) < k && KW(i, j - 1) < k) ls[IC(i, j, k)] = (ls[IC(i + 1, j, k)] + ls[IC(i, j - 1, k)]) * 0.5;
          else if (KW(i + 1, j) < k && KW(i, j + 1) < k) ls[IC(i, j, k)] = (ls[IC(i + 1, j, k)] + ls[IC(i, j + 1, k)]) * 0.5;
          else if (KW(i - 1, j) < k && KW(i, j + 1) < k) ls[IC(i, j, k)] = (ls[IC(i - 1, j, k)] + ls[IC(i, j + 1, k)]) * 0.5;
          else if (KW(i - 1, j) < k && KW(i, j - 1) >= k && KW(i, j + 1) >= k) ls[IC(i, j, k)] = ls[IC(i - 1, j, k)];
          else if (KW(i, j - 1) < k && KW(i - 1, j) >= k && KW(i + 1, j) >= k) ls[IC(i, j, k)] = ls[IC(i, j - 1, k)];
          else if (KW(i + 1, j) < k && KW(i, j - 1) >= k && KW(i, j + 1) >= k) ls[IC(i, j, k)] = ls[IC(i + 1, j, k)];
          else if (KW(i, j + 1) < k && KW(i - 1, j) >= k && KW(i + 1, j) >= k) ls[IC(i, j, k)] = ls[IC(i, j + 1, k)];
          else if (k == KW(i, j) - 1 && k < nz) ls[IC(i, j, k)] = ls[IC(i, j, k + 1)];
        }
    or_heaviside(s);
  }
#undef KU
#undef KV
#undef KW
}

/* graddin, 9_grad.f90:10-278: symmetrically scaled CG on the
 * variable-density pressure equation, warm-started from prd1, stopped when
 * the sum of squared scaled residuals drops to 1e-4/(nx*ny*nz) or after
 * 10000 iterations (or after cg_max_it, when set: a fixed-iteration run for
 * comparisons where the reference operator is not symmetric and its CG does
 * not stop). */
void or_graddin(or_state *s) {
  DIMS;
  const int nx1 = nx + 1, ny1 = ny + 1, nz1 = nz + 1;
  const double dt = s->dt, dx = s->dx, dy = s->dy, dz = s->dz;
  const int64_t FX = (int64_t)nx1 * ny * nz, FY = (int64_t)nx * ny1 * nz, FZ = (int64_t)nx * ny * nz1;
  const int64_t C = (int64_t)nx * ny * nz, G = (int64_t)(nx + 2) * (ny + 2) * (nz + 2);
  double *mi = alloc_n(FX), *mj = alloc_n(FY), *mk = alloc_n(FZ);
  double *rhox = alloc_n(FX), *rhoy = alloc_n(FY), *rhoz = alloc_n(FZ);
  double *q = alloc_n(C), *aip = alloc_n(C), *ain = alloc_n(C), *ajp = alloc_n(C), *ajn = alloc_n(C),
         *akp = alloc_n(C), *akn = alloc_n(C);
  double *d = alloc_n(G), *e = alloc_n(G), *r = alloc_n(G), *pp = alloc_n(G), *mp = alloc_n(G);
  for (int64_t c = 0; c < FX; c++) mi[c] = dt / (dx * dx);
  for (int64_t c = 0; c < FY; c++) mj[c] = dt / (dy * dy);
  for (int64_t c = 0; c < FZ; c++) mk[c] = dt / (dz * dz);
  interpx_cf(s->rho, nx, ny, nz, rhox);
  interpy_cf(s->rho, nx, ny, nz, rhoy);
  interpz_cf(s->rho, nx, ny, nz, rhoz);
  for (int64_t c = 0; c < FX; c++) mi[c] = mi[c] / rhox[c];
  for (int64_t c = 0; c < FY; c++) mj[c] = mj[c] / rhoy[c];
  for (int64_t c = 0; c < FZ; c++) mk[c] = mk[c] / rhoz[c];
  const double *u = s->u, *v = s->v, *w = s->w;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        d[IP(i, j, k)] = mi[IFX(i + 1, j, k)] + mi[IFX(i, j, k)] + mj[IFY(i, j + 1, k)] + mj[IFY(i, j, k)] +
                         mk[IFZ(i, j, k + 1)] + mk[IFZ(i, j, k)];
        q[IC(i, j, k)] = (u[IU(i, j, k)] - u[IU(i + 1, j, k)]) / dx + (v[IV(i, j, k)] - v[IV(i, j + 1, k)]) / dy +
                         (w[IW(i, j, k)] - w[IW(i, j, k + 1)]) / dz;
      }
  /* mirror a ghosted vector (0:nx1,0:ny1,0:nz1), :98-116 */
#define MIRROR(a)                                                                              \
  do {                                                                                         \
    for (int k_ = 0; k_ <= nz1; k_++)                                                          \
      for (int j_ = 0; j_ <= ny1; j_++) {                                                      \
        if (s->ccx0 == 0) { a[IP(0, j_, k_)] = a[IP(nx, j_, k_)]; a[IP(nx1, j_, k_)] = a[IP(1, j_, k_)]; } \
        else { a[IP(0, j_, k_)] = a[IP(1, j_, k_)]; a[IP(nx1, j_, k_)] = a[IP(nx, j_, k_)]; }  \
      }                                                                                        \
    for (int k_ = 0; k_ <= nz1; k_++)                                                          \
      for (int i_ = 0; i_ <= nx1; i_++) {                                                      \
        if (s->ccy0 == 0) { a[IP(i_, 0, k_)] = a[IP(i_, ny, k_)]; a[IP(i_, ny1, k_)] = a[IP(i_, 1, k_)]; } \
        else { a[IP(i_, 0, k_)] = a[IP(i_, 1, k_)]; a[IP(i_, ny1, k_)] = a[IP(i_, ny, k_)]; }  \
      }                                                                                        \
    for (int j_ = 0; j_ <= ny1; j_++)                                                          \
      for (int i_ = 0; i_ <= nx1; i_++) { a[IP(i_, j_, 0)] = a[IP(i_, j_, 1)]; a[IP(i_, j_, nz1)] = a[IP(i_, j_, nz)]; } \
  } while (0)
  MIRROR(d);
  for (int64_t c = 0; c < G; c++) e[c] = s->prd1[c] * sqrt(d[c]);
  double alfamupr = 0.;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        int64_t c = IC(i, j, k), g = IP(i, j, k);
        aip[c] = mi[IFX(i + 1, j, k)] / sqrt(d[g] * d[IP(i + 1, j, k)]);
        ain[c] = mi[IFX(i, j, k)] / sqrt(d[g] * d[IP(i - 1, j, k)]);
        ajp[c] = mj[IFY(i, j + 1, k)] / sqrt(d[g] * d[IP(i, j + 1, k)]);
        ajn[c] = mj[IFY(i, j, k)] / sqrt(d[g] * d[IP(i, j - 1, k)]);
        akp[c] = mk[IFZ(i, j, k + 1)] / sqrt(d[g] * d[IP(i, j, k + 1)]);
        akn[c] = mk[IFZ(i, j, k)] / sqrt(d[g] * d[IP(i, j, k - 1)]);
        r[g] = e[g] - aip[c] * e[IP(i + 1, j, k)] - ain[c] * e[IP(i - 1, j, k)] - ajp[c] * e[IP(i, j + 1, k)] -
               ajn[c] * e[IP(i, j - 1, k)] - akp[c] * e[IP(i, j, k + 1)] - akn[c] * e[IP(i, j, k - 1)] -
               q[c] / sqrt(d[g]);
        alfamupr = alfamupr + r[g] * r[g];
      }
  MIRROR(r);
  memcpy(pp, r, sizeof(double) * (size_t)G);
  int cont = 0;
  const double thr = 0.0001 / (nx * ny * nz);
  const int cap = s->cg_max_it > 0 ? s->cg_max_it : 10000;
  while (fabs(alfamupr) > thr && cont < cap) {
    cont = cont + 1;
    double alfadipr = 0., betamupr = 0.;
    alfamupr = 0.;
    memset(mp, 0, sizeof(double) * (size_t)G);
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          int64_t c = IC(i, j, k), g = IP(i, j, k);
          mp[g] = pp[g] - pp[IP(i + 1, j, k)] * aip[c] - pp[IP(i - 1, j, k)] * ain[c] - pp[IP(i, j + 1, k)] * ajp[c] -
                  pp[IP(i, j - 1, k)] * ajn[c] - pp[IP(i, j, k + 1)] * akp[c] - pp[IP(i, j, k - 1)] * akn[c];
          alfamupr = alfamupr + r[g] * r[g];
          alfadipr = alfadipr + pp[g] * mp[g];
        }
    double alfapr = alfamupr / alfadipr;
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          int64_t g = IP(i, j, k);
          e[g] = e[g] - alfapr * pp[g];
          r[g] = r[g] - alfapr * mp[g];
          betamupr = betamupr + r[g] * r[g];
        }
    double betapr = betamupr / alfamupr;
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          int64_t g = IP(i, j, k);
          pp[g] = r[g] + betapr * pp[g];
        }
    MIRROR(e);
    MIRROR(pp);
  }
  s->cg_iters = cont;
  s->cg_res2 = alfamupr;
  for (int64_t c = 0; c < G; c++) s->prd1[c] = e[c] / sqrt(d[c]);
#undef MIRROR
  free(mi); free(mj); free(mk); free(rhox); free(rhoy); free(rhoz);
  free(q); free(aip); free(ain); free(ajp); free(ajn); free(akp); free(akn);
  free(d); free(e); free(r); free(pp); free(mp);
}

/* posdin, 10_pos.f90:1-96 */
void or_posdin(or_state *s) {
  DIMS;
  const int nx1 = nx + 1, ny1 = ny + 1, nz1 = nz + 1;
  const double dt = s->dt;
  const int64_t FX = (int64_t)nx1 * ny * nz, FY = (int64_t)nx * ny1 * nz, FZ = (int64_t)nx * ny * nz1;
  double *rhox = alloc_n(FX), *rhoy = alloc_n(FY), *rhoz = alloc_n(FZ);
  interpx_cf(s->rho, nx, ny, nz, rhox);
  interpy_cf(s->rho, nx, ny, nz, rhoy);
  interpz_cf(s->rho, nx, ny, nz, rhoz);
  const double aux1 = dt / s->dx, aux2 = dt / s->dy;
  double *p = s->prd1;
  for (int k = 1; k <= nz; k++) {
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx1; i++)
        s->u[IU(i, j, k)] = s->u[IU(i, j, k)] - aux1 * (p[IP(i, j, k)] - p[IP(i - 1, j, k)]) / rhox[IFX(i, j, k)];
    for (int j = 1; j <= ny1; j++)
      for (int i = 1; i <= nx; i++)
        s->v[IV(i, j, k)] = s->v[IV(i, j, k)] - aux2 * (p[IP(i, j, k)] - p[IP(i, j - 1, k)]) / rhoy[IFY(i, j, k)];
  }
  for (int k = 1; k <= nz1; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        s->w[IW(i, j, k)] = s->w[IW(i, j, k)] - dt / s->dz * (p[IP(i, j, k)] - p[IP(i, j, k - 1)]) / rhoz[IFZ(i, j, k)];
  free(rhox); free(rhoy); free(rhoz);
}

/* est, 4_plot.f90:500-525: max abs divergence */
double or_est_div(const or_state *s) {
  DIMS;
  double m = 0.;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        double dv = (s->u[IU(i + 1, j, k)] - s->u[IU(i, j, k)]) / s->dx + (s->v[IV(i, j + 1, k)] - s->v[IV(i, j, k)]) / s->dy +
                    (s->w[IW(i, j, k + 1)] - s->w[IW(i, j, k)]) / s->dz;
        if (fabs(dv) > m) m = fabs(dv);
      }
  return m;
}

/* one RK2 stage of 1_principal.f90:77-95 */
void or_stage(or_state *s, int tt, double dt_base) {
  s->tt = tt;
  s->dt = or_a_dt(s->t_tempo, tt, dt_base);
  or_visco(s);
  or_convdiff(s);
  or_tempo(s);
  if (s->wave_t > 0) or_boundary_waves(s); /* 1_principal.f90:84 */
  or_contorno(s, 2);
  or_graddin(s);
  or_posdin(s);
  or_contorno(s, 1);
}

/* waves_coef, 3_contorno.f90:1070-1205: wavelength, wave number and angular
 * frequency of the linear dispersion relation by 1000 fixed-point passes
 * (Stokes I, II); for Stokes V (Fenton 1985) the wavelength is the one of
 * 10000 candidates l0/2 + l0 (ii/10000) 2.5 whose two estimates of the
 * steepness parameter (10000 fixed-point passes each) agree best, then the
 * fifth-order coefficients.  Real exponents are pow() as the reference's
 * x**n.; integer ones are products. */
static double p_(double x, double e) { return pow(x, e); }
static void stokes5_b(double c, double s, double *b33, double *b35, double *b55, double *c1, double *c2) {
  const double c2_ = c * c;
  const double c4 = c2_ * c2_, c8 = c4 * c4, c12 = c8 * c4;  /* c**12: integer power */
  *b33 = 3. * (8. * p_(c, 6.) + 1.) / (64. * p_(s, 6.));
  *b35 = (88128. * p_(c, 14.) - 208224. * c12 + 70848. * p_(c, 10.) + 54000. * p_(c, 8.) - 21816. * p_(c, 6.) +
          6264. * p_(c, 4.) - 54. * c * c - 81.) /
         (12288. * p_(s, 12.) * (6. * c * c - 1.));
  *b55 = (192000. * p_(c, 16.) - 262720. * p_(c, 14.) + 83680. * c12 + 20160. * p_(c, 10.) - 7280. * p_(c, 8.) +
          7160. * p_(c, 6.) - 1800. * p_(c, 4.) - 1050. * c * c + 225.) /
         (12288 * p_(s, 10.) * (6. * c * c - 1.) * (8. * p_(c, 4.) - 11. * c * c + 3.));
  *c1 = (8. * p_(c, 4.) - 8. * c * c + 9.) / (8. * p_(s, 4.));
  *c2 = (3840. * p_(c, 12.) - 4096. * p_(c, 10.) + 2592. * p_(c, 8.) - 1008. * p_(c, 6.) + 5944. * p_(c, 4.) -
         1830. * c * c + 147.) /
        (512. * p_(s, 10.) * (6. * c * c - 1.));
}
void or_waves_coef(or_state *s) {
  const double pi = M_PI, gz = s->gz, a_w = s->wave_a, p_w = s->wave_p, h0_f = s->wave_h0;
  const double l0_w = gz * p_w * p_w / (2. * pi);
  double l_w = 0., lamb = 0.;
  double *wc = s->wc;
  for (int q = 0; q < 14; q++) wc[q] = 0.;
  if (s->wave_t <= 2) {
    l_w = gz * p_w * p_w / (2. * pi);
    for (int i = 1; i <= 1000; i++) l_w = gz * p_w * p_w / (2. * pi) * tanh(2. * pi * h0_f / l_w);
  } else {
    double erro0 = 999.;
    const int nxii = 10000;
    for (int ii = 1; ii <= nxii; ii++) {
      double lamb1 = 0., lamb2 = 0.;
      const double l_w1 = l0_w * 0.5 + l0_w * ((double)ii / nxii) * 2.5;
      const double sh = sinh(2. * pi * h0_f / l_w1), ch = cosh(2. * pi * h0_f / l_w1);
      double b33, b35, b55, c1, c2;
      stokes5_b(ch, sh, &b33, &b35, &b55, &c1, &c2);
      for (int i = 1; i <= 10000; i++) lamb1 = pi * a_w * 2. / l_w1 - p_(lamb1, 3.) * b33 - p_(lamb1, 5.) * (b35 + b55);
      for (int i = 1; i <= 10000; i++)
        lamb2 = sqrt((l_w1 / (l0_w * tanh(2. * pi * h0_f / l_w1)) - 1.) / (c1 + p_(lamb2, 2.) * c2));
      const double erro = fabs(lamb1 - lamb2);
      if (erro < erro0) {
        erro0 = erro;
        lamb = lamb1;
        l_w = l_w1;
      }
    }
  }
  const double c_w = l_w / p_w, f_w = 2. * pi / p_w, n_w = 2. * pi / l_w;
  wc[0] = l_w; wc[1] = n_w; wc[2] = f_w; wc[3] = c_w;
  if (s->wave_t == 5) {
    const double sh = sinh(2. * pi * h0_f / l_w), c = cosh(2. * pi * h0_f / l_w);
    double b33, b35, b55, c1, c2;
    stokes5_b(c, sh, &b33, &b35, &b55, &c1, &c2);
    const double a11 = 1. / sh;
    const double a13 = -c * c * (5. * c * c + 1.) / (8. * p_(sh, 5.));
    const double a15 = -(1184. * p_(c, 10.) - 1440. * p_(c, 8.) - 1992. * p_(c, 6.) + 2641. * p_(c, 4.) -
                         249. * p_(c, 2.) + 18.) / (1536. * p_(sh, 11.));
    const double a22 = 3. / (8. * p_(sh, 4.));
    const double a24 = (192. * p_(c, 8.) - 424. * p_(c, 6.) - 312. * p_(c, 4.) + 480. * c * c - 17.) / (768. * p_(sh, 10.));
    const double a33 = (13. - 4. * c * c) / (64. * p_(sh, 7.));
    const double a35 = (512. * p_(c, 12.) + 4224. * p_(c, 10.) - 6800. * p_(c, 8.) - 12808. * p_(c, 6.) +
                        16704. * p_(c, 4.) - 3154. * c * c + 107.) / (4096. * p_(sh, 13.) * (6. * c * c - 1.));
    const double a44 = (80. * p_(c, 6.) - 816. * p_(c, 4.) + 1338. * c * c - 197.) /
                       (1536. * p_(sh, 10.) * (6. * c * c - 1.));
    const double a55 = -(2880. * p_(c, 10.) - 72480. * p_(c, 8.) + 324000. * p_(c, 6.) - 432000. * p_(c, 4.) +
                         163470. * c * c - 16245.) /
                       (61440. * p_(sh, 11.) * (6. * c * c - 1.) * (8. * p_(c, 4.) - 11. * c * c + 3.));
    const double b22 = c * (2. * c * c + 1.) / (4. * p_(sh, 3.));
    const double b24 = c * (272. * p_(c, 8.) - 504. * p_(c, 6.) - 192. * p_(c, 4.) + 322 * c * c + 21.) /
                       (384. * p_(sh, 9.));
    const double b44 = c * (768. * p_(c, 10.) - 448. * p_(c, 8.) - 48. * p_(c, 6.) + 48. * p_(c, 4.) + 106. * c * c -
                            21.) / (384. * p_(sh, 9.) * (6. * c * c - 1.));
    wc[4] = lamb / n_w;
    wc[5] = (lamb * lamb * b22 + p_(lamb, 4.) * b24) / n_w;
    wc[6] = (p_(lamb, 3.) * b33 + p_(lamb, 5.) * b35) / n_w;
    wc[7] = p_(lamb, 4.) * b44 / n_w;
    wc[8] = p_(lamb, 5.) * b55 / n_w;
    wc[9] = 1. * 2. * pi / (p_w * n_w) * (lamb * a11 + p_(lamb, 3.) * a13 + p_(lamb, 5.) * a15);
    wc[10] = 2. * 2. * pi / (p_w * n_w) * (p_(lamb, 2.) * a22 + p_(lamb, 4.) * a24);
    wc[11] = 3. * 2. * pi / (p_w * n_w) * (p_(lamb, 3.) * a33 + p_(lamb, 5.) * a35);
    wc[12] = 4. * 2. * pi / (p_w * n_w) * (p_(lamb, 4.) * a44);
    wc[13] = 5. * 2. * pi / (p_w * n_w) * (p_(lamb, 5.) * a55);
  }
}

/* boundary_waves, 3_contorno.f90:952-1068: the free surface h_f and the
 * velocities u, w of the wave at i = 0, 1 (x = (i - 1/2) dx for the surface
 * and w, (i - 1) dx for u) at the time t, zero above the surface; the level
 * set of the first cell column becomes h_f(1) - z, and the whole level set
 * passes through real(4) (the reference's ls1 is single precision).  The
 * planes bxx0, bxx1, bxz0 take u(0), u(1), w(0); bxy0 takes v1, which the
 * reference never sets (zero here, as the flang-built reference leaves it). */
void or_boundary_waves(or_state *s) {
  DIMS;
  const double *wc = s->wc;
  const double l_w = wc[0], n_w = wc[1], f_w = wc[2];
  const double a_w = s->wave_a, h0_f = s->wave_h0, t = s->t, dx = s->dx, dz = s->dz;
  const int nx1 = nx + 1, nz1 = nz + 1;
  (void)l_w;
  double *ls = s->ls;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) ls[IC(i, j, k)] = (double)(float)ls[IC(i, j, k)];
  const double aux1 = tanh(n_w * h0_f);
  for (int k = 0; k <= nz + 1; k++) {
    const double kp = (k - 0.5) * dz;
    for (int j = 0; j <= ny + 1; j++)
      for (int i = 0; i <= 1; i++) {
        const double aux2 = n_w * dx * (i - 0.5) - f_w * t, aux3 = n_w * dx * (i - 1.) - f_w * t;
        double h_f, u1 = 0., w1 = 0.;
        if (s->wave_t == 1) h_f = h0_f + a_w * cos(n_w * dx * (i - 0.5) - f_w * t);
        else if (s->wave_t == 2)
          h_f = h0_f + a_w * cos(aux2) + n_w * a_w * a_w * (3. - aux1 * aux1) / (4. * aux1 * aux1 * aux1) * cos(2. * aux2);
        else
          h_f = h0_f + wc[4] * cos(aux2) + wc[5] * cos(2. * aux2) + wc[6] * cos(3. * aux2) + wc[7] * cos(4. * aux2) +
                wc[8] * cos(5. * aux2);
        const float ls1 = (float)(h_f - kp);
        if (ls1 >= 0.f) {
          if (s->wave_t == 1) {
            u1 = a_w * f_w * cosh(n_w * kp) / sinh(n_w * h0_f) * cos(n_w * dx * (i - 1.) - f_w * t);
            w1 = a_w * f_w * sinh(n_w * (kp - 0.5 * dz)) / sinh(n_w * h0_f) * sin(n_w * dx * (i - 0.5) - f_w * t);
          } else if (s->wave_t == 2) {
            u1 = a_w * f_w * cosh(n_w * kp) / sinh(n_w * h0_f) * cos(aux3) +
                 3. / 4. * a_w * a_w * f_w * n_w * cosh(2. * n_w * kp) / pow(sinh(n_w * h0_f), 4.) * cos(2 * aux3);
            w1 = a_w * f_w * sinh(n_w * (kp - 0.5 * dz)) / sinh(n_w * h0_f) * sin(aux2) +
                 3. / 4. * a_w * a_w * f_w * n_w * sinh(2. * n_w * (kp - 0.5 * dz)) / pow(sinh(n_w * h0_f), 4.) *
                     sin(2. * aux3);
          } else {
            const double aux4 = n_w * kp, aux5 = n_w * (kp - 0.5 * dz);
            u1 = wc[9] * cosh(aux4) * cos(aux3) + wc[10] * cosh(2 * aux4) * cos(2 * aux3) +
                 wc[11] * cosh(3 * aux4) * cos(3 * aux3) + wc[12] * cosh(4 * aux4) * cos(4 * aux3) +
                 wc[13] * cosh(5 * aux4) * cos(5 * aux3);
            w1 = wc[9] * sinh(aux5) * sin(aux2) + wc[10] * sinh(2 * aux5) * sin(2 * aux2) +
                 wc[11] * sinh(3 * aux5) * sin(3 * aux2) + wc[12] * sinh(4 * aux5) * sin(4 * aux2) +
                 wc[13] * sinh(5 * aux5) * sin(5 * aux2);
          }
        }
        if (i == 1 && j >= 1 && j <= ny && k >= 1 && k <= nz) ls[IC(1, j, k)] = (double)ls1;
        if (i == 0) {
          BP(BXX0, j, k, ny + 2) = u1;
          BP(BXZ0, j, k, ny + 2) = w1;
        } else {
          BP(BXX1, j, k, ny + 2) = u1;
        }
      }
  }
  for (int j = 0; j <= ny + 1; j++) BP(BXZ0, j, nz1 + 1, ny + 2) = 0.;
  for (int k = 0; k <= nz + 1; k++)
    for (int j = 0; j <= ny + 2; j++) BP(BXY0, j, k, ny + 3) = 0.;
  (void)nx1;
}

/* one stage of the manufactured-solution mode (mms_t > 0): no projection,
 * 1_principal.f90:76-94 */
void or_stage_noproj(or_state *s, int tt, double dt_base) {
  s->tt = tt;
  s->dt = or_a_dt(s->t_tempo, tt, dt_base);
  or_visco(s);
  or_convdiff(s);
  or_tempo(s);
  or_contorno(s, 2);
}

/* one time step, 1_principal.f90:74-95 (level set with the full dt) */
void or_step(or_state *s, double dt_base) {
  s->dt = dt_base;
  or_level_set(s);
  or_contorno(s, 3);
  for (int tt = 1; tt <= or_ntt(s->t_tempo); tt++) or_stage(s, tt, dt_base);
  s->dt = dt_base;
}
