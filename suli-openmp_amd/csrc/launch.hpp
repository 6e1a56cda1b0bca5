// Host-side launchers of the templated stencil kernels.  Each is defined in
// stencils_impl.hpp and instantiated for double (inst_f64.hip, built without
// FMA contraction) and for float (inst_f32.hip).
#pragma once
#include "common.hpp"

struct MomPtrs {
  const void *u, *v, *w, *u0, *v0, *w0;
  const void *rho, *lsnu, *nut, *hsx, *hsy, *hsz, *kurv;
  const void *tfu, *tfv, *tfw;
  void *un, *vn, *wn, *u0o, *v0o, *w0o;
  void *Fu, *Fv, *Fw;
  void *fu0, *fv0, *fw0, *fu1, *fv1, *fw1;
  const void *xnf, *ynf, *znf;  // face viscosities of the directional model, else null
  int adv;                      // 10*adv_type + der (11, 12, 13, 21, 31)
  const int *ku, *kv, *kw;      // bottom-obstacle columns (plane layout), else null
};

// outputs of the FCG face set-up (k_fcg_faces)
struct FcgFaceOut {
  double *ax, *ay, *az;
  float *ax32, *ay32, *az32;
  uint16_t *hx, *hy, *hz;
};

template <typename T>
struct Launch {
  static void ls_adv(const Grid &g, hipStream_t s, const void *phi, const void *phin, const void *u,
                     const void *v, const void *w, void *out, double ca, double cb, double cg, double dt);
  static void reinit_sign(const Grid &g, const Phys &ph, hipStream_t s, const void *phi, void *S);
  static void reinit_stage(const Grid &g, const Phys &ph, hipStream_t s, const void *phi,
                           const void *phit, const void *S, void *out, double ca, double cb,
                           double cgdt, int band, RedSlot rs);
  // reinit_sign fused with the first stage of the first reinitialisation
  // iteration (ca = 1, cb = 0, phit = phi): S and that stage's output
  static void reinit_first(const Grid &g, const Phys &ph, hipStream_t s, const void *phi, void *S, void *out,
                           double cgdt);
  static void heavi(const Grid &g, const Phys &ph, hipStream_t s, const void *phi, void *hs, void *hsx,
                    void *hsy, void *hsz, void *rho, void *lsnu, void *modls, void *nrx, void *nry,
                    void *nrz, RedSlot rs);
  static void curv(const Grid &g, hipStream_t s, const void *nrx, const void *nry, const void *nrz,
                   void *ddx, void *ddy, void *ddz, void *kurv);
  static void masscorr(const Grid &g, hipStream_t s, void *phi, const void *modls, double f, double vol0);
  static void visco(const Grid &g, const Phys &ph, hipStream_t s, const void *u, const void *v,
                    const void *w, const void *rho, void *nut);
  // mode: MomUpdate (stencils.hpp)
  static void momentum(const Grid &g, const Phys &ph, hipStream_t s, const MomPtrs &m, double dt,
                       int mode);
  // bp: the boundary planes of wall code 3 (layout bplanes_layout; read only
  // where a wall has code 3, or 4 at x = 0)
  static void contorno(const Grid &g, hipStream_t s, int nlock, void *u, void *v, void *w, double *p,
                       const void *bp, const int *ku = nullptr, const int *kv = nullptr,
                       const int *kw = nullptr);
  static void obst_ls(const Grid &g, hipStream_t s, void *ls, const int *kw);
  // directional Smagorinsky face viscosities (m_turb = 2)
  static void visco_dir(const Grid &g, const Phys &ph, hipStream_t s, const void *u, const void *v, const void *w,
                        const void *lsnu, void *xn, void *yn, void *zn);
  static void posdin(const Grid &g, hipStream_t s, double dt, const double *p, const void *rho,
                     void *u, void *v, void *w);
  static void divmax(const Grid &g, hipStream_t s, const void *u, const void *v, const void *w,
                     RedSlot rs);
  static void dambreak(const Grid &g, hipStream_t s, double width, double height, void *phi);
  static void cast_from_f64(hipStream_t s, const double *src, void *dst, int64_t n);
  static void cast_to_f64(hipStream_t s, const void *src, double *dst, int64_t n);
  // projection set-up shared by the Poisson solvers
  // (ghosts of D and sqrt(D) mirrored unless mirror is false)
  static void pD(const Grid &g, hipStream_t s, double dt, const void *u, const void *v, const void *w,
                 const void *rho, double *D, double *sD, double *q, bool mirror = true);
  static void pcoef(const Grid &g, hipStream_t s, double dt, const void *rho, const double *D,
                    double *ax, double *ay, double *az);
  // the face arrays of the FCG solve (k_fcg_faces, poisson_ref.hpp): pcoef's
  // coefficients with their FP32 copies and the BF16 multigrid conductances
  static void fcg_faces(const Grid &g, hipStream_t s, double dt, const void *rho, const double *D,
                        const FcgFaceOut &o);
  // the record of a field-output file (k_frame_record)
  static void frame_record(const Grid &g, hipStream_t s, int t_plot, const void *u, const void *v, const void *w,
                           const void *ls, const double *p, const void *nut, const int *kw, const float *xs,
                           const float *ys, const float *zs, float *out);
};

// grid helpers
inline dim3 cell_grid(const Grid &g) {
  return dim3((g.nx + SULI_BX - 1) / SULI_BX, (g.ny + SULI_BY - 1) / SULI_BY, g.nz);
}
inline dim3 blk2() { return dim3(SULI_BX, SULI_BY, 1); }
