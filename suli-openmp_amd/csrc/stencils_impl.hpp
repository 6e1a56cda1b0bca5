// Definitions of the Launch<T> members; included once per precision.
#pragma once
#include "launch.hpp"
#include "poisson_ref.hpp"
#include "stencils.hpp"

#define SULI_LAUNCH(kernel, grid, block, stream, ...) \
  hipLaunchKernelGGL(kernel, grid, block, 0, stream, __VA_ARGS__)

template <typename T>
void Launch<T>::ls_adv(const Grid &g, hipStream_t s, const void *phi, const void *phin, const void *u,
                       const void *v, const void *w, void *out, double ca, double cb, double cg,
                       double dt) {
  SULI_LAUNCH(k_ls_adv<T>, cell_grid(g), blk2(), s, g, (const T *)phi, (const T *)phin, (const T *)u,
              (const T *)v, (const T *)w, (T *)out, (T)ca, (T)cb, (T)cg, (T)dt);
}

template <typename T>
void Launch<T>::reinit_sign(const Grid &g, const Phys &ph, hipStream_t s, const void *phi, void *S) {
  SULI_LAUNCH(k_reinit_sign<T>, cell_grid(g), blk2(), s, g, (T)ph.dx1, (const T *)phi, (T *)S);
}

template <typename T>
void Launch<T>::reinit_stage(const Grid &g, const Phys &ph, hipStream_t s, const void *phi,
                             const void *phit, const void *S, void *out, double ca, double cb,
                             double cgdt, int band, RedSlot rs) {
  const double bandw = 2. * ph.alpha1 * ph.dx1;
  SULI_LAUNCH(k_reinit_stage<T>, cell_grid(g), blk2(), s, g, (const T *)phi, (const T *)phit, (T *)S, (T *)out,
              (T)ca, (T)cb, (T)cgdt, band, bandw, rs, T(0));
  if (band) launch_finish<2>(s, cell_grid(g), rs);
}

template <typename T>
void Launch<T>::reinit_first(const Grid &g, const Phys &ph, hipStream_t s, const void *phi, void *S, void *out,
                             double cgdt) {
  SULI_LAUNCH((k_reinit_stage<T, true>), cell_grid(g), blk2(), s, g, (const T *)phi, (const T *)phi, (T *)S,
              (T *)out, T(1.), T(0.), (T)cgdt, 0, 0.0, RedSlot{nullptr, nullptr}, (T)ph.dx1);
}

template <typename T>
void Launch<T>::heavi(const Grid &g, const Phys &ph, hipStream_t s, const void *phi, void *hs, void *hsx,
                      void *hsy, void *hsz, void *rho, void *lsnu, void *modls, void *nrx, void *nry,
                      void *nrz, RedSlot rs) {
  SULI_LAUNCH(k_heavi<T>, cell_grid(g), blk2(), s, g, ph, (const T *)phi, (T *)hs, (T *)hsx, (T *)hsy,
              (T *)hsz, (T *)rho, (T *)lsnu, (T *)modls, (T *)nrx, (T *)nry, (T *)nrz, rs);
  launch_finish<1>(s, cell_grid(g), rs);
}

template <typename T>
void Launch<T>::curv(const Grid &g, hipStream_t s, const void *nrx, const void *nry, const void *nrz,
                     void *ddx, void *ddy, void *ddz, void *kurv) {
  SULI_LAUNCH(k_curv<T>, cell_grid(g), blk2(), s, g, (const T *)nrx, (const T *)nry, (const T *)nrz,
              (T *)ddx, (T *)ddy, (T *)ddz, (T *)kurv);
}

template <typename T>
void Launch<T>::masscorr(const Grid &g, hipStream_t s, void *phi, const void *modls, double f,
                         double vol0) {
  SULI_LAUNCH(k_masscorr<T>, cell_grid(g), blk2(), s, g, (T *)phi, (const T *)modls, (T)f, (T)vol0);
}

template <typename T>
void Launch<T>::visco(const Grid &g, const Phys &ph, hipStream_t s, const void *u, const void *v,
                      const void *w, const void *rho, void *nut) {
  SULI_LAUNCH(k_visco<T>, cell_grid(g), blk2(), s, g, (T)ph.c23, (T)ph.csmag, (const T *)u,
              (const T *)v, (const T *)w, (const T *)rho, (T *)nut);
}

template <typename T>
void Launch<T>::momentum(const Grid &g, const Phys &ph, hipStream_t s, const MomPtrs &m, double dt,
                         int mode) {
  MomArgs<T> a;
  a.u = (const T *)m.u; a.v = (const T *)m.v; a.w = (const T *)m.w;
  a.u0 = (const T *)m.u0; a.v0 = (const T *)m.v0; a.w0 = (const T *)m.w0;
  a.rho = (const T *)m.rho; a.lsnu = (const T *)m.lsnu; a.nut = (const T *)m.nut;
  a.hsx = (const T *)m.hsx; a.hsy = (const T *)m.hsy; a.hsz = (const T *)m.hsz;
  a.kurv = (const T *)m.kurv;
  a.tfu = (const T *)m.tfu; a.tfv = (const T *)m.tfv; a.tfw = (const T *)m.tfw;
  a.un = (T *)m.un; a.vn = (T *)m.vn; a.wn = (T *)m.wn;
  a.u0o = (T *)m.u0o; a.v0o = (T *)m.v0o; a.w0o = (T *)m.w0o;
  a.Fu = (T *)m.Fu; a.Fv = (T *)m.Fv; a.Fw = (T *)m.Fw;
  a.fu0 = (T *)m.fu0; a.fv0 = (T *)m.fv0; a.fw0 = (T *)m.fw0;
  a.fu1 = (T *)m.fu1; a.fv1 = (T *)m.fv1; a.fw1 = (T *)m.fw1;
  a.xnf = (const T *)m.xnf; a.ynf = (const T *)m.ynf; a.znf = (const T *)m.znf;
  a.ku = m.ku; a.kv = m.kv; a.kw = m.kw;
  a.dt = (T)dt;
  a.mode = mode;
  // the inner box (i in 1..nx, j and k in 2..n-1; grid gi) and the shell of
  // the frame around it (grid)
  const dim3 grid((unsigned)((mom_shell_cells(g) + 255) / 256), 3);
  const dim3 gi((g.nx + SULI_BX - 1) / SULI_BX, (g.ny - 2 + SULI_BY - 1) / SULI_BY, g.nz - 2);
  const bool faces = m.xnf != nullptr;  // directional model: stored face viscosities
  if (m.adv == 11 && !faces && g.fast) {
    // LDS tiles of 64 x MT_TJ cells walking chunks of MOM_KC planes (8: 3 % faster
    // than 16, 20 % faster than 32 at 256x128x128)
#ifndef SULI_MOM_KC
#define SULI_MOM_KC 8
#endif
    constexpr int MOM_KC = SULI_MOM_KC;
    const dim3 gt((g.nx + SULI_BX - 1) / SULI_BX, (g.ny - 2 + MT_TJ - 1) / MT_TJ,
                  (g.nz - 2 + MOM_KC - 1) / MOM_KC);
    const dim3 bt(SULI_BX, MT_TJ);
    if (mode == MOM_RK_FIRST && !a.fu0) k_momentum_tile<T, MOM_RK_FIRST><<<gt, bt, 0, s>>>(g, ph, a, MOM_KC);
    else if (mode == MOM_FROM_U0) k_momentum_tile<T, MOM_FROM_U0><<<gt, bt, 0, s>>>(g, ph, a, MOM_KC);
    else k_momentum_tile<T, 0><<<gt, bt, 0, s>>>(g, ph, a, MOM_KC);
    k_momentum<T, 11, false><<<grid, 256, 0, s>>>(g, ph, a);
    return;
  }
#define SULI_MOM(ADV)                                                                        \
  (faces ? ((void)(k_momentum_in<T, ADV, true><<<gi, blk2(), 0, s>>>(g, ph, a)),              \
            (void)(k_momentum<T, ADV, true><<<grid, 256, 0, s>>>(g, ph, a)))                  \
         : ((void)(k_momentum_in<T, ADV, false><<<gi, blk2(), 0, s>>>(g, ph, a)),             \
            (void)(k_momentum<T, ADV, false><<<grid, 256, 0, s>>>(g, ph, a))))
  switch (m.adv) {
    case 12: SULI_MOM(12); break;
    case 13: SULI_MOM(13); break;
    case 21: SULI_MOM(21); break;
    case 31: SULI_MOM(31); break;
    default: SULI_MOM(11); break;
  }
#undef SULI_MOM
}

template <typename T>
void Launch<T>::obst_ls(const Grid &g, hipStream_t s, void *ls, const int *kw) {
  SULI_LAUNCH(k_obst_ls<T>, dim3((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz), dim3(64, 4, 1), s, g, (T *)ls, kw);
}

template <typename T>
void Launch<T>::visco_dir(const Grid &g, const Phys &ph, hipStream_t s, const void *u, const void *v, const void *w,
                          const void *lsnu, void *xn, void *yn, void *zn) {
  dim3 grid((g.nx + 1 + SULI_BX - 1) / SULI_BX, (g.ny + 1 + SULI_BY - 1) / SULI_BY, g.nz + 1);
  SULI_LAUNCH(k_visco_dir<T>, grid, blk2(), s, g, (T)ph.csmag, (const T *)u, (const T *)v, (const T *)w,
              (const T *)lsnu, (T *)xn, (T *)yn, (T *)zn);
}

template <typename T>
void Launch<T>::contorno(const Grid &g, hipStream_t s, int nlock, void *u, void *v, void *w, double *p,
                         const void *bp, const int *ku, const int *kv, const int *kw) {
  if (nlock != 1 && nlock != 2) return;
  const dim3 b(64, 4, 1);
  auto gr = [&](int a, int c) { return dim3((a + 63) / 64, (c + 3) / 4, 1); };
  const BPlanes L = bplanes_layout(g.nx, g.ny, g.nz);
  const T *P = (const T *)bp;
  SULI_LAUNCH(k_bc_y<T>, gr(g.nx + 3, g.nz + 3), b, s, g, (T *)u, (T *)v, (T *)w, p, P, L);
  {
    dim3 gx = gr(g.ny + 3, g.nz + 3);
    gx.z = 8;  // field x side
    SULI_LAUNCH(k_bc_x<T>, gx, b, s, g, (T *)u, (T *)v, (T *)w, p, P, L);
  }
  SULI_LAUNCH(k_bc_z<T>, gr(g.nx + 3, g.ny + 3), b, s, g, (T *)u, (T *)v, (T *)w, p, P, L);
  if (ku) SULI_LAUNCH(k_bc_obst<T>, gr(g.nx + 3, g.ny + 3), b, s, g, (T *)u, (T *)v, (T *)w, ku, kv, kw);
  if (nlock == 2) {
    SULI_LAUNCH(k_bcn_y<T>, gr(g.nx + 2, g.nz + 2), b, s, g, (T *)v, P, L);
    SULI_LAUNCH(k_bcn_x<T>, gr(g.ny + 3, g.nz + 3), b, s, g, (T *)u, (T *)v, (T *)w, P, L);
    SULI_LAUNCH(k_bcn_z<T>, gr(g.nx + 2, g.ny + 2), b, s, g, (T *)w, P, L);
  }
}

template <typename T>
void Launch<T>::posdin(const Grid &g, hipStream_t s, double dt, const double *p, const void *rho, void *u,
                       void *v, void *w) {
  dim3 grid((g.nx + 1 + SULI_BX - 1) / SULI_BX, (g.ny + 1 + SULI_BY - 1) / SULI_BY, g.nz + 1);
  SULI_LAUNCH(k_posdin<T>, grid, blk2(), s, g, dt, p, (const T *)rho, (T *)u, (T *)v, (T *)w);
}

template <typename T>
void Launch<T>::divmax(const Grid &g, hipStream_t s, const void *u, const void *v, const void *w,
                       RedSlot rs) {
  SULI_LAUNCH(k_divmax<T>, cell_grid(g), blk2(), s, g, (const T *)u, (const T *)v, (const T *)w, rs);
  launch_finish<1, true>(s, cell_grid(g), rs);
}

// Dam-break level set: signed distance (positive in water) to the column
// x <= width, z <= height spanning the tank in y, at the cell centres
// ((i-0.5)dx, (k-0.5)dz) of 14_level_set.f90:53-63.  The column faces lying
// on the walls are not interfaces, so the distance is to x = width, z = height.
template <typename T>
__global__ void k_dambreak(Grid g, double width, double height, T *__restrict__ phi) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const double x = (i - 0.5) * g.dx, z = (k - 0.5) * g.dz;
  const double ax = x - width, az = z - height;
  double d;
  if (ax <= 0 && az <= 0) d = (-ax < -az) ? -ax : -az;
  else {
    const double px = ax > 0 ? ax : 0, pz = az > 0 ? az : 0;
    d = -sqrt(px * px + pz * pz);
  }
  phi[g.idx(i, j, k)] = (T)d;
}
template <typename T>
void Launch<T>::dambreak(const Grid &g, hipStream_t s, double width, double height, void *phi) {
  SULI_LAUNCH(k_dambreak<T>, cell_grid(g), blk2(), s, g, width, height, (T *)phi);
}

template <typename T>
__global__ void k_cast_from(const double *__restrict__ a, T *__restrict__ b, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    b[q] = (T)a[q];
}
template <typename T>
__global__ void k_cast_to(const T *__restrict__ a, double *__restrict__ b, int64_t n) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    b[q] = (double)a[q];
}
template <typename T>
void Launch<T>::cast_from_f64(hipStream_t s, const double *src, void *dst, int64_t n) {
  SULI_LAUNCH(k_cast_from<T>, dim3(2048), dim3(256), s, src, (T *)dst, n);
}
template <typename T>
void Launch<T>::cast_to_f64(hipStream_t s, const void *src, double *dst, int64_t n) {
  SULI_LAUNCH(k_cast_to<T>, dim3(2048), dim3(256), s, (const T *)src, dst, n);
}

template <typename T>
void Launch<T>::pD(const Grid &g, hipStream_t s, double dt, const void *u, const void *v, const void *w,
                   const void *rho, double *D, double *sD, double *q, bool mirror) {
  SULI_LAUNCH(k_pD<T>, cell_grid(g), blk2(), s, g, dt, (const T *)u, (const T *)v, (const T *)w,
              (const T *)rho, D, sD, q);
  if (!mirror) return;
  dim3 b(64, 4, 1);
  dim3 gr((g.nx + 2 + 63) / 64, (g.ny + 2 + 3) / 4, g.nz + 2);
  SULI_LAUNCH(k_mirror_ghosts<0>, gr, b, s, g, D, sD);
}

template <typename T>
void Launch<T>::fcg_faces(const Grid &g, hipStream_t s, double dt, const void *rho, const double *D,
                          const FcgFaceOut &o) {
  dim3 gf((g.nx + 1 + SULI_BX - 1) / SULI_BX, (g.ny + 1 + SULI_BY - 1) / SULI_BY, g.nz + 1);
  SULI_LAUNCH(k_fcg_faces<T>, gf, blk2(), s, g, dt, (const T *)rho, D, o);
}

template <typename T>
void Launch<T>::pcoef(const Grid &g, hipStream_t s, double dt, const void *rho, const double *D, double *ax,
                      double *ay, double *az) {
  dim3 grid((g.nx + 1 + SULI_BX - 1) / SULI_BX, (g.ny + 1 + SULI_BY - 1) / SULI_BY, g.nz + 1);
  SULI_LAUNCH(k_pcoef<T>, grid, blk2(), s, g, dt, (const T *)rho, D, ax, ay, az);
}

template <typename T>
void Launch<T>::frame_record(const Grid &g, hipStream_t s, int t_plot, const void *u, const void *v, const void *w,
                             const void *ls, const double *p, const void *nut, const int *kw, const float *xs,
                             const float *ys, const float *zs, float *out) {
  const dim3 grid((g.nx + 1 + SULI_BX - 1) / SULI_BX, (g.ny + 1 + SULI_BY - 1) / SULI_BY, g.nz + 1);
  SULI_LAUNCH(k_frame_record<T>, grid, blk2(), s, g, t_plot, (const T *)u, (const T *)v, (const T *)w,
              (const T *)ls, p, (const T *)nut, kw, xs, ys, zs, out);
}
