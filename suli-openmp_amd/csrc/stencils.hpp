// Stencil kernels of the SuLi-IC time step, templated on the field type T
// (double for the FP64 path, float for the mixed-precision path).  The two
// instantiations live in separate translation units: the FP64 one is built
// without FMA contraction so that it evaluates every expression in the
// reference's order.  Reference lines are cited as file:line in SuLi-IC/.
#pragma once
#include "common.hpp"

#define SULI_PI 3.141592653589793115997963468544185161590576171875

// ---------------------------------------------------------------------------
// Level set (14_level_set.f90)

// One SSP-RK3 stage of the level-set advection: conv_weno (:308-341) fused
// with intt_ls (:287-306).  out = ca*phin + cb*phi + (cg*dt)*L(phi).
template <typename T>
__global__ __launch_bounds__(256) void k_ls_adv(Grid g, const T *__restrict__ phi,
                                                const T *__restrict__ phin, const T *__restrict__ u,
                                                const T *__restrict__ v, const T *__restrict__ w,
                                                T *__restrict__ out, T ca, T cb, T cg, T dt) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const fidx c = g.idx(i, j, k);
  T ta, tb, tc, td, te, tf;
  der_weno_cell<T>(phi, g, i, j, k, c, false, ta, tb, tc, td, te, tf);
  const T a = (u[c] + u[c + 1]) * T(0.5);
  const T b = (v[c] + v[c + g.sx]) * T(0.5);
  const T cc = (w[c] + w[c + g.sz]) * T(0.5);
  const T apos = tmax(a, T(0)), aneg = tmin(a, T(0));
  const T bpos = tmax(b, T(0)), bneg = tmin(b, T(0));
  const T cpos = tmax(cc, T(0)), cneg = tmin(cc, T(0));
  const T sy7 = -(apos * td + aneg * ta + bpos * te + bneg * tb + cpos * tf + cneg * tc);
  out[c] = ca * phin[c] + cb * phi[c] + cg * dt * sy7;
}

// Godunov upwinding with the 1e-8 floors of reinic_weno (:406-430).
template <typename T>
__device__ __forceinline__ T godunov_mod(T lsv, T ta, T tb, T tc, T td, T te, T tf) {
  const T fl = T(0.00000001);
  T x, y, z;
  if (lsv > T(0)) {
    x = tmax(tmax(td, fl), -tmin(ta, fl));
    y = tmax(tmax(te, fl), -tmin(tb, fl));
    z = tmax(tmax(tf, fl), -tmin(tc, fl));
  } else {
    x = tmax(tmax(ta, fl), -tmin(td, fl));
    y = tmax(tmax(tb, fl), -tmin(te, fl));
    z = tmax(tmax(tc, fl), -tmin(tf, fl));
  }
  return sqrt(x * x + y * y + z * z);
}

// Sign function S(phi0) of the reinitialisation (:400-435).
template <typename T>
__global__ __launch_bounds__(256) void k_reinit_sign(Grid g, T dx1, const T *__restrict__ phi,
                                                     T *__restrict__ S) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const fidx c = g.idx(i, j, k);
  T ta, tb, tc, td, te, tf;
  der_weno_cell<T>(phi, g, i, j, k, c, false, ta, tb, tc, td, te, tf);
  const T lsv = phi[c];
  const T m = godunov_mod<T>(lsv, ta, tb, tc, td, te, tf);
  S[c] = lsv / sqrt(lsv * lsv + m * m * dx1 * dx1);
}

// One SSP-RK3 stage of the reinitialisation (:443-484).  On the last stage
// (band != 0) it also sums, in FP64, the change phi_new - phi_old and the
// number of cells in the band abs(phi_new) <= 2*alpha1*dx1 (:493-503).
// SIGN: the first stage of the first iteration, whose phi is phi0: the sign
// function of k_reinit_sign comes from the same derivatives and the same
// |grad phi|, so it is computed here and stored in S (dx1: as there).
template <typename T, bool SIGN = false>
__global__ __launch_bounds__(256) void k_reinit_stage(Grid g, const T *__restrict__ phi,
                                                      const T *__restrict__ phit,
                                                      T *__restrict__ S, T *__restrict__ out,
                                                      T ca, T cb, T cgdt, int band, double bandw,
                                                      RedSlot rs, T dx1 = T(0)) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  double acc[2] = {0.0, 0.0};
  if (i <= g.nx && j <= g.ny) {
    const fidx c = g.idx(i, j, k);
    T ta, tb, tc, td, te, tf;
    der_weno_cell<T>(phi, g, i, j, k, c, false, ta, tb, tc, td, te, tf);
    const T lsv = phi[c];
    const T m = godunov_mod<T>(lsv, ta, tb, tc, td, te, tf);
    T sv;
    if constexpr (SIGN) {
      sv = lsv / sqrt(lsv * lsv + m * m * dx1 * dx1);
      S[c] = sv;
    } else {
      sv = S[c];
    }
    const T sy7 = sv * (T(1.) - m);
    const T nv = ca * phit[c] + cb * lsv + cgdt * sy7;
    out[c] = nv;
    if (band) {
      const double nd = (double)nv;
      if ((nd < 0 ? -nd : nd) <= bandw) {
        acc[0] = nd - (double)lsv;
        acc[1] = 1.0;
      }
    }
  }
  if (band) {
    __shared__ double lds[2 * SULI_LDS_RED];
    block_reduce<2>(acc, lds);
    block_partial<2>(acc, rs.partials);
  }
}

// heaviside (:612-707) fused with the first half of mod_ls1 (:717-759) and the
// volume integral (:246-253): both take the larger-magnitude one-sided WENO
// derivative in each direction.  Writes H, grad H, rho, mu, |grad phi| and the
// unit normal; sums (1-(1-H))*dV in FP64.
template <typename T>
__global__ __launch_bounds__(256) void k_heavi(Grid g, Phys ph, const T *__restrict__ phi,
                                               T *__restrict__ hs, T *__restrict__ hsx,
                                               T *__restrict__ hsy, T *__restrict__ hsz,
                                               T *__restrict__ rho, T *__restrict__ lsnu,
                                               T *__restrict__ modls, T *__restrict__ nrx,
                                               T *__restrict__ nry, T *__restrict__ nrz,
                                               RedSlot rs) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  double acc[1] = {0.0};
  if (i <= g.nx && j <= g.ny) {
    const fidx c = g.idx(i, j, k);
    T ta, tb, tc, td, te, tf;
    der_weno_cell<T>(phi, g, i, j, k, c, false, ta, tb, tc, td, te, tf);
    if (!(tabs(ta) > tabs(td))) ta = td;
    if (!(tabs(tb) > tabs(te))) tb = te;
    if (!(tabs(tc) > tabs(tf))) tc = tf;
    const T m = sqrt(ta * ta + tb * tb + tc * tc);
    modls[c] = m;
    if (m == T(0)) {
      nrx[c] = T(0); nry[c] = T(0); nrz[c] = T(0);
    } else {
      nrx[c] = ta / m; nry[c] = tb / m; nrz[c] = tc / m;
    }
    const T lsv = phi[c];
    const T a = T(ph.alpha1), dx1 = T(ph.dx1);
    const T pi = T(SULI_PI);
    T h, hx, hy, hz;
    if (lsv < -a * dx1) {
      h = T(0); hx = T(0); hy = T(0); hz = T(0);
    } else if (lsv > a * dx1) {
      h = T(1); hx = T(0); hy = T(0); hz = T(0);
    } else {
      h = T(0.5) * (T(1.) + lsv / (a * dx1) + T(1.) / pi * sin(pi * lsv / (a * dx1)));
      const T cs = cos(pi * lsv / (a * dx1));
      hx = T(0.5) * ta * (T(1.) + cs) / (a * dx1);
      hy = T(0.5) * tb * (T(1.) + cs) / (a * dx1);
      hz = T(0.5) * tc * (T(1.) + cs) / (a * dx1);
    }
    hs[c] = h; hsx[c] = hx; hsy[c] = hy; hsz[c] = hz;
    rho[c] = T(ph.rho_f1) * (T(1.) - h) + T(ph.rho_f2) * h;
    lsnu[c] = T(ph.mi_f1) * (T(1.) - h) + T(ph.mi_f2) * h;
    acc[0] = (1. - (1. - (double)h)) * g.dx * g.dy * g.dz;
  }
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1>(acc, lds);
  block_partial<1>(acc, rs.partials);
}

// Second half of mod_ls1 (:761-803): WENO (ihs = 1 everywhere) of each normal
// component along its own direction, larger-magnitude selection, curvature.
template <typename T>
__global__ __launch_bounds__(256) void k_curv(Grid g, const T *__restrict__ nrx,
                                              const T *__restrict__ nry, const T *__restrict__ nrz,
                                              T *__restrict__ ddx, T *__restrict__ ddy,
                                              T *__restrict__ ddz, T *__restrict__ kurv) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const fidx c = g.idx(i, j, k);
  T w[7], ta, tb, tc, td, te, tf;
  window7<T>(nrx, c, 1, i, g.nx, 1, w);
  weno_pm<T>(w, T(g.dx), ta, td);
  window7<T>(nry, c, g.sx, j, g.ny, 1, w);
  weno_pm<T>(w, T(g.dy), tb, te);
  window7<T>(nrz, c, g.sz, k, g.nz, 1, w);
  weno_pm<T>(w, T(g.dz), tc, tf);
  const T a = tabs(td) > tabs(ta) ? td : ta;
  const T b = tabs(te) > tabs(tb) ? te : tb;
  const T cc = tabs(tf) > tabs(tc) ? tf : tc;
  ddx[c] = a; ddy[c] = b; ddz[c] = cc;
  kurv[c] = a + b + cc;
}

// Volume correction step (:259-265): phi -= (dt1*dV)*|grad phi|/V0.
template <typename T>
__global__ __launch_bounds__(256) void k_masscorr(Grid g, T *__restrict__ phi,
                                                  const T *__restrict__ modls, T f, T vol0) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const fidx c = g.idx(i, j, k);
  phi[c] = phi[c] - f * modls[c] / vol0;
}

// ---------------------------------------------------------------------------
// Smagorinsky eddy viscosity, visco m_turb = 1 (5_viscosidade.f90:61-97):
// central derivatives at the faces (8_derivadas.f90:1-68) averaged to the
// cell centre (8_derivadas.f90:331-455) and nu_t = c23*Cs^2*sqrt(2 S^2)*rho.
template <typename T>
__global__ __launch_bounds__(256) void k_visco(Grid g, T c23, T cs, const T *__restrict__ u,
                                               const T *__restrict__ v, const T *__restrict__ w,
                                               const T *__restrict__ rho, T *__restrict__ nut) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const fidx c = g.idx(i, j, k);
  const fidx X = 1, Y = g.sx, Z = g.sz;
  const T h2x = T(2.) * T(g.dx), h2y = T(2.) * T(g.dy), h2z = T(2.) * T(g.dz);
  // u at faces i (c) and i+1 (c+X)
  const T dudx = ((u[c + X] - u[c - X]) / h2x + (u[c + 2 * X] - u[c]) / h2x) * T(0.5);
  const T dudy = ((u[c + Y] - u[c - Y]) / h2y + (u[c + X + Y] - u[c + X - Y]) / h2y) * T(0.5);
  const T dudz = ((u[c + Z] - u[c - Z]) / h2z + (u[c + X + Z] - u[c + X - Z]) / h2z) * T(0.5);
  // v at faces j (c) and j+1 (c+Y)
  const T dvdx = ((v[c + X] - v[c - X]) / h2x + (v[c + Y + X] - v[c + Y - X]) / h2x) * T(0.5);
  const T dvdy = ((v[c + Y] - v[c - Y]) / h2y + (v[c + 2 * Y] - v[c]) / h2y) * T(0.5);
  const T dvdz = ((v[c + Z] - v[c - Z]) / h2z + (v[c + Y + Z] - v[c + Y - Z]) / h2z) * T(0.5);
  // w at faces k (c) and k+1 (c+Z)
  const T dwdx = ((w[c + X] - w[c - X]) / h2x + (w[c + Z + X] - w[c + Z - X]) / h2x) * T(0.5);
  const T dwdy = ((w[c + Y] - w[c - Y]) / h2y + (w[c + Z + Y] - w[c + Z - Y]) / h2y) * T(0.5);
  const T dwdz = ((w[c + Z] - w[c - Z]) / h2z + (w[c + 2 * Z] - w[c]) / h2z) * T(0.5);
  const T p1 = dudx + T(0.5) * (dudy + dvdx + dudz + dwdx);
  const T p2 = dvdy + T(0.5) * (dvdx + dudy + dvdz + dwdy);
  const T p3 = dwdz + T(0.5) * (dwdy + dvdz + dudz + dwdx);
  nut[c] = c23 * cs * cs * sqrt(T(2.) * (p1 + p2 + p3) * (p1 + p2 + p3)) * rho[c];
}

// ---------------------------------------------------------------------------
// Momentum: convdiff (6_convdiff.f90:14-189) with first-order upwind
// advection classico der=1 (7_advectivo.f90:33-113), fused with the stage
// update of tempo (11_diversos.f90:10-228, every t_tempo scheme).  The face viscosities
// xnut/ynut/znut (5_viscosidade.f90:178-211) and their edge averages
// (6_convdiff.f90:118-145) are rebuilt on the fly from the cell arrays nut
// and ls_nu with the reference's end-face copies.
//
// With a bottom obstacle the friction plane of each column is the face just
// above it, k = ku(i,j)+1 (3_contorno.f90:273-275), and the wall planes are
// not used.
// The friction term -(gz/C^2)*|ub|*u (6_convdiff.f90:156,168,183) reads ub
// from u itself: contorno copies u into ub on the planes k=2, j=2, j=ny
// (3_contorno.f90:247-264) and contorno(1) always runs between the last
// change of u and convdiff (1_principal.f90:46,92), so ub equals u there and
// is zero elsewhere.
template <typename T>
struct MomArgs {
  const T *u, *v, *w, *u0, *v0, *w0;
  const T *rho, *lsnu, *nut, *hsx, *hsy, *hsz, *kurv;
  const T *tfu, *tfv, *tfw;  // optional source terms (may be null)
  T *un, *vn, *wn, *u0o, *v0o, *w0o;
  T *Fu, *Fv, *Fw;           // optional force output (may be null)
  T *fu0, *fv0, *fw0;        // kept forces of RK3 stage 1 / AB2 (null when unused)
  T *fu1, *fv1, *fw1;        // RK3 stage-2 forces (null when unused)
  const T *xnf, *ynf, *znf;  // face viscosities of the directional model (null: rebuilt from nut, ls_nu)
  const int *ku, *kv, *kw;   // bottom-obstacle columns, plane index i + sx*j (null: no obstacle)
  T dt;
  int mode;                  // MomUpdate
};

// The velocity update of one stage of tempo (11_diversos.f90:10-228), per
// face: uc the velocity the forces F were evaluated at, u0 the start-of-step
// copy, fa/fb the kept forces (fu0/fu1).  Fortran evaluation order, so that
// the FP64 instantiation rounds as the reference does.
template <typename T>
__device__ __forceinline__ void stage_update(int mode, T dt, fidx c, T uc, T F, const T *__restrict__ u0,
                                             T *__restrict__ u0o, T *__restrict__ un, T *fa, T *fb) {
  switch (mode) {
    case MOM_RK_FIRST:
      u0o[c] = uc;
      if (fa) fa[c] = F;
      un[c] = uc + dt * F;
      break;
    case MOM_FROM_U0: un[c] = u0[c] + dt * F; break;
    case MOM_EULER:
      if (fa) fa[c] = F;
      un[c] = uc + dt * F;
      break;
    case MOM_RK3_2:
      fb[c] = F;
      un[c] = u0[c] - dt * fa[c] + (T(2.) * dt) * F;
      break;
    case MOM_RK3_3: un[c] = u0[c] + (dt / T(6.)) * (fa[c] + T(4.) * fb[c] + F); break;
    default: {  // MOM_AB2
      const T f0 = fa[c];
      fa[c] = F;
      un[c] = uc + dt * T(0.5) * (T(3.) * F - f0);
    }
  }
}

// IN: the caller guarantees 2 <= f <= n (interior), so the clamps of the
// end faces and cells are no-ops and left out (the same values are read)
template <typename T, bool IN = false, typename P = const T *>
__device__ __forceinline__ T interp_face(P a, fidx cf, fidx st, int f,
                                         int n) {
  // interp*_cf at face f (8_derivadas.f90:304-434): average of cells fc-1, fc
  // with fc = clamp(f, 2, n); cf is the address of face f's cell (f).
  const int fc = IN ? f : (f < 2 ? 2 : (f > n ? n : f));
  const fidx cc = cf + (fidx)(fc - f) * st;
  return (a[cc] + a[cc - st]) * T(0.5);
}
template <typename T, bool IN = false, typename P = const T *>
__device__ __forceinline__ T cell_clamp(P a, fidx cf, fidx st, int f,
                                        int n) {
  const int fc = IN ? f : (f < 1 ? 1 : (f > n ? n : f));
  return a[cf + (fidx)(fc - f) * st];
}
// face viscosity along d1 at (face f1 along d1), e.g. xnut(i,j,k) = ls_nux + nutx
template <typename T, bool IN = false, typename P = const T *>
__device__ __forceinline__ T face_visc(P lsnu, P nut,
                                       fidx cf, fidx st, int f, int n) {
  return interp_face<T, IN, P>(lsnu, cf, st, f, n) + interp_face<T, IN, P>(nut, cf, st, f, n);
}
// edge value: interp along d2 of the face viscosity along d1, divided by h2
// (IN1: no clamp along d1, IN2: none along d2)
template <typename T, bool IN1 = false, bool IN2 = false, typename P = const T *>
__device__ __forceinline__ T edge_visc(P lsnu, P nut,
                                       fidx ce, fidx st1, int f1, int n1, fidx st2, int f2,
                                       int n2, T h2) {
  const int fc = IN2 ? f2 : (f2 < 2 ? 2 : (f2 > n2 ? n2 : f2));
  const fidx cc = ce + (fidx)(fc - f2) * st2;
  return ((face_visc<T, IN1, P>(lsnu, nut, cc, st1, f1, n1) + face_visc<T, IN1, P>(lsnu, nut, cc - st2, st1, f1, n1)) *
          T(0.5)) / h2;
}
// edge viscosity from stored face viscosities fa (FACES: the directional
// model) or rebuilt from the cell arrays
template <typename T, bool FACES, bool IN1 = false, bool IN2 = false, typename P = const T *>
__device__ __forceinline__ T evisc(P fa, P lsnu, P nut,
                                   fidx ce, fidx st1, int f1, int n1, fidx st2, int f2, int n2, T h2) {
  if constexpr (!FACES) return edge_visc<T, IN1, IN2, P>(lsnu, nut, ce, st1, f1, n1, st2, f2, n2, h2);
  const int fc = IN2 ? f2 : (f2 < 2 ? 2 : (f2 > n2 ? n2 : f2));
  const fidx cc = ce + (fidx)(fc - f2) * st2;
  return ((fa[cc] + fa[cc - st2]) * T(0.5)) / h2;
}
template <typename T, typename P = const T *>
__device__ __forceinline__ T surf_cell(P hd, P kurv, P rho, fidx cc) {
  return hd[cc] * kurv[cc] / rho[cc];
}

// ---------------------------------------------------------------------------
// Advective terms of convdiff (6_convdiff.f90:70-82), one face of one
// velocity component each.  ADV = 10*adv_type + der: 11 classico upwind
// (7_advectivo.f90:33-113), 12 classico central (:116-173), 13 classico
// upwind second order (:176-291), 21 rotacional (:296-440), 31 antissim
// (:443-653).  Each sum in the reference's order.

// second-order one-sided differences of 8_derivadas.f90:83-290 along stride
// st at face f of n: backward (central at f = 1), forward (central at f = n)
template <typename T, typename P>
__device__ __forceinline__ T upw2(P q, fidx c, fidx st, int f, int n, T h, T a) {
  const T qc = q[c], qm = q[c - st], qp = q[c + st];
  const T dn = f >= 2 ? T(0.5) * (q[c - 2 * st] - T(4.) * qm + T(3.) * qc) / h : T(0.5) * (qp - qm) / h;
  const T dp = f <= n - 1 ? T(0.5) * (-q[c + 2 * st] + T(4.) * qp - T(3.) * qc) / h : T(0.5) * (qp - qm) / h;
  return tmax(a, T(0)) * dn + tmin(a, T(0)) * dp;
}
// flux-form upwind term along one direction (rotacional), or its average
// with the advective form (antissim); am, ac, ap: advecting velocity at the
// faces behind, here, ahead
template <typename T, bool ANTI, typename P>
__device__ __forceinline__ T flux1(P q, fidx c, fidx st, T am, T ac, T ap, T h) {
  const T qc = q[c], qm = q[c - st], qp = q[c + st];
  const T apc = tmax(ac, T(0)), apm = tmax(am, T(0)), anp = tmin(ap, T(0)), anc = tmin(ac, T(0));
  if (!ANTI) return (apc * qc - apm * qm) / h + (anp * qp - anc * qc) / h;
  return ((apc * qc - apm * qm) / h + (anp * qp - anc * qc) / h + apc * (qc - qm) / h + anc * (qp - qc) / h) *
         T(0.5);
}
template <typename T>
__device__ __forceinline__ int clamp2(int f, int n) { return f < 2 ? 2 : (f > n ? n : f); }

template <typename T, int ADV, typename P>
__device__ __forceinline__ T adv_u(const Grid &g, P u, P v,
                                   P w, fidx c, int i, int j, int k, T dx, T dy, T dz) {
  const fidx X = 1, Y = g.sx, Z = g.sz;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const T uc = u[c];
  if constexpr (ADV == 11) {
    const T bma_a = (v[c] + v[c - X] + v[c + Y] + v[c - X + Y]) * T(0.25);
    const T dma_a = (w[c] + w[c - X] + w[c + Z] + w[c - X + Z]) * T(0.25);
    T a1 = tmax(uc, T(0)), a2 = tmin(uc, T(0));
    const T dudx = a1 * (uc - u[c - X]) / dx + a2 * (u[c + X] - uc) / dx;
    a1 = tmax(bma_a, T(0)); a2 = tmin(bma_a, T(0));
    const T dudy = a1 * (uc - u[c - Y]) / dy + a2 * (u[c + Y] - uc) / dy;
    a1 = tmax(dma_a, T(0)); a2 = tmin(dma_a, T(0));
    const T dudz = a1 * (uc - u[c - Z]) / dz + a2 * (u[c + Z] - uc) / dz;
    const T uint_ = dudx + dudy + dudz;
    return uint_;
  } else if constexpr (ADV == 12 || ADV == 13) {
    const fidx cc = c + (fidx)(clamp2<T>(i, nx) - i);  // interpx_cf of the cell averages
    const T bma = ((v[cc] + v[cc + Y]) * T(0.5) + (v[cc - X] + v[cc - X + Y]) * T(0.5)) * T(0.5);
    const T dma = ((w[cc] + w[cc + Z]) * T(0.5) + (w[cc - X] + w[cc - X + Z]) * T(0.5)) * T(0.5);
    if constexpr (ADV == 12)
      return uc * ((u[c + X] - u[c - X]) / (T(2.) * dx)) + bma * ((u[c + Y] - u[c - Y]) / (T(2.) * dy)) +
             dma * ((u[c + Z] - u[c - Z]) / (T(2.) * dz));
    else
      return upw2<T>(u, c, X, i, nx + 1, dx, uc) + upw2<T>(u, c, Y, j, ny, dy, bma) + upw2<T>(u, c, Z, k, nz, dz, dma);
  } else {
    constexpr bool A = ADV == 31;
    auto bma = [&](fidx e) { return (v[e] + v[e - X] + v[e + Y] + v[e - X + Y]) * T(0.25); };
    auto dma = [&](fidx e) { return (w[e] + w[e - X] + w[e + Z] + w[e - X + Z]) * T(0.25); };
    const T tx = flux1<T, A>(u, c, X, A && i == 1 ? uc : u[c - X], uc, A && i == nx + 1 ? uc : u[c + X], dx);
    const T bc = bma(c), dc = dma(c);
    const T ty = flux1<T, A>(u, c, Y, A && j == 1 ? bc : bma(c - Y), bc, bma(c + Y), dy);
    const T tz = flux1<T, A>(u, c, Z, A && k == 1 ? dc : dma(c - Z), dc, dma(c + Z), dz);
    return tx + ty + tz;
  }
}

template <typename T, int ADV, typename P>
__device__ __forceinline__ T adv_v(const Grid &g, P u, P v,
                                   P w, fidx c, int i, int j, int k, T dx, T dy, T dz) {
  const fidx X = 1, Y = g.sx, Z = g.sz;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const T vc = v[c];
  if constexpr (ADV == 11) {
    const T amb_a = (u[c] + u[c + X] + u[c - Y] + u[c + X - Y]) * T(0.25);
    const T dmb_a = (w[c] + w[c - Y] + w[c + Z] + w[c - Y + Z]) * T(0.25);
    T a1 = tmax(amb_a, T(0)), a2 = tmin(amb_a, T(0));
    const T dvdx = a1 * (vc - v[c - X]) / dx + a2 * (v[c + X] - vc) / dx;
    a1 = tmax(vc, T(0)); a2 = tmin(vc, T(0));
    const T dvdy = a1 * (vc - v[c - Y]) / dy + a2 * (v[c + Y] - vc) / dy;
    a1 = tmax(dmb_a, T(0)); a2 = tmin(dmb_a, T(0));
    const T dvdz = a1 * (vc - v[c - Z]) / dz + a2 * (v[c + Z] - vc) / dz;
    const T vint_ = dvdx + dvdy + dvdz;
    return vint_;
  } else if constexpr (ADV == 12 || ADV == 13) {
    const fidx cc = c + (fidx)(clamp2<T>(j, ny) - j) * Y;  // interpy_cf of the cell averages
    const T amb = ((u[cc] + u[cc + X]) * T(0.5) + (u[cc - Y] + u[cc - Y + X]) * T(0.5)) * T(0.5);
    const T dmb = ((w[cc] + w[cc + Z]) * T(0.5) + (w[cc - Y] + w[cc - Y + Z]) * T(0.5)) * T(0.5);
    if constexpr (ADV == 12)
      return amb * ((v[c + X] - v[c - X]) / (T(2.) * dx)) + vc * ((v[c + Y] - v[c - Y]) / (T(2.) * dy)) +
             dmb * ((v[c + Z] - v[c - Z]) / (T(2.) * dz));
    else
      return upw2<T>(v, c, X, i, nx, dx, amb) + upw2<T>(v, c, Y, j, ny + 1, dy, vc) + upw2<T>(v, c, Z, k, nz, dz, dmb);
  } else {
    constexpr bool A = ADV == 31;
    auto amb = [&](fidx e) { return (u[e] + u[e + X] + u[e - Y] + u[e + X - Y]) * T(0.25); };
    auto dmb = [&](fidx e) { return (w[e] + w[e - Y] + w[e + Z] + w[e - Y + Z]) * T(0.25); };
    const T ac = amb(c), dc = dmb(c);
    const T tx = flux1<T, A>(v, c, X, A && i == 1 ? ac : amb(c - X), ac, amb(c + X), dx);
    const T ty = flux1<T, A>(v, c, Y, A && j == 1 ? vc : v[c - Y], vc, A && j == ny + 1 ? vc : v[c + Y], dy);
    const T tz = flux1<T, A>(v, c, Z, A && k == 1 ? dc : dmb(c - Z), dc, dmb(c + Z), dz);
    return tx + ty + tz;
  }
}

template <typename T, int ADV, typename P>
__device__ __forceinline__ T adv_w(const Grid &g, P u, P v,
                                   P w, fidx c, int i, int j, int k, T dx, T dy, T dz) {
  const fidx X = 1, Y = g.sx, Z = g.sz;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const T wc = w[c];
  if constexpr (ADV == 11) {
    const T amd_a = (u[c] + u[c + X] + u[c - Z] + u[c + X - Z]) * T(0.25);
    const T bmd_a = (v[c] + v[c + Y] + v[c - Z] + v[c + Y - Z]) * T(0.25);
    T a1 = tmax(amd_a, T(0)), a2 = tmin(amd_a, T(0));
    const T dwdx = a1 * (wc - w[c - X]) / dx + a2 * (w[c + X] - wc) / dx;
    a1 = tmax(bmd_a, T(0)); a2 = tmin(bmd_a, T(0));
    const T dwdy = a1 * (wc - w[c - Y]) / dy + a2 * (w[c + Y] - wc) / dy;
    a1 = tmax(wc, T(0)); a2 = tmin(wc, T(0));
    const T dwdz = a1 * (wc - w[c - Z]) / dz + a2 * (w[c + Z] - wc) / dz;
    const T wint_ = dwdx + dwdy + dwdz;
    return wint_;
  } else if constexpr (ADV == 12 || ADV == 13) {
    const fidx cc = c + (fidx)(clamp2<T>(k, nz) - k) * Z;  // interpz_cf of the cell averages
    const T amd = ((u[cc] + u[cc + X]) * T(0.5) + (u[cc - Z] + u[cc - Z + X]) * T(0.5)) * T(0.5);
    const T bmd = ((v[cc] + v[cc + Y]) * T(0.5) + (v[cc - Z] + v[cc - Z + Y]) * T(0.5)) * T(0.5);
    if constexpr (ADV == 12)
      return amd * ((w[c + X] - w[c - X]) / (T(2.) * dx)) + bmd * ((w[c + Y] - w[c - Y]) / (T(2.) * dy)) +
             wc * ((w[c + Z] - w[c - Z]) / (T(2.) * dz));
    else
      return upw2<T>(w, c, X, i, nx, dx, amd) + upw2<T>(w, c, Y, j, ny, dy, bmd) + upw2<T>(w, c, Z, k, nz + 1, dz, wc);
  } else {
    // rotacional's dn at k = nz+2 is never set by the reference (it assigns
    // dp there); min(w(nz+2), 0) is used, as for every other entry
    constexpr bool A = ADV == 31;
    auto amd = [&](fidx e) { return (u[e] + u[e + X] + u[e - Z] + u[e + X - Z]) * T(0.25); };
    auto bmd = [&](fidx e) { return (v[e] + v[e + Y] + v[e - Z] + v[e + Y - Z]) * T(0.25); };
    const T ac = amd(c), bc = bmd(c);
    const T tx = flux1<T, A>(w, c, X, A && i == 1 ? ac : amd(c - X), ac, amd(c + X), dx);
    const T ty = flux1<T, A>(w, c, Y, A && j == 1 ? bc : bmd(c - Y), bc, bmd(c + Y), dy);
    const T tz = flux1<T, A>(w, c, Z, A && k == 1 ? wc : w[c - Z], wc, A && k == nz + 1 ? wc : w[c + Z], dz);
    return tx + ty + tz;
  }
}

// Directional Smagorinsky model (m_turb = 2, 5_viscosidade.f90:104-170,
// 178-211): the face viscosities xnut, ynut, znut from the strain at each
// face (central differences of the velocities, 8_derivadas.f90:1-68, moved
// to the face by the interpolations of :104-131) plus ls_nu interpolated to
// the face; nut is left unchanged.  One thread per (i,j,k) of
// [1,nx+1]x[1,ny+1]x[1,nz+1], writing each face kind it owns.
template <typename T>
__global__ __launch_bounds__(256) void k_visco_dir(Grid g, T cs, const T *__restrict__ u, const T *__restrict__ v,
                                                   const T *__restrict__ w, const T *__restrict__ lsnu,
                                                   T *__restrict__ xn, T *__restrict__ yn, T *__restrict__ zn) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx + 1 || j > ny + 1) return;
  const fidx X = 1, Y = g.sx, Z = g.sz;
  const fidx c = g.idx(i, j, k);
  const T dx = (T)g.dx, dy = (T)g.dy, dz = (T)g.dz;
  auto cdx = [&](const T *q, fidx e) { return (q[e + X] - q[e - X]) / (T(2.) * dx); };
  auto cdy = [&](const T *q, fidx e) { return (q[e + Y] - q[e - Y]) / (T(2.) * dy); };
  auto cdz = [&](const T *q, fidx e) { return (q[e + Z] - q[e - Z]) / (T(2.) * dz); };
  if (j <= ny && k <= nz) {  // u face
    const fidx ci = c + (fidx)(clamp2<T>(i, nx) - i);  // interpx_cf column
    auto dvdx_a = [&](fidx e) { return (cdx(v, e) + cdx(v, e - X)) * T(0.5); };
    auto dwdx_a = [&](fidx e) { return (cdx(w, e) + cdx(w, e - X)) * T(0.5); };
    const T dvdx_x = (dvdx_a(ci) + dvdx_a(ci + Y)) * T(0.5);
    const T dwdx_x = (dwdx_a(ci) + dwdx_a(ci + Z)) * T(0.5);
    const T p1 = cdx(u, c) + T(0.5) * (cdy(u, c) + dvdx_x + cdz(u, c) + dwdx_x);
    xn[c] = interp_face<T>(lsnu, c, X, i, nx) + dx * dx * cs * cs * sqrt(T(2.) * p1 * p1);
  }
  if (i <= nx && k <= nz) {  // v face
    const fidx cj = c + (fidx)(clamp2<T>(j, ny) - j) * Y;
    auto dudy_a = [&](fidx e) { return (cdy(u, e) + cdy(u, e - Y)) * T(0.5); };
    auto dwdy_a = [&](fidx e) { return (cdy(w, e) + cdy(w, e - Y)) * T(0.5); };
    const T dudy_y = (dudy_a(cj) + dudy_a(cj + X)) * T(0.5);
    const T dwdy_y = (dwdy_a(cj) + dwdy_a(cj + Z)) * T(0.5);
    const T p1 = cdy(v, c) + T(0.5) * (cdx(v, c) + dudy_y + cdz(v, c) + dwdy_y);
    yn[c] = interp_face<T>(lsnu, c, Y, j, ny) + dy * dy * cs * cs * sqrt(T(2.) * p1 * p1);
  }
  if (i <= nx && j <= ny) {  // w face
    const fidx ck = c + (fidx)(clamp2<T>(k, nz) - k) * Z;
    auto dvdz_a = [&](fidx e) { return (cdz(v, e) + cdz(v, e - Z)) * T(0.5); };
    auto dudz_a = [&](fidx e) { return (cdz(u, e) + cdz(u, e - Z)) * T(0.5); };
    const T dvdz_z = (dvdz_a(ck) + dvdz_a(ck + Y)) * T(0.5);
    const T dudz_z = (dudz_a(ck) + dudz_a(ck + X)) * T(0.5);
    const T p1 = cdz(w, c) + T(0.5) * (cdy(w, c) + dvdz_z + dudz_z + cdx(w, c));
    zn[c] = interp_face<T>(lsnu, c, Z, k, nz) + dz * dz * cs * cs * sqrt(T(2.) * p1 * p1);
  }
}

// comp: the face kind to update (0 u, 1 v, 2 w) or -1 for all three
template <typename T, int ADV, bool FACES, bool INX = false, bool INYZ = false>
__device__ __forceinline__ void momentum_cell(const Grid &g, const Phys &ph, const MomArgs<T> &A, int i, int j,
                                              int k, int comp = -1) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const fidx c = g.idx(i, j, k);
  const fidx X = 1, Y = g.sx, Z = g.sz;
  const T dx = T(g.dx), dy = T(g.dy), dz = T(g.dz);
  // (raw pointers: reading these through 32-bit byte offsets, as the WENO
  // windows do, measured 2 % slower here)
  const T *__restrict__ u = A.u, *__restrict__ v = A.v, *__restrict__ w = A.w;
  const T *__restrict__ nut = A.nut, *__restrict__ lsnu = A.lsnu, *__restrict__ rho = A.rho;
  const T *__restrict__ hsx = A.hsx, *__restrict__ hsy = A.hsy, *__restrict__ hsz = A.hsz, *__restrict__ kurv = A.kurv;
  const T *__restrict__ xnf = A.xnf, *__restrict__ ynf = A.ynf, *__restrict__ znf = A.znf;
  const T fr = T(ph.gz / (ph.chezy * ph.chezy));
  const T sigma = T(ph.sigma);
  // ---- u faces: i in [1,nx+1], j in [1,ny], k in [1,nz]
  if ((comp < 0 || comp == 0) && i <= nx + 2 && j <= ny + 1 && k <= nz + 1) {
    if (i >= 1 && i <= nx + 1 && j >= 1 && j <= ny && k >= 1 && k <= nz) {
      const T uc = u[c];
      const T uint_ = adv_u<T, ADV>(g, u, v, w, c, i, j, k, dx, dy, dz);
      const T amai = cell_clamp<T, INX>(nut, c, X, i, nx) / dx;
      const T amam = cell_clamp<T, INX>(nut, c - X, X, i - 1, nx) / dx;
      const T bmap = evisc<T, FACES, INX, INYZ>(xnf, lsnu, nut, c + Y, X, i, nx, Y, j + 1, ny, dy);
      const T bmac = evisc<T, FACES, INX, INYZ>(xnf, lsnu, nut, c, X, i, nx, Y, j, ny, dy);
      const T dmap = evisc<T, FACES, INX, INYZ>(xnf, lsnu, nut, c + Z, X, i, nx, Z, k + 1, nz, dz);
      const T dmac = evisc<T, FACES, INX, INYZ>(xnf, lsnu, nut, c, X, i, nx, Z, k, nz, dz);
      const T rhox = interp_face<T, INX>(rho, c, X, i, nx);
      const int ic = INX ? i : (i < 2 ? 2 : (i > nx ? nx : i));
      const fidx cc = c + (fidx)(ic - i);
      const T dhsdx = (surf_cell<T>(hsx, kurv, rho, cc) + surf_cell<T>(hsx, kurv, rho, cc - X)) * T(0.5);
      const bool pl = A.ku ? k == A.ku[i + g.sx * j] + 1
                           : (g.ccz0 == 2 && k == 2) || (g.ccy0 == 2 && j == 2) || (g.ccyf == 2 && j == ny);
      const T ub = pl ? uc : T(0);
      const T tf = A.tfu ? A.tfu[c] : T(0);
      const T F = -uint_ +
                  ((amai * T(2.) * (u[c + X] - uc) - amam * T(2.) * (uc - u[c - X])) / dx +
                   bmap * ((u[c + Y] - uc) / dy + (v[c + Y] - v[c - X + Y]) / dx) -
                   bmac * ((uc - u[c - Y]) / dy + (v[c] - v[c - X]) / dx) +
                   dmap * ((u[c + Z] - uc) / dz + (w[c + Z] - w[c - X + Z]) / dx) -
                   dmac * ((uc - u[c - Z]) / dz + (w[c] - w[c - X]) / dx)) /
                      rhox +
                  T(ph.gx) - sigma * dhsdx + tf - fr * sqrt(ub * ub) * uc;
      if (A.Fu) A.Fu[c] = F;
      stage_update<T>(A.mode, A.dt, c, uc, F, A.u0, A.u0o, A.un, A.fu0, A.fu1);
    } else {
      A.un[c] = u[c];
    }
  }
  // ---- v faces: i in [1,nx], j in [1,ny+1], k in [1,nz]
  if ((comp < 0 || comp == 1) && i <= nx + 1 && j <= ny + 2 && k <= nz + 1) {
    if (i >= 1 && i <= nx && j >= 1 && j <= ny + 1 && k >= 1 && k <= nz) {
      const T vc = v[c];
      const T vint_ = adv_v<T, ADV>(g, u, v, w, c, i, j, k, dx, dy, dz);
      const T ambp = evisc<T, FACES, INYZ, INX>(ynf, lsnu, nut, c + X, Y, j, ny, X, i + 1, nx, dx);
      const T ambc = evisc<T, FACES, INYZ, INX>(ynf, lsnu, nut, c, Y, j, ny, X, i, nx, dx);
      const T bmbj = cell_clamp<T, INYZ>(nut, c, Y, j, ny) / dy;
      const T bmbm = cell_clamp<T, INYZ>(nut, c - Y, Y, j - 1, ny) / dy;
      const T dmbp = evisc<T, FACES, INYZ, INYZ>(ynf, lsnu, nut, c + Z, Y, j, ny, Z, k + 1, nz, dz);
      const T dmbc = evisc<T, FACES, INYZ, INYZ>(ynf, lsnu, nut, c, Y, j, ny, Z, k, nz, dz);
      const T rhoy = interp_face<T, INYZ>(rho, c, Y, j, ny);
      const int jc = INYZ ? j : (j < 2 ? 2 : (j > ny ? ny : j));
      const fidx cc = c + (fidx)(jc - j) * Y;
      const T dhsdy = (surf_cell<T>(hsy, kurv, rho, cc) + surf_cell<T>(hsy, kurv, rho, cc - Y)) * T(0.5);
      const bool pl = A.kv ? k == A.kv[i + g.sx * j] + 1
                           : (g.ccz0 == 2 && k == 2) || (g.ccy0 == 2 && j == 2) || (g.ccyf == 2 && j == ny);
      const T vb = pl ? vc : T(0);
      const T tf = A.tfv ? A.tfv[c] : T(0);
      const T F = -vint_ +
                  (ambp * ((v[c + X] - vc) / dx + (u[c + X] - u[c + X - Y]) / dy) -
                   ambc * ((vc - v[c - X]) / dx + (u[c] - u[c - Y]) / dy) +
                   (bmbj * T(2.) * (v[c + Y] - vc) - bmbm * T(2.) * (vc - v[c - Y])) / dy +
                   dmbp * ((v[c + Z] - vc) / dz + (w[c + Z] - w[c - Y + Z]) / dy) -
                   dmbc * ((vc - v[c - Z]) / dz + (w[c] - w[c - Y]) / dy)) /
                      rhoy -
                  sigma * dhsdy + tf - fr * sqrt(vb * vb) * vc;
      if (A.Fv) A.Fv[c] = F;
      stage_update<T>(A.mode, A.dt, c, vc, F, A.v0, A.v0o, A.vn, A.fv0, A.fv1);
    } else {
      A.vn[c] = v[c];
    }
  }
  // ---- w faces: i in [1,nx], j in [1,ny], k in [1,nz+1]
  if ((comp < 0 || comp == 2) && i <= nx + 1 && j <= ny + 1 && k <= nz + 2) {
    if (i >= 1 && i <= nx && j >= 1 && j <= ny && k >= 1 && k <= nz + 1) {
      const T wc = w[c];
      const T wint_ = adv_w<T, ADV>(g, u, v, w, c, i, j, k, dx, dy, dz);
      const T amdp = evisc<T, FACES, INYZ, INX>(znf, lsnu, nut, c + X, Z, k, nz, X, i + 1, nx, dx);
      const T amdc = evisc<T, FACES, INYZ, INX>(znf, lsnu, nut, c, Z, k, nz, X, i, nx, dx);
      const T bmdp = evisc<T, FACES, INYZ, INYZ>(znf, lsnu, nut, c + Y, Z, k, nz, Y, j + 1, ny, dy);
      const T bmdc = evisc<T, FACES, INYZ, INYZ>(znf, lsnu, nut, c, Z, k, nz, Y, j, ny, dy);
      const T dmdk = cell_clamp<T, INYZ>(nut, c, Z, k, nz) / dz;
      const T dmdm = cell_clamp<T, INYZ>(nut, c - Z, Z, k - 1, nz) / dz;
      const T rhoz = interp_face<T, INYZ>(rho, c, Z, k, nz);
      const int kc = INYZ ? k : (k < 2 ? 2 : (k > nz ? nz : k));
      const fidx cc = c + (fidx)(kc - k) * Z;
      const T dhsdz = (surf_cell<T>(hsz, kurv, rho, cc) + surf_cell<T>(hsz, kurv, rho, cc - Z)) * T(0.5);
      const bool pl = A.kw ? k == A.kw[i + g.sx * j] + 1
                           : (g.ccz0 == 2 && k == 2) || (g.ccy0 == 2 && j == 2) || (g.ccyf == 2 && j == ny);
      const T wb = pl ? wc : T(0);
      const T tf = A.tfw ? A.tfw[c] : T(0);
      const T epis_z = T(0);
      const T F = -wint_ +
                  (amdp * ((w[c + X] - wc) / dx + (u[c + X] - u[c + X - Z]) / dz) -
                   amdc * ((wc - w[c - X]) / dx + (u[c] - u[c - Z]) / dz) +
                   bmdp * ((w[c + Y] - wc) / dy + (v[c + Y] - v[c + Y - Z]) / dz) -
                   bmdc * ((wc - w[c - Y]) / dy + (v[c] - v[c - Z]) / dz) +
                   (dmdk * T(2.) * (w[c + Z] - wc) - dmdm * T(2.) * (wc - w[c - Z])) / dz) /
                      rhoz -
                  (wc - T(0.)) * epis_z - T(ph.gz) - sigma * dhsdz + tf - fr * sqrt(wb * wb) * wc;
      if (A.Fw) A.Fw[c] = F;
      stage_update<T>(A.mode, A.dt, c, wc, F, A.w0, A.w0o, A.wn, A.fw0, A.fw1);
    } else {
      A.wn[c] = w[c];
    }
  }
}

// All three momentum updates of one RK stage, one k-plane per block.
// waves_per_eu(6): 80 VGPRs instead of 82 lifts the occupancy from 5 to 6
// waves per SIMD without spills (8 would spill)
//
// Two launches per stage: the inner box 1 <= i <= nx, 2 <= j <= ny-1,
// 2 <= k <= nz-1 (k_momentum_tile for the default advection form, else
// k_momentum_in), where none of the end-face and end-cell clamps along j
// and k acts, so they are compiled out; and the shell of the frame around
// it (k_momentum), which holds every clamp.  The same values are computed
// either way.
//
// Shell cells in thread order: the planes k in {0, 1, nz, nz+1, nz+2} of
// the frame; for 2 <= k <= nz-1 the rows j in {0, 1, ny, ny+1, ny+2}; for
// 2 <= k <= nz-1, 2 <= j <= ny-1 the cells i in {0, nx+1, nx+2}.
__host__ __device__ __forceinline__ int64_t mom_shell_cells(const Grid &g) {
  const int64_t W = g.nx + 3, H = g.ny + 3;
  const int64_t kin = g.nz > 2 ? g.nz - 2 : 0, jin = g.ny > 2 ? g.ny - 2 : 0;
  return 5 * W * H + kin * 5 * W + kin * jin * 3;
}
__device__ __forceinline__ void mom_shell_cell(const Grid &g, int t, int &i, int &j, int &k) {
  const int W = g.nx + 3, H = g.ny + 3, nZ = 5 * W * H, nY = (g.nz > 2 ? g.nz - 2 : 0) * 5 * W;
  const auto side = [](int q, int n) { return q < 2 ? q : n + (q - 2); };  // 0, 1, n, n+1, n+2
  if (t < nZ) {
    k = side(t / (W * H), g.nz);
    const int r = t % (W * H);
    j = r / W;
    i = r % W;
  } else if (t < nZ + nY) {
    const int t2 = t - nZ, r = t2 % (5 * W);
    k = 2 + t2 / (5 * W);
    j = side(r / W, g.ny);
    i = r % W;
  } else {
    const int t3 = t - nZ - nY, jin = g.ny - 2, r = t3 % (3 * jin);
    k = 2 + t3 / (3 * jin);
    j = 2 + r / 3;
    i = r % 3 == 0 ? 0 : g.nx + r % 3;  // 0, nx+1, nx+2
  }
}
// the shell: one face kind per block row (blockIdx.y = 0 u, 1 v, 2 w), so
// that each thread walks one of the three stencils (the shell is a few
// hundred thousand cells: latency, not throughput, sets its time)
template <typename T, int ADV, bool FACES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_momentum(Grid g, Phys ph,
                                                                                          MomArgs<T> A) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= mom_shell_cells(g)) return;
  int i, j, k;
  mom_shell_cell(g, t, i, j, k);
  momentum_cell<T, ADV, FACES>(g, ph, A, i, j, k, (int)blockIdx.y);
}
template <typename T, int ADV, bool FACES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_momentum_in(Grid g, Phys ph,
                                                                                             MomArgs<T> A) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 2 + blockIdx.y * SULI_BY + threadIdx.y;
  if (i > g.nx || j > g.ny - 1) return;
  momentum_cell<T, ADV, FACES, false, true>(g, ph, A, i, j, 2 + blockIdx.z);
}

// ---------------------------------------------------------------------------
// LDS-tiled momentum predictor of the inner box 1 <= i <= nx,
// 2 <= j <= ny-1, 2 <= k <= nz-1 for the default advection form (classico
// upwind, ADV = 11) with the viscosities rebuilt from the cell arrays: the
// same values as momentum_cell<T, 11, false, false, true>, term for term in
// the same order.  The end-cell clamps along x of that function act only at
// i = 1 and i = nx, on the face interpolations of nut, ls_nu, rho and the
// surface terms: there the pair of cells (1,0) becomes (2,1) and (nx+1,nx)
// becomes (nx,nx-1).  The tile therefore holds those four fields of column
// i = 0 mirrored from column 2 and of column nx+1 from column nx-1: the
// unclamped expression then adds the same two values in the other order,
// which IEEE addition rounds identically.  (The one plain read of cell i-1,
// nut at i = 1, is clamped to cell i through the lane's LDS offset; a
// select there would change how the FP32 instantiation contracts the
// expression into FMAs.)
//
// A work-group of 64 x MT_TJ threads owns a tile of 64 i by MT_TJ j and
// walks a chunk of k planes.  Every input value is read from HBM once per
// tile (plus its one-cell halo) instead of once per stencil use:
// - ring: u, v, w, nut, ls_nu of planes k-1, k, k+1 in three LDS slots, the
//   slot of plane k-1 refilled with plane k+2 after plane k is done;
// - plane k of rho and of the surface-tension cell terms
//   hs_x kurv/rho, hs_y kurv/rho, hs_z kurv/rho (their -x, -y neighbours);
//   the -z neighbour of rho and hs_z kurv/rho is the thread's own value of
//   the previous plane, kept in registers.
// The loads of the next planes are issued before the arithmetic of the
// current one and land in LDS after it, so their latency hides behind it.
#ifndef MT_TJ
#define MT_TJ 4
#endif
template <typename T>
struct MomTile {
  static constexpr int LW = SULI_BX + 2;            // row of the tile with its i-1 and i+64 halo
  static constexpr int LH = MT_TJ + 2;              // rows with the j-1 and j+MT_TJ halo
  static constexpr int HP = LW * LH;                // one plane of one field
  static constexpr int HS = HP + 1;                 // its stride in LDS: a spare slot for idle lanes
  static constexpr int NT = SULI_BX * MT_TJ;        // threads
  static constexpr int NL = (HP + NT - 1) / NT;     // plane elements per thread
  static constexpr int NR = 5;                      // ring fields: u, v, w, nut, ls_nu
  static constexpr int NS = 4;                      // plane fields: rho, sx, sy, sz
};

// MODE: the stage update as a compile-time MomUpdate (MOM_RK_FIRST and
// MOM_FROM_U0, the two RK2 stages), or 0 for A.mode at run time.
template <typename T, int MODE>
__global__ __launch_bounds__(SULI_BX *MT_TJ) __attribute__((amdgpu_waves_per_eu(4))) void k_momentum_tile(
    Grid g, Phys ph, MomArgs<T> A, int kc) {
  using M = MomTile<T>;
  constexpr int LW = M::LW, HP = M::HP, HS = M::HS, NL = M::NL, NR = M::NR, NS = M::NS, NT = M::NT;
  __shared__ T ring[3][NR][HS];
  __shared__ T pk[NS][HS];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + SULI_BX * ty;
  // XCD-band order: the j-neighbour tiles, which read each other's halo
  // rows, run on the same XCD and share its L2
  int bx, by, bz;
  xcd_band_block(bx, by, bz);
  const int i0 = 1 + bx * SULI_BX, j0 = 2 + by * MT_TJ, kb = 2 + bz * kc;
  const int ke = min(kb + kc, nz);  // planes kb..ke-1 computed, kb-1..ke read
  const int i = i0 + tx, j = j0 + ty;
  const bool act = i <= nx && j <= ny - 1;
  const fidx Z = g.sz;
  const Fld<T> ringsrc[NR] = {{A.u}, {A.v}, {A.w}, {A.nut}, {A.lsnu}};
  const Fld<T> plsrc[5] = {{A.rho}, {A.hsx}, {A.hsy}, {A.hsz}, {A.kurv}};

  // the plane elements this thread loads: frame offset in plane 0 (clamped
  // into the frame: the values of positions outside it are never used) and
  // LDS position (the spare slot HP for lanes past the plane)
  // and, for nut, ls_nu and the plane fields, with columns 0 and nx+1
  // mirrored from 2 and nx-1
  fidx goff[NL], gofm[NL];
  int lpos[NL];
#pragma unroll
  for (int q = 0; q < NL; q++) {
    const int e = tid + q * NT;
    const int gi = min(i0 - 1 + e % LW, nx + 2), gj = min(j0 - 1 + e / LW, ny + 2);
    const int gm = gi == 0 ? 2 : (gi == nx + 1 ? nx - 1 : gi);
    goff[q] = g.idx(gi, gj, 0);
    gofm[q] = g.idx(gm, gj, 0);
    lpos[q] = e < HP ? e : HP;
  }
  T rr[NR][NL], rs[5][NL];  // in flight: ring fields of one plane, plane fields of another
  auto load_ring = [&](int kk) {
#pragma unroll
    for (int f = 0; f < NR; f++)
#pragma unroll
      for (int q = 0; q < NL; q++) rr[f][q] = ringsrc[f][(f < 3 ? goff[q] : gofm[q]) + (fidx)kk * Z];
  };
  auto store_ring = [&](int slot) {
#pragma unroll
    for (int f = 0; f < NR; f++)
#pragma unroll
      for (int q = 0; q < NL; q++) ring[slot][f][lpos[q]] = rr[f][q];
  };
  auto load_plane = [&](int kk) {
#pragma unroll
    for (int f = 0; f < 5; f++)
#pragma unroll
      for (int q = 0; q < NL; q++) rs[f][q] = plsrc[f][gofm[q] + (fidx)kk * Z];
  };
  auto store_plane = [&]() {
#pragma unroll
    for (int q = 0; q < NL; q++) {
      const T r = rs[0][q];
      pk[0][lpos[q]] = r;
      // surf_cell: hs_d * kurv / rho
      pk[1][lpos[q]] = rs[1][q] * rs[4][q] / r;
      pk[2][lpos[q]] = rs[2][q] * rs[4][q] / r;
      pk[3][lpos[q]] = rs[3][q] * rs[4][q] / r;
    }
  };

  // prologue: planes kb-1, kb, kb+1 in slots 0, 1, 2; plane kb's cell terms
  load_ring(kb - 1);
  store_ring(0);
  load_ring(kb);
  store_ring(1);
  load_ring(kb + 1);
  store_ring(2);
  load_plane(kb);
  store_plane();
  const fidx c00 = g.idx(min(i, nx), min(j, ny), 0);  // (idle lanes: a cell inside the frame)
  T rho_m, sz_m;  // rho and hs_z kurv/rho of the thread's cell in the plane below
  {
    const fidx cm = c00 + (fidx)(kb - 1) * Z;
    rho_m = A.rho[cm];
    sz_m = A.hsz[cm] * A.kurv[cm] / rho_m;
  }
  __syncthreads();

  const T dx = T(g.dx), dy = T(g.dy), dz = T(g.dz);
  const T fr = T(ph.gz / (ph.chezy * ph.chezy));
  const T sigma = T(ph.sigma);
  const int l = (ty + 1) * LW + (tx + 1);  // the thread's cell in a tile plane
  const int am = i == 1 ? 0 : -1;          // x offset of nut's cell i-1, clamped to 1
  const int cp = i + g.sx * j;             // its column in the obstacle planes
  for (int k = kb; k < ke; k++) {
    const int n = k - kb;
    const fidx c = c00 + (fidx)k * Z;
    // the thread's own inputs from memory first, then the next planes
    T u0c = T(0), v0c = T(0), w0c = T(0);
    if (MODE == MOM_FROM_U0) { u0c = A.u0[c]; v0c = A.v0[c]; w0c = A.w0[c]; }
    const T tfu = A.tfu ? A.tfu[c] : T(0), tfv = A.tfv ? A.tfv[c] : T(0), tfw = A.tfw ? A.tfw[c] : T(0);
    const int kuc = A.ku && act ? A.ku[cp] : -9, kvc = A.kv && act ? A.kv[cp] : -9,
              kwc = A.kw && act ? A.kw[cp] : -9;
    if (k + 2 <= ke) load_ring(k + 2);
    if (k + 1 < ke) load_plane(k + 1);
    const T *Pm = &ring[n % 3][0][l], *P0 = &ring[(n + 1) % 3][0][l], *Pp = &ring[(n + 2) % 3][0][l];
    // ring field f at (i+a, j+b, k+d)
    auto R = [&](int f, int a, int b, int d) -> T {
      const T *B = d < 0 ? Pm : (d > 0 ? Pp : P0);
      return B[f * HS + b * LW + a];
    };
    auto u = [&](int a, int b, int d) { return R(0, a, b, d); };
    auto v = [&](int a, int b, int d) { return R(1, a, b, d); };
    auto w = [&](int a, int b, int d) { return R(2, a, b, d); };
    auto nu = [&](int a, int b, int d) { return R(3, a, b, d); };
    auto ln = [&](int a, int b, int d) { return R(4, a, b, d); };
    auto P = [&](int f, int a, int b) { return pk[f][l + b * LW + a]; };
    // face viscosity (ls_nu + nut interpolated to the face) of the face of
    // cell (a,b,d) along x, y, z: face_visc
    auto fvx = [&](int a, int b, int d) {
      return (ln(a, b, d) + ln(a - 1, b, d)) * T(0.5) + (nu(a, b, d) + nu(a - 1, b, d)) * T(0.5);
    };
    auto fvy = [&](int a, int b, int d) {
      return (ln(a, b, d) + ln(a, b - 1, d)) * T(0.5) + (nu(a, b, d) + nu(a, b - 1, d)) * T(0.5);
    };
    auto fvz = [&](int a, int b, int d) {
      return (ln(a, b, d) + ln(a, b, d - 1)) * T(0.5) + (nu(a, b, d) + nu(a, b, d - 1)) * T(0.5);
    };
    auto update = [&](T uc, T F, T u0v, const T *u0, T *u0o, T *un, T *fa, T *fb) {
      if (MODE == MOM_RK_FIRST) {
        u0o[c] = uc;
        un[c] = uc + A.dt * F;
      } else if (MODE == MOM_FROM_U0) {
        un[c] = u0v + A.dt * F;
      } else {
        stage_update<T>(A.mode, A.dt, c, uc, F, u0, u0o, un, fa, fb);
      }
    };
    const T rho0 = P(0, 0, 0), sz0 = P(3, 0, 0);
    const bool pl = (g.ccz0 == 2 && k == 2) || (g.ccy0 == 2 && j == 2) || (g.ccyf == 2 && j == ny);
    if (act) {
      // ---- u face (momentum_cell, adv_u with ADV = 11)
      {
        const T uc = u(0, 0, 0);
        const T bma_a = (v(0, 0, 0) + v(-1, 0, 0) + v(0, 1, 0) + v(-1, 1, 0)) * T(0.25);
        const T dma_a = (w(0, 0, 0) + w(-1, 0, 0) + w(0, 0, 1) + w(-1, 0, 1)) * T(0.25);
        T a1 = tmax(uc, T(0)), a2 = tmin(uc, T(0));
        const T dudx = a1 * (uc - u(-1, 0, 0)) / dx + a2 * (u(1, 0, 0) - uc) / dx;
        a1 = tmax(bma_a, T(0)); a2 = tmin(bma_a, T(0));
        const T dudy = a1 * (uc - u(0, -1, 0)) / dy + a2 * (u(0, 1, 0) - uc) / dy;
        a1 = tmax(dma_a, T(0)); a2 = tmin(dma_a, T(0));
        const T dudz = a1 * (uc - u(0, 0, -1)) / dz + a2 * (u(0, 0, 1) - uc) / dz;
        const T uint_ = dudx + dudy + dudz;
        const T amai = nu(0, 0, 0) / dx;
        const T amam = nu(am, 0, 0) / dx;  // cell_clamp of cell i-1
        const T bmap = ((fvx(0, 1, 0) + fvx(0, 0, 0)) * T(0.5)) / dy;
        const T bmac = ((fvx(0, 0, 0) + fvx(0, -1, 0)) * T(0.5)) / dy;
        const T dmap = ((fvx(0, 0, 1) + fvx(0, 0, 0)) * T(0.5)) / dz;
        const T dmac = ((fvx(0, 0, 0) + fvx(0, 0, -1)) * T(0.5)) / dz;
        const T rhox = (rho0 + P(0, -1, 0)) * T(0.5);
        const T dhsdx = (P(1, 0, 0) + P(1, -1, 0)) * T(0.5);
        const bool plu = A.ku ? k == kuc + 1 : pl;
        const T ub = plu ? uc : T(0);
        const T F = -uint_ +
                    ((amai * T(2.) * (u(1, 0, 0) - uc) - amam * T(2.) * (uc - u(-1, 0, 0))) / dx +
                     bmap * ((u(0, 1, 0) - uc) / dy + (v(0, 1, 0) - v(-1, 1, 0)) / dx) -
                     bmac * ((uc - u(0, -1, 0)) / dy + (v(0, 0, 0) - v(-1, 0, 0)) / dx) +
                     dmap * ((u(0, 0, 1) - uc) / dz + (w(0, 0, 1) - w(-1, 0, 1)) / dx) -
                     dmac * ((uc - u(0, 0, -1)) / dz + (w(0, 0, 0) - w(-1, 0, 0)) / dx)) /
                        rhox +
                    T(ph.gx) - sigma * dhsdx + tfu - fr * sqrt(ub * ub) * uc;
        if (A.Fu) A.Fu[c] = F;
        update(uc, F, u0c, A.u0, A.u0o, A.un, A.fu0, A.fu1);
      }
      // ---- v face (adv_v with ADV = 11)
      {
        const T vc = v(0, 0, 0);
        const T amb_a = (u(0, 0, 0) + u(1, 0, 0) + u(0, -1, 0) + u(1, -1, 0)) * T(0.25);
        const T dmb_a = (w(0, 0, 0) + w(0, -1, 0) + w(0, 0, 1) + w(0, -1, 1)) * T(0.25);
        T a1 = tmax(amb_a, T(0)), a2 = tmin(amb_a, T(0));
        const T dvdx = a1 * (vc - v(-1, 0, 0)) / dx + a2 * (v(1, 0, 0) - vc) / dx;
        a1 = tmax(vc, T(0)); a2 = tmin(vc, T(0));
        const T dvdy = a1 * (vc - v(0, -1, 0)) / dy + a2 * (v(0, 1, 0) - vc) / dy;
        a1 = tmax(dmb_a, T(0)); a2 = tmin(dmb_a, T(0));
        const T dvdz = a1 * (vc - v(0, 0, -1)) / dz + a2 * (v(0, 0, 1) - vc) / dz;
        const T vint_ = dvdx + dvdy + dvdz;
        const T ambp = ((fvy(1, 0, 0) + fvy(0, 0, 0)) * T(0.5)) / dx;
        const T ambc = ((fvy(0, 0, 0) + fvy(-1, 0, 0)) * T(0.5)) / dx;
        const T bmbj = nu(0, 0, 0) / dy;
        const T bmbm = nu(0, -1, 0) / dy;
        const T dmbp = ((fvy(0, 0, 1) + fvy(0, 0, 0)) * T(0.5)) / dz;
        const T dmbc = ((fvy(0, 0, 0) + fvy(0, 0, -1)) * T(0.5)) / dz;
        const T rhoy = (rho0 + P(0, 0, -1)) * T(0.5);
        const T dhsdy = (P(2, 0, 0) + P(2, 0, -1)) * T(0.5);
        const bool plv = A.kv ? k == kvc + 1 : pl;
        const T vb = plv ? vc : T(0);
        const T F = -vint_ +
                    (ambp * ((v(1, 0, 0) - vc) / dx + (u(1, 0, 0) - u(1, -1, 0)) / dy) -
                     ambc * ((vc - v(-1, 0, 0)) / dx + (u(0, 0, 0) - u(0, -1, 0)) / dy) +
                     (bmbj * T(2.) * (v(0, 1, 0) - vc) - bmbm * T(2.) * (vc - v(0, -1, 0))) / dy +
                     dmbp * ((v(0, 0, 1) - vc) / dz + (w(0, 0, 1) - w(0, -1, 1)) / dy) -
                     dmbc * ((vc - v(0, 0, -1)) / dz + (w(0, 0, 0) - w(0, -1, 0)) / dy)) /
                        rhoy -
                    sigma * dhsdy + tfv - fr * sqrt(vb * vb) * vc;
        if (A.Fv) A.Fv[c] = F;
        update(vc, F, v0c, A.v0, A.v0o, A.vn, A.fv0, A.fv1);
      }
      // ---- w face (adv_w with ADV = 11)
      {
        const T wc = w(0, 0, 0);
        const T amd_a = (u(0, 0, 0) + u(1, 0, 0) + u(0, 0, -1) + u(1, 0, -1)) * T(0.25);
        const T bmd_a = (v(0, 0, 0) + v(0, 1, 0) + v(0, 0, -1) + v(0, 1, -1)) * T(0.25);
        T a1 = tmax(amd_a, T(0)), a2 = tmin(amd_a, T(0));
        const T dwdx = a1 * (wc - w(-1, 0, 0)) / dx + a2 * (w(1, 0, 0) - wc) / dx;
        a1 = tmax(bmd_a, T(0)); a2 = tmin(bmd_a, T(0));
        const T dwdy = a1 * (wc - w(0, -1, 0)) / dy + a2 * (w(0, 1, 0) - wc) / dy;
        a1 = tmax(wc, T(0)); a2 = tmin(wc, T(0));
        const T dwdz = a1 * (wc - w(0, 0, -1)) / dz + a2 * (w(0, 0, 1) - wc) / dz;
        const T wint_ = dwdx + dwdy + dwdz;
        const T amdp = ((fvz(1, 0, 0) + fvz(0, 0, 0)) * T(0.5)) / dx;
        const T amdc = ((fvz(0, 0, 0) + fvz(-1, 0, 0)) * T(0.5)) / dx;
        const T bmdp = ((fvz(0, 1, 0) + fvz(0, 0, 0)) * T(0.5)) / dy;
        const T bmdc = ((fvz(0, 0, 0) + fvz(0, -1, 0)) * T(0.5)) / dy;
        const T dmdk = nu(0, 0, 0) / dz;
        const T dmdm = nu(0, 0, -1) / dz;
        const T rhoz = (rho0 + rho_m) * T(0.5);
        const T dhsdz = (sz0 + sz_m) * T(0.5);
        const bool plw = A.kw ? k == kwc + 1 : pl;
        const T wb = plw ? wc : T(0);
        const T epis_z = T(0);
        const T F = -wint_ +
                    (amdp * ((w(1, 0, 0) - wc) / dx + (u(1, 0, 0) - u(1, 0, -1)) / dz) -
                     amdc * ((wc - w(-1, 0, 0)) / dx + (u(0, 0, 0) - u(0, 0, -1)) / dz) +
                     bmdp * ((w(0, 1, 0) - wc) / dy + (v(0, 1, 0) - v(0, 1, -1)) / dz) -
                     bmdc * ((wc - w(0, -1, 0)) / dy + (v(0, 0, 0) - v(0, 0, -1)) / dz) +
                     (dmdk * T(2.) * (w(0, 0, 1) - wc) - dmdm * T(2.) * (wc - w(0, 0, -1))) / dz) /
                        rhoz -
                    (wc - T(0.)) * epis_z - T(ph.gz) - sigma * dhsdz + tfw - fr * sqrt(wb * wb) * wc;
        if (A.Fw) A.Fw[c] = F;
        update(wc, F, w0c, A.w0, A.w0o, A.wn, A.fw0, A.fw1);
      }
    }
    rho_m = rho0;
    sz_m = sz0;
    __syncthreads();  // plane k-1's slot and plane k's cell terms are free
    if (k + 2 <= ke) store_ring(n % 3);
    if (k + 1 < ke) store_plane();
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Boundary conditions contorno(nlock) (3_contorno.f90:29-443) with prd_corr
// (:872-950, whose corrections are all zero).  Wall codes: 0 periodic,
// 1 free slip, 2 no slip, 3 prescribed (ghosts and wall-normal faces from the
// boundary planes, :80-84,109-113,136-140,185-189,214-218,236-240 and their
// nlock = 2 counterparts), 4 outflow at x = xf and, at x = 0, the inflow of
// :143-146 (v from the plane bxy0, w copied).  The reference fills the
// ghost planes in the order y walls, x walls, z walls (pressure: x, y, z;
// plain copies commute), so each phase is one launch over its ghost planes.
// Bottom obstacle in contorno(1), contorno(2) (3_contorno.f90:266-291):
// the velocities of each column are zeroed up to k = ku(i,j) (kv, kw),
// for i, j in [0, n+1] and the extra face planes u(nx+2, :), v(:, ny+2).
// One thread per column.
template <typename T>
__global__ void k_bc_obst(Grid g, T *__restrict__ u, T *__restrict__ v, T *__restrict__ w,
                          const int *__restrict__ ku, const int *__restrict__ kv, const int *__restrict__ kw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = g.nx, ny = g.ny;
  if (i > nx + 2 || j > ny + 2) return;
  const int q = i + g.sx * j;
  if (j <= ny + 1) {  // u: i in [0, nx+2]
    for (int k = 0; k <= ku[q]; k++) u[g.idx(i, j, k)] = T(0);
  }
  if (i <= nx + 1) {  // v: j in [0, ny+2]
    for (int k = 0; k <= kv[q]; k++) v[g.idx(i, j, k)] = T(0);
  }
  if (i <= nx + 1 && j <= ny + 1) {
    for (int k = 0; k <= kw[q]; k++) w[g.idx(i, j, k)] = T(0);
  }
}

// contorno(3) with a bottom obstacle (3_contorno.f90:445-486): the level set
// of the cells below kw(i,j) from the neighbouring columns, in the
// reference's order of cases.  Every value read belongs to a cell outside
// the obstacle (a neighbour column lower than k, or the cell above), so the
// update is in place without races.
template <typename T>
__global__ void k_obst_ls(Grid g, T *__restrict__ ls, const int *__restrict__ kw) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const int sx = g.sx;
  const int q = i + sx * j;
  if (k > kw[q] - 1) return;
  const fidx c = g.idx(i, j, k), X = 1, Y = g.sx;
  const int wm = kw[q - 1], wp = kw[q + 1], sm = kw[q - sx], sp = kw[q + sx];
  T val = ls[c];
  if (wm < k && sm < k) val = (ls[c - X] + ls[c - Y]) * T(0.5);
  else if (wp < k && sm < k) val = (ls[c + X] + ls[c - Y]) * T(0.5);
  else if (wp < k && sp < k) val = (ls[c + X] + ls[c + Y]) * T(0.5);
  else if (wm < k && sp < k) val = (ls[c - X] + ls[c + Y]) * T(0.5);
  else if (wm < k && sm >= k && sp >= k) val = ls[c - X];
  else if (sm < k && wm >= k && wp >= k) val = ls[c - Y];
  else if (wp < k && sm >= k && sp >= k) val = ls[c + X];
  else if (sp < k && wm >= k && wp >= k) val = ls[c + Y];
  else if (k == kw[q] - 1 && k < g.nz) val = ls[c + g.sz];
  ls[c] = val;
}

template <typename T>
__global__ void k_bc_y(Grid g, T *__restrict__ u, T *__restrict__ v, T *__restrict__ w,
                       double *__restrict__ p, const T *__restrict__ bp, BPlanes L) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx + 2 || k > nz + 2) return;
  const bool per = g.ccy0 == 0 && g.ccyf == 0;
  if (i <= nx + 1 && k <= nz + 1) {
    if (per) { p[g.idx(i, 0, k)] = p[g.idx(i, ny, k)]; p[g.idx(i, ny + 1, k)] = p[g.idx(i, 1, k)]; }
    else { p[g.idx(i, 0, k)] = p[g.idx(i, 1, k)]; p[g.idx(i, ny + 1, k)] = p[g.idx(i, ny, k)]; }
  }
  if (k <= nz + 1) {  // u: i in [0,nx+2]
    if (per) u[g.idx(i, 0, k)] = u[g.idx(i, ny, k)];
    else if (g.ccy0 == 1) u[g.idx(i, 0, k)] = u[g.idx(i, 1, k)];
    else if (g.ccy0 == 2) u[g.idx(i, 0, k)] = -u[g.idx(i, 1, k)];
    else if (g.ccy0 == 3) u[g.idx(i, 0, k)] = bp[L.at(BYX0, i, k)];
    if (per) u[g.idx(i, ny + 1, k)] = u[g.idx(i, 1, k)];
    else if (g.ccyf == 1) u[g.idx(i, ny + 1, k)] = u[g.idx(i, ny, k)];
    else if (g.ccyf == 2) u[g.idx(i, ny + 1, k)] = -u[g.idx(i, ny, k)];
    else if (g.ccyf == 3) u[g.idx(i, ny + 1, k)] = bp[L.at(BYXF, i, k)];
  }
  if (i <= nx + 1 && k <= nz + 1) {  // v
    if (per) { v[g.idx(i, 0, k)] = v[g.idx(i, ny, k)]; v[g.idx(i, 1, k)] = v[g.idx(i, ny + 1, k)]; }
    else if (g.ccy0 == 1) v[g.idx(i, 0, k)] = -v[g.idx(i, 2, k)];
    else if (g.ccy0 == 2) v[g.idx(i, 0, k)] = v[g.idx(i, 2, k)];
    else if (g.ccy0 == 3) v[g.idx(i, 0, k)] = bp[L.at(BYY0, i, k)];
    if (per) v[g.idx(i, ny + 2, k)] = v[g.idx(i, 2, k)];
    else if (g.ccyf == 1) v[g.idx(i, ny + 2, k)] = -v[g.idx(i, ny, k)];
    else if (g.ccyf == 2) v[g.idx(i, ny + 2, k)] = v[g.idx(i, ny, k)];
    else if (g.ccyf == 3) v[g.idx(i, ny + 2, k)] = bp[L.at(BYYF1, i, k)];
  }
  if (i <= nx + 1) {  // w: k in [0,nz+2]
    if (per) w[g.idx(i, 0, k)] = w[g.idx(i, ny, k)];
    else if (g.ccy0 == 1) w[g.idx(i, 0, k)] = w[g.idx(i, 1, k)];
    else if (g.ccy0 == 2) w[g.idx(i, 0, k)] = -w[g.idx(i, 1, k)];
    else if (g.ccy0 == 3) w[g.idx(i, 0, k)] = bp[L.at(BYZ0, i, k)];
    if (per) w[g.idx(i, ny + 1, k)] = w[g.idx(i, 1, k)];
    else if (g.ccyf == 1) w[g.idx(i, ny + 1, k)] = w[g.idx(i, ny, k)];
    else if (g.ccyf == 2) w[g.idx(i, ny + 1, k)] = -w[g.idx(i, ny, k)];
    else if (g.ccyf == 3) w[g.idx(i, ny + 1, k)] = bp[L.at(BYZF, i, k)];
  }
}

template <typename T>
__global__ void k_bc_x(Grid g, T *__restrict__ u, T *__restrict__ v, T *__restrict__ w,
                       double *__restrict__ p, const T *__restrict__ bp, BPlanes L) {
  // blockIdx.z: field (p, u, v, w) x side (x = 0, x = xf).  Every write reads
  // only cells that no part writes, so the eight parts are independent; one
  // thread per row touches a cache line per row, and the split gives the
  // launch enough wavefronts to hide that latency.
  const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  const int fld = blockIdx.z >> 1, hi = blockIdx.z & 1;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (j > ny + 2 || k > nz + 2) return;
  const bool per = g.ccx0 == 0 && g.ccxf == 0;
  if (fld == 0 && j <= ny + 1 && k <= nz + 1) {
    if (!hi) p[g.idx(0, j, k)] = per ? p[g.idx(nx, j, k)] : p[g.idx(1, j, k)];
    else p[g.idx(nx + 1, j, k)] = per ? p[g.idx(1, j, k)] : p[g.idx(nx, j, k)];
  }
  if (fld == 1 && j <= ny + 1 && k <= nz + 1) {  // u: j in [0,ny+1], k in [0,nz+1]
    if (!hi) {
      if (per) u[g.idx(0, j, k)] = u[g.idx(nx, j, k)];
      else if (g.ccx0 == 1) u[g.idx(0, j, k)] = -u[g.idx(2, j, k)];
      else if (g.ccx0 == 2) u[g.idx(0, j, k)] = u[g.idx(2, j, k)];
      else if (g.ccx0 == 3) u[g.idx(0, j, k)] = bp[L.at(BXX0, j, k)];
    } else {
      if (per) u[g.idx(nx + 2, j, k)] = u[g.idx(2, j, k)];
      else if (g.ccxf == 1) u[g.idx(nx + 2, j, k)] = -u[g.idx(nx, j, k)];
      else if (g.ccxf == 2) u[g.idx(nx + 2, j, k)] = u[g.idx(nx, j, k)];
      else if (g.ccxf == 3) u[g.idx(nx + 2, j, k)] = bp[L.at(BXXF1, j, k)];
      else if (g.ccxf == 4) u[g.idx(nx + 2, j, k)] = u[g.idx(nx + 1, j, k)];
    }
  }
  if (fld == 2 && k <= nz + 1) {  // v: j in [0,ny+2]
    if (!hi) {
      if (per) v[g.idx(0, j, k)] = v[g.idx(nx, j, k)];
      else if (g.ccx0 == 1) v[g.idx(0, j, k)] = v[g.idx(1, j, k)];
      else if (g.ccx0 == 2) v[g.idx(0, j, k)] = -v[g.idx(1, j, k)];
      else if (g.ccx0 == 3 || g.ccx0 == 4) v[g.idx(0, j, k)] = bp[L.at(BXY0, j, k)];
    } else {
      if (per) v[g.idx(nx + 1, j, k)] = v[g.idx(1, j, k)];
      else if (g.ccxf == 1 || g.ccxf == 4) v[g.idx(nx + 1, j, k)] = v[g.idx(nx, j, k)];
      else if (g.ccxf == 2) v[g.idx(nx + 1, j, k)] = -v[g.idx(nx, j, k)];
      else if (g.ccxf == 3) v[g.idx(nx + 1, j, k)] = bp[L.at(BXYF, j, k)];
    }
  }
  if (fld == 3 && j <= ny + 1) {  // w: k in [0,nz+2]
    if (!hi) {
      if (per) w[g.idx(0, j, k)] = w[g.idx(nx, j, k)];
      else if (g.ccx0 == 1) w[g.idx(0, j, k)] = w[g.idx(1, j, k)];
      else if (g.ccx0 == 2) w[g.idx(0, j, k)] = -w[g.idx(1, j, k)];
      else if (g.ccx0 == 3) w[g.idx(0, j, k)] = bp[L.at(BXZ0, j, k)];
      else if (g.ccx0 == 4) w[g.idx(0, j, k)] = w[g.idx(1, j, k)];
    } else {
      if (per) w[g.idx(nx + 1, j, k)] = w[g.idx(1, j, k)];
      else if (g.ccxf == 1 || g.ccxf == 4) w[g.idx(nx + 1, j, k)] = w[g.idx(nx, j, k)];
      else if (g.ccxf == 2) w[g.idx(nx + 1, j, k)] = -w[g.idx(nx, j, k)];
      else if (g.ccxf == 3) w[g.idx(nx + 1, j, k)] = bp[L.at(BXZF, j, k)];
    }
  }
}

template <typename T>
__global__ void k_bc_z(Grid g, T *__restrict__ u, T *__restrict__ v, T *__restrict__ w,
                       double *__restrict__ p, const T *__restrict__ bp, BPlanes L) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx + 2 || j > ny + 2) return;
  if (i <= nx + 1 && j <= ny + 1) {
    p[g.idx(i, j, 0)] = p[g.idx(i, j, 1)];
    p[g.idx(i, j, nz + 1)] = p[g.idx(i, j, nz)];
  }
  if (j <= ny + 1) {  // u: i in [0,nx+2]
    if (g.ccz0 == 1) u[g.idx(i, j, 0)] = u[g.idx(i, j, 1)];
    else if (g.ccz0 == 2) u[g.idx(i, j, 0)] = -u[g.idx(i, j, 1)];
    else if (g.ccz0 == 3) u[g.idx(i, j, 0)] = bp[L.at(BZX0, i, j)];
    if (g.cczf == 1) u[g.idx(i, j, nz + 1)] = u[g.idx(i, j, nz)];
    else if (g.cczf == 2) u[g.idx(i, j, nz + 1)] = -u[g.idx(i, j, nz)];
    else if (g.cczf == 3) u[g.idx(i, j, nz + 1)] = bp[L.at(BZXF, i, j)];
  }
  if (i <= nx + 1) {  // v: j in [0,ny+2]
    if (g.ccz0 == 1) v[g.idx(i, j, 0)] = v[g.idx(i, j, 1)];
    else if (g.ccz0 == 2) v[g.idx(i, j, 0)] = -v[g.idx(i, j, 1)];
    else if (g.ccz0 == 3) v[g.idx(i, j, 0)] = bp[L.at(BZY0, i, j)];
    if (g.cczf == 1) v[g.idx(i, j, nz + 1)] = v[g.idx(i, j, nz)];
    else if (g.cczf == 2) v[g.idx(i, j, nz + 1)] = -v[g.idx(i, j, nz)];
    else if (g.cczf == 3) v[g.idx(i, j, nz + 1)] = bp[L.at(BZYF, i, j)];
  }
  if (i <= nx + 1 && j <= ny + 1) {  // w
    if (g.ccz0 == 1) w[g.idx(i, j, 0)] = -w[g.idx(i, j, 2)];
    else if (g.ccz0 == 2) w[g.idx(i, j, 0)] = w[g.idx(i, j, 2)];
    else if (g.ccz0 == 3) w[g.idx(i, j, 0)] = bp[L.at(BZZ0, i, j)];
    if (g.cczf == 1) w[g.idx(i, j, nz + 2)] = -w[g.idx(i, j, nz)];
    else if (g.cczf == 2) w[g.idx(i, j, nz + 2)] = w[g.idx(i, j, nz)];
    else if (g.cczf == 3) w[g.idx(i, j, nz + 2)] = bp[L.at(BZZF1, i, j)];
  }
}

// contorno(2) wall-normal velocities (3_contorno.f90:297-443), in three
// launches: y walls, x walls (with the ccxf=2 ghost refresh), z walls.
template <typename T>
__global__ void k_bcn_y(Grid g, T *__restrict__ v, const T *__restrict__ bp, BPlanes L) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > g.nx + 1 || k > g.nz + 1) return;
  const bool per = g.ccy0 == 0 && g.ccyf == 0;
  if (per) v[g.idx(i, 1, k)] = v[g.idx(i, g.ny + 1, k)];
  else if (g.ccy0 == 1 || g.ccy0 == 2) v[g.idx(i, 1, k)] = T(0.) + T(0.);
  else if (g.ccy0 == 3) v[g.idx(i, 1, k)] = bp[L.at(BYY1, i, k)] + T(0.);
  if (per) v[g.idx(i, g.ny + 1, k)] = v[g.idx(i, 1, k)];
  else if (g.ccyf == 1 || g.ccyf == 2) v[g.idx(i, g.ny + 1, k)] = T(0.) + T(0.);
  else if (g.ccyf == 3) v[g.idx(i, g.ny + 1, k)] = bp[L.at(BYYF, i, k)] + T(0.);
}
template <typename T>
__global__ void k_bcn_x(Grid g, T *__restrict__ u, T *__restrict__ v, T *__restrict__ w,
                        const T *__restrict__ bp, BPlanes L) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = g.nx;
  if (j > g.ny + 2 || k > g.nz + 2) return;
  const bool per = g.ccx0 == 0 && g.ccxf == 0;
  if (j <= g.ny + 1 && k <= g.nz + 1) {
    if (per) u[g.idx(1, j, k)] = u[g.idx(nx + 1, j, k)];
    else if (g.ccx0 == 1 || g.ccx0 == 2) u[g.idx(1, j, k)] = T(0.) + T(0.);
    else if (g.ccx0 == 3) u[g.idx(1, j, k)] = bp[L.at(BXX1, j, k)] + T(0.);
    if (per) u[g.idx(nx + 1, j, k)] = u[g.idx(1, j, k)];
    else if (g.ccxf == 1) u[g.idx(nx + 1, j, k)] = T(0.) + T(0.);
    else if (g.ccxf == 3) u[g.idx(nx + 1, j, k)] = bp[L.at(BXXF, j, k)] + T(0.);
    else if (g.ccxf == 2) {
      u[g.idx(nx + 1, j, k)] = T(0.) + T(0.);
      u[g.idx(nx + 2, j, k)] = u[g.idx(nx, j, k)];
    } else if (g.ccxf == 4) u[g.idx(nx + 1, j, k)] = u[g.idx(nx, j, k)];
  }
  if (g.ccxf == 2) {
    if (k <= g.nz + 1) v[g.idx(nx + 1, j, k)] = -v[g.idx(nx, j, k)];
    if (j <= g.ny + 1) w[g.idx(nx + 1, j, k)] = -w[g.idx(nx, j, k)];
  }
}
template <typename T>
__global__ void k_bcn_z(Grid g, T *__restrict__ w, const T *__restrict__ bp, BPlanes L) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > g.nx + 1 || j > g.ny + 1) return;
  if (g.ccz0 == 1 || g.ccz0 == 2) w[g.idx(i, j, 1)] = T(0.) + T(0.);
  else if (g.ccz0 == 3) w[g.idx(i, j, 1)] = bp[L.at(BZZ1, i, j)] + T(0.);
  if (g.cczf == 1 || g.cczf == 2) w[g.idx(i, j, g.nz + 1)] = T(0.) + T(0.);
  else if (g.cczf == 3) w[g.idx(i, j, g.nz + 1)] = bp[L.at(BZZF, i, j)] + T(0.);
}

// ---------------------------------------------------------------------------
// Velocity correction posdin (10_pos.f90:17-45): u -= (dt/h)*(p_i - p_{i-1})/rho_face
// over every face including the walls.  The pressure is FP64; the update is
// done in FP64 and rounded to T.
template <typename T>
__global__ __launch_bounds__(256) void k_posdin(Grid g, double dt, const double *__restrict__ p,
                                                const T *__restrict__ rho, T *__restrict__ u,
                                                T *__restrict__ v, T *__restrict__ w) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx + 1 || j > ny + 1) return;
  const fidx c = g.idx(i, j, k);
  const double a1 = dt / g.dx, a2 = dt / g.dy, a3 = dt / g.dz;
  if (j <= ny && k <= nz) {
    const double rx = (double)interp_face<T>(rho, c, 1, i, nx);
    u[c] = (T)((double)u[c] - a1 * (p[c] - p[c - 1]) / rx);
  }
  if (i <= nx && k <= nz) {
    const double ry = (double)interp_face<T>(rho, c, g.sx, j, ny);
    v[c] = (T)((double)v[c] - a2 * (p[c] - p[c - g.sx]) / ry);
  }
  if (i <= nx && j <= ny) {
    const double rz = (double)interp_face<T>(rho, c, g.sz, k, nz);
    w[c] = (T)((double)w[c] - a3 * (p[c] - p[c - g.sz]) / rz);
  }
}

// max abs(div u) over the cells (est, 4_plot.f90:509-514), FP64.
template <typename T>
__global__ __launch_bounds__(256) void k_divmax(Grid g, const T *__restrict__ u,
                                                const T *__restrict__ v, const T *__restrict__ w,
                                                RedSlot rs) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  double acc[1] = {0.0};
  if (i <= g.nx && j <= g.ny) {
    const fidx c = g.idx(i, j, k);
    const double d = ((double)u[c + 1] - (double)u[c]) / g.dx + ((double)v[c + g.sx] - (double)v[c]) / g.dy +
                     ((double)w[c + g.sz] - (double)w[c]) / g.dz;
    acc[0] = d < 0 ? -d : d;
  }
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1, true>(acc, lds);
  block_partial<1>(acc, rs.partials);
}

// Field output of plot_f (4_plot.f90:266-213 of the t_plot branches): the one
// record of an arquivos/campos_NNNNN file, FP32, in the reference's order
// and extents (Fortran order within each array):
//   x11, y11, z11, uaux, vaux, waux        (nx+1, ny+1, nz+1) nodes
//   ls, kaux                               (nx, ny, nz)
//   t_plot = 1 adds prdaux, vorti, vortj, vortk (nx, ny, nz),
//                   xnuta (nx+1, ny, nz), ynuta (nx, ny+1, nz), znuta (nx, ny, nz+1).
// Node velocities average the four faces around the node's edge
// (4_plot.f90:269-272), vorticity takes forward differences (:190-201), the
// face viscosities interp*_cf(nut) with the end faces copied (:206-208,
// 8_derivadas.f90:306-323); kaux marks the cells below kw(i,j) of a bottom
// obstacle (0 without one); the coordinates come
// from per-axis tables (the reference's running sums, rounded to FP32).
// One thread per node; every value is computed in T and rounded once.
template <typename T>
__global__ __launch_bounds__(256) void k_frame_record(Grid g, int t_plot, const T *__restrict__ u,
                                                      const T *__restrict__ v, const T *__restrict__ w,
                                                      const T *__restrict__ ls, const double *__restrict__ p,
                                                      const T *__restrict__ nut, const int *__restrict__ kw,
                                                      const float *__restrict__ xs,
                                                      const float *__restrict__ ys, const float *__restrict__ zs,
                                                      float *__restrict__ out) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx + 1 || j > ny + 1 || k > nz + 1) return;
  const int64_t n1 = (int64_t)(nx + 1) * (ny + 1) * (nz + 1), n = (int64_t)nx * ny * nz;
  const int64_t f1 = (i - 1) + (int64_t)(nx + 1) * ((j - 1) + (int64_t)(ny + 1) * (k - 1));
  const fidx c = g.idx(i, j, k);
  const fidx X = 1, Y = g.sx, Z = g.sz;
  const T q = (T)0.25;
  out[f1] = xs[i];
  out[n1 + f1] = ys[j];
  out[2 * n1 + f1] = zs[k];
  out[3 * n1 + f1] = (float)((((u[c] + u[c - Y]) + u[c - Z]) + u[c - Y - Z]) * q);
  out[4 * n1 + f1] = (float)((((v[c] + v[c - X]) + v[c - Z]) + v[c - X - Z]) * q);
  out[5 * n1 + f1] = (float)((((w[c] + w[c - Y]) + w[c - X]) + w[c - X - Y]) * q);
  float *o = out + 6 * n1;
  if (i <= nx && j <= ny && k <= nz) {
    const int64_t f = (i - 1) + (int64_t)nx * ((j - 1) + (int64_t)ny * (k - 1));
    o[f] = (float)ls[c];
    o[n + f] = (kw && k <= kw[i + g.sx * j] - 1) ? 1.f : 0.f;  // kaux, 4_plot.f90:285-297
    if (t_plot == 1) {
      o[2 * n + f] = (float)p[c];
      const T dudy = (u[c + Y] - u[c]) / (T)g.dy, dudz = (u[c + Z] - u[c]) / (T)g.dz;
      const T dvdx = (v[c + X] - v[c]) / (T)g.dx, dvdz = (v[c + Z] - v[c]) / (T)g.dz;
      const T dwdx = (w[c + X] - w[c]) / (T)g.dx, dwdy = (w[c + Y] - w[c]) / (T)g.dy;
      o[3 * n + f] = (float)(dwdy - dvdz);
      o[4 * n + f] = (float)(dudz - dwdx);
      o[5 * n + f] = (float)(dvdx - dudy);
    }
  }
  if (t_plot != 1) return;
  const T h = (T)0.5;
  float *ox = o + 6 * n, *oy = ox + (int64_t)(nx + 1) * ny * nz, *oz = oy + (int64_t)nx * (ny + 1) * nz;
  if (j <= ny && k <= nz) {  // xnuta(i), i = 1..nx+1: the face between cells ia-1 and ia
    const int ia = i < 2 ? 2 : (i > nx ? nx : i);
    const fidx ca = g.idx(ia, j, k);
    ox[(i - 1) + (int64_t)(nx + 1) * ((j - 1) + (int64_t)ny * (k - 1))] = (float)((nut[ca] + nut[ca - X]) * h);
  }
  if (i <= nx && k <= nz) {
    const int ja = j < 2 ? 2 : (j > ny ? ny : j);
    const fidx ca = g.idx(i, ja, k);
    oy[(i - 1) + (int64_t)nx * ((j - 1) + (int64_t)(ny + 1) * (k - 1))] = (float)((nut[ca] + nut[ca - Y]) * h);
  }
  if (i <= nx && j <= ny) {
    const int ka = k < 2 ? 2 : (k > nz ? nz : k);
    const fidx ca = g.idx(i, j, ka);
    oz[(i - 1) + (int64_t)nx * ((j - 1) + (int64_t)ny * (k - 1))] = (float)((nut[ca] + nut[ca - Z]) * h);
  }
}
