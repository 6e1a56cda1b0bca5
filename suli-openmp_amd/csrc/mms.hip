// The manufactured solution of Weng (2009) (13_mms.f90): initial fields
// (mms_i, :12-212), the time-dependent boundary planes of the
// prescribed-velocity walls (mms_bc, :468-816), the momentum sources
// (termo_fonte1 / termo_fonte2, :819-1322) and the error norms (mms,
// :218-465), all on the device.
//
// The solution is a free-surface flow over a depth h: level set h - z and
// velocities built from sin/cos products of x, y and z/h.
//   mms_t = 1 (steady): h = a sin x sin y + h0,
//     u = sin x cos y sin(pi z/2h) - cos x sin y (cos(2 pi z/h) - 1)
//     v = -cos x sin y sin(pi z/2h) + sin x cos y (cos(2 pi z/h) - 1)
//     w = -a (sin^2x cos^2y - cos^2x sin^2y)(sin(2 pi z/h) - (2 pi z/h) cos(2 pi z/h) + 2 pi)/(2 pi)
//   mms_t = 2 (unsteady), X = x - pi, Y = y - pi: h = a cos(X/2) cos(Y/2) cos t + h0,
//     c = cos(pi/2 (1 - z/h)) sin t, u = sin Y cos^2(X/2) c, v = -sin X cos^2(Y/2) c,
//     w = -a cos(X/2) cos(Y/2) c, p = -coef cos X cos Y cos(pi z/2h)
//
// The sources are the exact residual of the momentum equation that convdiff
// discretises (6_convdiff.f90:150-186) for these fields:
//   tf_u = [du/dt] + u u_x + v u_y + w u_z - nu (2 u_xx + u_yy + u_zz + v_xy + w_xz)
// (v, w alike), nu = 0.00100798/998 as in 13_mms.f90:841.  Their first and
// second derivatives come from second-order Taylor jets: the field
// expressions are evaluated once in jet arithmetic, which carries the value,
// the gradient and the Hessian through every operation exactly (up to
// rounding).  The reference's hand-derived expressions equal these for
// a = 0, its setting (0_modules.f90:208); for a != 0 its tf_u (steady) and
// its time derivatives (unsteady: the dependence of h on t is left out)
// differ from the exact ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "ctx.hpp"

#define NEEDCTX()                                              \
  do {                                                         \
    if (!g_ctx) {                                              \
      suli_set_error("library not initialised (suli_init)");   \
      return 2;                                                \
    }                                                          \
  } while (0)

namespace {

struct MmsPar {
  int mms_t;
  double a, h0, coef;
};

// -- second-order Taylor jets in N variables ------------------------------------
template <int N>
struct Jet {
  static constexpr int H = N * (N + 1) / 2;
  double v, g[N], h[H];  // value, gradient, Hessian (upper triangle, row major)
};
template <int N>
__host__ __device__ constexpr int hix(int i, int j) { return i * N - i * (i - 1) / 2 + (j - i); }  // i <= j

template <int N>
__host__ __device__ inline Jet<N> jconst(double c) {
  Jet<N> r;
  r.v = c;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = 0.0;
#pragma unroll
  for (int i = 0; i < Jet<N>::H; i++) r.h[i] = 0.0;
  return r;
}
template <int N>
__host__ __device__ inline Jet<N> jvar(int k, double x) {
  Jet<N> r = jconst<N>(x);
  r.g[k] = 1.0;
  return r;
}
template <int N>
__host__ __device__ inline Jet<N> operator+(const Jet<N> &a, const Jet<N> &b) {
  Jet<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = a.g[i] + b.g[i];
#pragma unroll
  for (int i = 0; i < Jet<N>::H; i++) r.h[i] = a.h[i] + b.h[i];
  return r;
}
template <int N>
__host__ __device__ inline Jet<N> operator*(double s, const Jet<N> &a) {
  Jet<N> r;
  r.v = s * a.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = s * a.g[i];
#pragma unroll
  for (int i = 0; i < Jet<N>::H; i++) r.h[i] = s * a.h[i];
  return r;
}
template <int N>
__host__ __device__ inline Jet<N> operator-(const Jet<N> &a, const Jet<N> &b) { return a + (-1.0) * b; }
template <int N>
__host__ __device__ inline Jet<N> operator+(const Jet<N> &a, double c) {
  Jet<N> r = a;
  r.v += c;
  return r;
}
template <int N>
__host__ __device__ inline Jet<N> operator*(const Jet<N> &a, const Jet<N> &b) {
  Jet<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = a.v * b.g[i] + b.v * a.g[i];
#pragma unroll
  for (int i = 0; i < N; i++)
#pragma unroll
    for (int j = i; j < N; j++)
      r.h[hix<N>(i, j)] = a.v * b.h[hix<N>(i, j)] + b.v * a.h[hix<N>(i, j)] + a.g[i] * b.g[j] + a.g[j] * b.g[i];
  return r;
}
// f(a) from f, f', f'' at a.v
template <int N>
__host__ __device__ inline Jet<N> chain(const Jet<N> &a, double f0, double f1, double f2) {
  Jet<N> r;
  r.v = f0;
#pragma unroll
  for (int i = 0; i < N; i++) r.g[i] = f1 * a.g[i];
#pragma unroll
  for (int i = 0; i < N; i++)
#pragma unroll
    for (int j = i; j < N; j++) r.h[hix<N>(i, j)] = f1 * a.h[hix<N>(i, j)] + f2 * a.g[i] * a.g[j];
  return r;
}
template <int N>
__host__ __device__ inline Jet<N> jsin(const Jet<N> &a) {
  const double s = sin(a.v), c = cos(a.v);
  return chain(a, s, c, -s);
}
template <int N>
__host__ __device__ inline Jet<N> jcos(const Jet<N> &a) {
  const double s = sin(a.v), c = cos(a.v);
  return chain(a, c, -s, -c);
}
template <int N>
__host__ __device__ inline Jet<N> operator/(const Jet<N> &a, const Jet<N> &b) {
  const double r = 1.0 / b.v;
  return a * chain(b, r, -r * r, 2.0 * r * r * r);
}
// the same field expressions in plain doubles
__host__ __device__ inline double jsin(double a) { return sin(a); }
__host__ __device__ inline double jcos(double a) { return cos(a); }

// -- the manufactured fields ------------------------------------------------------
// R: double or Jet<N>; x, y, z, t the point (jets: the variables).
template <typename R>
__host__ __device__ inline R weng_h(const MmsPar &m, const R &x, const R &y, const R &t) {
  if (m.mms_t == 1) return m.a * (jsin(x) * jsin(y)) + m.h0;
  const R X = x + (-M_PI), Y = y + (-M_PI);
  return m.a * (jcos(0.5 * X) * jcos(0.5 * Y) * jcos(t)) + m.h0;
}
// component c (0 u, 1 v, 2 w) of the velocity
template <typename R>
__host__ __device__ inline R weng_vel(const MmsPar &m, int c, const R &x, const R &y, const R &z, const R &t) {
  const R h = weng_h(m, x, y, t);
  const R zh = z / h;
  if (m.mms_t == 1) {
    const R sx = jsin(x), cx = jcos(x), sy = jsin(y), cy = jcos(y);
    const R s4 = jsin((0.5 * M_PI) * zh), c2 = jcos((2.0 * M_PI) * zh) + (-1.0);
    if (c == 0) return sx * cy * s4 - cx * sy * c2;
    if (c == 1) return (-1.0) * (cx * sy * s4) + sx * cy * c2;
    const R r = (2.0 * M_PI) * zh;
    return (-m.a / (2.0 * M_PI)) * ((sx * sx * cy * cy - cx * cx * sy * sy) * (jsin(r) - r * jcos(r) + 2.0 * M_PI));
  }
  const R X = x + (-M_PI), Y = y + (-M_PI);
  const R cc = jcos((0.5 * M_PI) * ((-1.0) * zh + 1.0)) * jsin(t);
  if (c == 0) {
    const R q = jcos(0.5 * X);
    return jsin(Y) * q * q * cc;
  }
  if (c == 1) {
    const R q = jcos(0.5 * Y);
    return (-1.0) * (jsin(X) * q * q * cc);
  }
  return (-m.a) * (jcos(0.5 * X) * jcos(0.5 * Y) * cc);
}
__host__ __device__ inline double weng_p(const MmsPar &m, double x, double y, double z, double t) {
  const double X = x - M_PI, Y = y - M_PI;
  return -m.coef * cos(X) * cos(Y) * cos(0.5 * M_PI * z / weng_h(m, x, y, t));
}

// point of face/cell index (i, j, k) of the staggered grid of component c
// (0 u, 1 v, 2 w, 3 cell centre): (i - 1) dx along the component's own axis,
// (i - 0.5) dx along the others
__host__ __device__ inline void stag_point(const Grid &g, int c, int i, int j, int k, double &x, double &y, double &z) {
  x = (i - (c == 0 ? 1.0 : 0.5)) * g.dx;
  y = (j - (c == 1 ? 1.0 : 0.5)) * g.dy;
  z = (k - (c == 2 ? 1.0 : 0.5)) * g.dz;
}

// -- kernels ------------------------------------------------------------------------
// mms_i: level set, interior-face velocities and (mms_t = 2) the pressure.
// One thread per (i, j, k) in [1, n+1]^3.
template <typename T>
__global__ void k_mms_init(Grid g, MmsPar m, double t, T *ls, T *u, T *v, T *w, double *prd1) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx + 1 || j > ny + 1 || k > nz + 1) return;
  const fidx c = g.idx(i, j, k);
  double x, y, z;
  if (j <= ny && k <= nz) {
    stag_point(g, 0, i, j, k, x, y, z);
    u[c] = (T)weng_vel<double>(m, 0, x, y, z, t);
  }
  if (i <= nx && k <= nz) {
    stag_point(g, 1, i, j, k, x, y, z);
    v[c] = (T)weng_vel<double>(m, 1, x, y, z, t);
  }
  if (i <= nx && j <= ny) {
    stag_point(g, 2, i, j, k, x, y, z);
    w[c] = (T)weng_vel<double>(m, 2, x, y, z, t);
  }
  if (i <= nx && j <= ny && k <= nz) {
    stag_point(g, 3, i, j, k, x, y, z);
    ls[c] = (T)(weng_h<double>(m, x, y, t) - z);
    if (m.mms_t == 2) prd1[c] = weng_p(m, x, y, z, t);
  }
}

// mms_bc: the 24 planes at time t.  blockIdx.z = plane; each covers the
// in-plane indices 0..n+1 of both axes (the loops of 13_mms.f90:485-486,
// :591-592, :683-684); the extra row of a plane's face axis stays as it was.
// In the unsteady case bzy0 and bzyf take the formula of w at the points of
// v (13_mms.f90:777-781, :795-799).
template <typename T>
__global__ void k_mms_planes(Grid g, MmsPar m, double t, T *bp, BPlanes L) {
  const int n = blockIdx.z;
  const int a = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y * blockDim.y + threadIdx.y;
  const int wall = bplane_wall(n), comp = bplane_comp(n);
  const int nn[3] = {g.nx, g.ny, g.nz};
  const int ax0 = wall == 0 ? 1 : 0, ax1 = wall == 2 ? 1 : 2;
  if (a > nn[ax0] + 1 || b > nn[ax1] + 1) return;
  // normal index of the plane within its wall group (BPlane order)
  const int p = n % 8, nw = nn[wall];
  const int qx[8] = {0, 1, 0, 0, nw + 1, nw + 2, nw + 1, nw + 1};
  const int qy[8] = {0, 0, 1, 0, nw + 1, nw + 1, nw + 2, nw + 1};
  const int qz[8] = {0, 0, 0, 1, nw + 1, nw + 1, nw + 1, nw + 2};
  const int q = wall == 0 ? qx[p] : wall == 1 ? qy[p] : qz[p];
  int idx[3];
  idx[wall] = q;
  idx[ax0] = a;
  idx[ax1] = b;
  double x, y, z;
  stag_point(g, comp, idx[0], idx[1], idx[2], x, y, z);
  const int f = (m.mms_t == 2 && (n == BZY0 || n == BZYF)) ? 2 : comp;
  bp[L.at(n, a, b)] = (T)weng_vel<double>(m, f, x, y, z, t);
}

// termo_fonte1 / termo_fonte2: the source of component c on its face grid
// (tf_u(1:nx+1,1:ny,1:nz), ...).  mms_t = 1 takes jets in (x, y, z),
// mms_t = 2 in (x, y, z, t).
template <typename T, int N>
__global__ void k_mms_src(Grid g, MmsPar m, double t, int c, T *out) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.nx + (c == 0) || j > g.ny + (c == 1) || k > g.nz + (c == 2)) return;
  double x, y, z;
  stag_point(g, c, i, j, k, x, y, z);
  const Jet<N> X = jvar<N>(0, x), Y = jvar<N>(1, y), Z = jvar<N>(2, z);
  const Jet<N> Tt = N > 3 ? jvar<N>(N - 1, t) : jconst<N>(t);
  Jet<N> V[3];
#pragma unroll
  for (int e = 0; e < 3; e++) V[e] = weng_vel(m, e, X, Y, Z, Tt);
  const Jet<N> &q = V[c];
  double adv = 0.0, visc = 0.0;
#pragma unroll
  for (int e = 0; e < 3; e++) {
    adv += V[e].v * q.g[e];
    visc += q.h[hix<N>(e, e)] + V[e].h[c <= e ? hix<N>(c, e) : hix<N>(e, c)];
  }
  double src = adv - (0.00100798 / 998.) * visc;
  if (N > 3) src += q.g[N - 1];
  out[g.idx(i, j, k)] = (T)src;
}

// mms(): squared errors against the exact solution at time t, over the
// faces whose level set interpolated to the face (end faces copied,
// 8_derivadas.f90:304-434) is positive and, for mms_t = 2, over the cells
// with positive level set, where erro_p (nx*ny*nz, persistent like the
// reference's module array) is rewritten; the sum of erro_p runs over every
// cell.  Per block: sums and counts of u, v, w, p (8 values).
template <typename T>
__global__ __launch_bounds__(256) void k_mms_err(Grid g, MmsPar m, double t, const T *ls, const T *u, const T *v,
                                                 const T *w, const double *prd1, double *erro_p, double *part) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int i = 1 + blockIdx.x * 64 + (threadIdx.x & 63), j = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
  const int k = 1 + blockIdx.z;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (i <= nx + 1 && j <= ny + 1 && k <= nz + 1) {
    auto L = [&](int ii, int jj, int kk) { return (double)ls[g.idx(ii, jj, kk)]; };
    auto cl = [](int f, int n) { return f < 2 ? 2 : (f > n ? n : f); };
    double x, y, z;
    if (j <= ny && k <= nz) {
      const int f = cl(i, nx);
      if ((L(f - 1, j, k) + L(f, j, k)) * 0.5 > 0.0) {
        stag_point(g, 0, i, j, k, x, y, z);
        const double d = weng_vel<double>(m, 0, x, y, z, t) - (double)u[g.idx(i, j, k)];
        acc[0] = d * d;
        acc[1] = 1.0;
      }
    }
    if (i <= nx && k <= nz) {
      const int f = cl(j, ny);
      if ((L(i, f - 1, k) + L(i, f, k)) * 0.5 > 0.0) {
        stag_point(g, 1, i, j, k, x, y, z);
        const double d = weng_vel<double>(m, 1, x, y, z, t) - (double)v[g.idx(i, j, k)];
        acc[2] = d * d;
        acc[3] = 1.0;
      }
    }
    if (i <= nx && j <= ny) {
      const int f = cl(k, nz);
      if ((L(i, j, f - 1) + L(i, j, f)) * 0.5 > 0.0) {
        stag_point(g, 2, i, j, k, x, y, z);
        const double d = weng_vel<double>(m, 2, x, y, z, t) - (double)w[g.idx(i, j, k)];
        acc[4] = d * d;
        acc[5] = 1.0;
      }
    }
    if (m.mms_t == 2 && i <= nx && j <= ny && k <= nz) {
      const int64_t q = (i - 1) + (int64_t)nx * ((j - 1) + (int64_t)ny * (k - 1));
      if (L(i, j, k) > 0.0) {
        stag_point(g, 3, i, j, k, x, y, z);
        const double d = weng_p(m, x, y, z, t) - prd1[g.idx(i, j, k)];
        erro_p[q] = d * d;
        acc[7] = 1.0;
      }
      acc[6] = erro_p[q];
    }
  }
  __shared__ double red[8][256];
#pragma unroll
  for (int q = 0; q < 8; q++) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q < 8; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < 8) {
    const int64_t blk = blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z);
    part[blk * 8 + threadIdx.x] = red[threadIdx.x][0];
  }
}

// sum of the block partials in block order, one thread per quantity
__global__ void k_mms_err_finish(const double *part, int64_t nblk, double *out) {
  const int q = threadIdx.x;
  if (q >= 8) return;
  double s = 0.0;
  for (int64_t b = 0; b < nblk; b++) s += part[b * 8 + q];
  out[q] = s;
}

bool mms_on(const Ctx &c) { return c.mms.t > 0; }

MmsPar par(const Ctx &c) { return MmsPar{c.mms.t, c.mms.a, c.mms.h0, c.mms.coef}; }

template <typename T>
int launch_planes(Ctx &c, double t) {
  const BPlanes L = bplanes_layout(c.g.nx, c.g.ny, c.g.nz);
  const int mx = std::max(c.g.nx, std::max(c.g.ny, c.g.nz)) + 2;
  hipLaunchKernelGGL(k_mms_planes<T>, dim3((mx + 63) / 64, (mx + 3) / 4, BP_N), dim3(64, 4, 1), 0, c.st, c.g,
                     par(c), t, (T *)c.bp, L);
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename T>
int launch_src(Ctx &c, double t) {
  const Grid &g = c.g;
  void *out[3] = {c.tfu, c.tfv, c.tfw};
  for (int comp = 0; comp < 3; comp++) {
    const dim3 grid((g.nx + 1 + 63) / 64, (g.ny + 1 + 3) / 4, g.nz + (comp == 2));
    if (c.mms.t == 1)
      hipLaunchKernelGGL((k_mms_src<T, 3>), grid, dim3(64, 4, 1), 0, c.st, g, par(c), t, comp, (T *)out[comp]);
    else
      hipLaunchKernelGGL((k_mms_src<T, 4>), grid, dim3(64, 4, 1), 0, c.st, g, par(c), t, comp, (T *)out[comp]);
    HIPCHK(hipGetLastError());
  }
  return 0;
}

}  // namespace

// plane helpers used by api.hip
int mms_planes_now(Ctx &c) {
  if (!mms_on(c)) return 0;
  if (c.mms.t == 1 && c.t_now != 0.0) return 0;  // the steady planes are set at t = 0 only (13_mms.f90:483)
  int rc = 0;
  DISPATCH(c, rc = launch_planes<TT>(c, c.t_now));
  return rc;
}

int mms_sources_now(Ctx &c) {
  if (!mms_on(c)) return 0;
  int rc = 0;
  DISPATCH(c, rc = launch_src<TT>(c, c.t_now));
  if (rc) return rc;
  c.tf_on = true;
  c.mms.src_done = true;
  return 0;
}

// Manufactured-solution mode: mms_t of 0_modules.f90:37 (0 off, 1 steady,
// 2 unsteady) with the constants a, h0, coef of module mms_m
// (0_modules.f90:208-212).  Resets the error array of the pressure and the
// "sources computed" flag of the steady case (termo_fonte1 runs at it = 1).
extern "C" int suli_mms_set(int mms_t, double a, double h0, double coef) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  if (mms_t < 0 || mms_t > 2) {
    suli_set_error("suli_mms_set: mms_t must be 0 (off), 1 (steady) or 2 (unsteady)");
    return 1;
  }
  if (mms_t > 0 && !(std::fabs(a) < h0)) {
    suli_set_error("suli_mms_set: the depth h0 must exceed |a| (h = h0 + a*(...) > 0)");
    return 1;
  }
  c.mms.t = mms_t;
  c.mms.a = a;
  c.mms.h0 = h0;
  c.mms.coef = coef;
  c.mms.src_done = false;
  if (mms_t > 0) {
    const int64_t ncell = (int64_t)c.g.nx * c.g.ny * c.g.nz;
    if (!c.mms.erro_p) {
      HIPCHK(hipMalloc((void **)&c.mms.erro_p, ncell * sizeof(double)));
      c.allocs.push_back(c.mms.erro_p);
      const int64_t nblk = (int64_t)((c.g.nx + 1 + 63) / 64) * ((c.g.ny + 1 + 3) / 4) * (c.g.nz + 1);
      HIPCHK(hipMalloc((void **)&c.mms.part, (nblk * 8 + 8) * sizeof(double)));
      c.allocs.push_back(c.mms.part);
    }
    HIPCHK(hipMemsetAsync(c.mms.erro_p, 0, ncell * sizeof(double), c.st));
    HIPCHK(hipStreamSynchronize(c.st));
  }
  return 0;
}

extern "C" int suli_mms_i(void) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  if (!mms_on(c)) {
    suli_set_error("suli_mms_i: manufactured-solution mode is off (suli_mms_set)");
    return 1;
  }
  const Grid &g = c.g;
  const dim3 grid((g.nx + 1 + 63) / 64, (g.ny + 1 + 3) / 4, g.nz + 1);
  DISPATCH(c, hipLaunchKernelGGL(k_mms_init<TT>, grid, dim3(64, 4, 1), 0, c.st, g, par(c), c.t_now, (TT *)c.ls,
                                 (TT *)c.u, (TT *)c.v, (TT *)c.w, c.prd1));
  HIPCHK(hipGetLastError());
  c.convdiff_pending = false;
  guess_reset(c);
  c.mms.src_done = false;
  return 0;
}

extern "C" int suli_mms_bc(void) {
  NEEDCTX();
  return mms_planes_now(*g_ctx);
}

extern "C" int suli_termo_fonte1(void) {
  NEEDCTX();
  if (g_ctx->mms.t != 1) {
    suli_set_error("suli_termo_fonte1: needs manufactured-solution mode 1");
    return 1;
  }
  return mms_sources_now(*g_ctx);
}

extern "C" int suli_termo_fonte2(void) {
  NEEDCTX();
  if (g_ctx->mms.t != 2) {
    suli_set_error("suli_termo_fonte2: needs manufactured-solution mode 2");
    return 1;
  }
  return mms_sources_now(*g_ctx);
}

extern "C" int suli_mms(double *err) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  if (!mms_on(c)) {
    suli_set_error("suli_mms: manufactured-solution mode is off (suli_mms_set)");
    return 1;
  }
  const Grid &g = c.g;
  const dim3 grid((g.nx + 1 + 63) / 64, (g.ny + 1 + 3) / 4, g.nz + 1);
  const int64_t nblk = (int64_t)grid.x * grid.y * grid.z;
  DISPATCH(c, hipLaunchKernelGGL(k_mms_err<TT>, grid, dim3(256, 1, 1), 0, c.st, g, par(c), c.t_now,
                                 (const TT *)c.ls, (const TT *)c.u, (const TT *)c.v, (const TT *)c.w, c.prd1,
                                 c.mms.erro_p, c.mms.part));
  HIPCHK(hipGetLastError());
  double *out = c.mms.part + nblk * 8;
  hipLaunchKernelGGL(k_mms_err_finish, dim3(1), dim3(64), 0, c.st, c.mms.part, nblk, out);
  HIPCHK(hipGetLastError());
  double s[8];
  HIPCHK(hipMemcpyAsync(s, out, sizeof(s), hipMemcpyDeviceToHost, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  for (int q = 0; q < 3; q++) err[q] = std::sqrt(s[2 * q] / s[2 * q + 1]);
  err[3] = c.mms.t == 2 ? std::sqrt(s[6] / s[7]) : NAN;  // the steady case has no exact pressure
  return 0;
}
