// Pressure projection graddin (9_grad.f90:10-278): set-up of the
// variable-density operator in symmetrically scaled form and the reference's
// conjugate-gradient loop, in FP64 with the reference's stopping rule (sum of
// squared scaled residuals <= 1e-4/(nx*ny*nz), at most 10000 iterations).
//
// Layout: the face coefficients a_f = c_f/sqrt(D_P*D_Q) are stored once per
// face in the frame (ax at (i,j,k) is the face between cells i-1 and i), so
// the six arrays matapri*/matapr*/mataprk* of the reference collapse to three
// (aip(i) = ax(i+1) and ain(i) = ax(i) bit for bit: the products D_P*D_Q are
// commutative).  Ghost values of the iteration vectors are never stored: the
// matrix-vector product reads the mirrored neighbour directly, which is the
// value the reference copies into its ghost planes each iteration
// (:234-262).
#pragma once
#include "common.hpp"

// mirrored index of a ghost position along one direction (9_grad.f90:99-116)
__device__ __forceinline__ int mirror_idx(int q, int n, bool periodic) {
  if (q < 1) return periodic ? n : 1;
  if (q > n) return periodic ? 1 : n;
  return q;
}

// c_f at a face (matspri/matsprj/matsprk, :46-76): dt/(h*h)/rho_face with
// rho_face = interp*_cf(rho) (end faces copied, 8_derivadas.f90:304-434).
template <typename T>
__device__ __forceinline__ double face_coef(const T *__restrict__ rho, int64_t cf, int64_t st, int f,
                                            int n, double dth2) {
  const int fc = f < 2 ? 2 : (f > n ? n : f);
  const int64_t cc = cf + (int64_t)(fc - f) * st;
  const double rf = ((double)rho[cc] + (double)rho[cc - st]) * 0.5;
  return dth2 / rf;
}

// D and q = -div u* (:80-87), sqrt(D).
template <typename T>
__global__ __launch_bounds__(256) void k_pD(Grid g, double dt, const T *__restrict__ u,
                                            const T *__restrict__ v, const T *__restrict__ w,
                                            const T *__restrict__ rho, double *__restrict__ D,
                                            double *__restrict__ sD, double *__restrict__ q) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const int64_t c = g.idx(i, j, k);
  const int64_t X = 1, Y = g.sx, Z = g.sz;
  const double mx = dt / (g.dx * g.dx), my = dt / (g.dy * g.dy), mz = dt / (g.dz * g.dz);
  const double d = face_coef<T>(rho, c + X, X, i + 1, g.nx, mx) + face_coef<T>(rho, c, X, i, g.nx, mx) +
                   face_coef<T>(rho, c + Y, Y, j + 1, g.ny, my) + face_coef<T>(rho, c, Y, j, g.ny, my) +
                   face_coef<T>(rho, c + Z, Z, k + 1, g.nz, mz) + face_coef<T>(rho, c, Z, k, g.nz, mz);
  D[c] = d;
  sD[c] = sqrt(d);
  q[c] = ((double)u[c] - (double)u[c + X]) / g.dx + ((double)v[c] - (double)v[c + Y]) / g.dy +
         ((double)w[c] - (double)w[c + Z]) / g.dz;
}

// Ghost cells of a P-extent array, (0:nx+1,0:ny+1,0:nz+1), filled with the
// value at the mirrored interior cell (composition of the x, y, z copies).
static __global__ void k_mirror_ghosts(Grid g, double *__restrict__ a, double *__restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y,
            k = blockIdx.z;
  if (i > g.nx + 1 || j > g.ny + 1 || k > g.nz + 1) return;
  if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz) return;
  const int mi = mirror_idx(i, g.nx, g.ccx0 == 0), mj = mirror_idx(j, g.ny, g.ccy0 == 0),
            mk = mirror_idx(k, g.nz, false);
  const int64_t c = g.idx(i, j, k), s = g.idx(mi, mj, mk);
  a[c] = a[s];
  if (b) b[c] = b[s];
}

// Scaled face coefficients a_f = c_f/sqrt(D_P*D_Q) (:133-140), faces 1..n+1.
template <typename T>
__global__ __launch_bounds__(256) void k_pcoef(Grid g, double dt, const T *__restrict__ rho,
                                               const double *__restrict__ D, double *__restrict__ ax,
                                               double *__restrict__ ay, double *__restrict__ az) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx + 1 || j > ny + 1 || k > nz + 1) return;
  const int64_t c = g.idx(i, j, k);
  const int64_t X = 1, Y = g.sx, Z = g.sz;
  if (j <= ny && k <= nz)
    ax[c] = face_coef<T>(rho, c, X, i, nx, dt / (g.dx * g.dx)) / sqrt(D[c] * D[c - X]);
  if (i <= nx && k <= nz)
    ay[c] = face_coef<T>(rho, c, Y, j, ny, dt / (g.dy * g.dy)) / sqrt(D[c] * D[c - Y]);
  if (i <= nx && j <= ny)
    az[c] = face_coef<T>(rho, c, Z, k, nz, dt / (g.dz * g.dz)) / sqrt(D[c] * D[c - Z]);
}

// e = prd1*sqrt(D) over the whole (0:nx+1,0:ny+1,0:nz+1) extent (:118-124).
static __global__ void k_pe(Grid g, const double *__restrict__ p, const double *__restrict__ sD,
                     double *__restrict__ e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y,
            k = blockIdx.z;
  if (i > g.nx + 1 || j > g.ny + 1 || k > g.nz + 1) return;
  const int64_t c = g.idx(i, j, k);
  e[c] = p[c] * sD[c];
}

// Loop control of the reference CG, kept in device memory so that the host
// can enqueue iterations without waiting for each residual.
struct CGCtl {
  double r0[1];     // initial sum of squared residuals
  double alfa[2];   // alfamupr, alfadipr of the current iteration
  double beta[1];   // betamupr
  double thr;
  int maxit;
  int active;       // 1 while the loop runs
  int next;         // loop condition for the following iteration
  int iters;        // iterations done (cont)
};

// Initial residual r = e - sum a_nb e_nb - q/sqrt(D) (:127-151), stored
// interior only; starts the loop if sum r^2 > thr.
static __global__ __launch_bounds__(256) void k_pr0(Grid g, const double *__restrict__ e,
                                             const double *__restrict__ ax, const double *__restrict__ ay,
                                             const double *__restrict__ az, const double *__restrict__ q,
                                             const double *__restrict__ sD, double *__restrict__ r,
                                             double *__restrict__ pp, CGCtl *ctl, RedSlot rs) {
#pragma clang fp contract(off)
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  double acc[1] = {0.0};
  if (i <= g.nx && j <= g.ny) {
    const int64_t c = g.idx(i, j, k);
    const int64_t X = 1, Y = g.sx, Z = g.sz;
    const double rv = e[c] - ax[c + X] * e[c + X] - ax[c] * e[c - X] - ay[c + Y] * e[c + Y] -
                      ay[c] * e[c - Y] - az[c + Z] * e[c + Z] - az[c] * e[c - Z] - q[c] / sD[c];
    r[c] = rv;
    pp[c] = rv;
    acc[0] = rv * rv;
  }
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1>(acc, lds);
  if (grid_reduce_finish<1>(acc, rs)) {
    if (threadIdx.x == 0 && threadIdx.y == 0) {
      const double a = ctl->r0[0];
      ctl->alfa[0] = a;
      ctl->iters = 0;
      ctl->active = (fabs(a) > ctl->thr && 0 < ctl->maxit) ? 1 : 0;
      ctl->next = ctl->active;
    }
  }
}

// neighbour address with the mirrored-ghost convention
__device__ __forceinline__ int64_t nb_addr(const Grid &g, int64_t c, int q, int n, int64_t st,
                                           int dir, bool periodic) {
  const int t = q + dir;
  const int m = mirror_idx(t, n, periodic);
  return c + (int64_t)(m - q) * st;
}

// mp = A pp; sums r.r and pp.mp (:192-203).
static __global__ __launch_bounds__(256) void k_cg1(Grid g, const double *__restrict__ ax,
                                             const double *__restrict__ ay, const double *__restrict__ az,
                                             const double *__restrict__ r, const double *__restrict__ pp,
                                             double *__restrict__ mp, CGCtl *ctl, RedSlot rs) {
#pragma clang fp contract(off)
  if (!ctl->active) return;
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  double acc[2] = {0.0, 0.0};
  if (i <= g.nx && j <= g.ny) {
    const int64_t c = g.idx(i, j, k);
    const bool px = g.ccx0 == 0, py = g.ccy0 == 0;
    const int64_t ie = nb_addr(g, c, i, g.nx, 1, 1, px), iw = nb_addr(g, c, i, g.nx, 1, -1, px);
    const int64_t jn = nb_addr(g, c, j, g.ny, g.sx, 1, py), js = nb_addr(g, c, j, g.ny, g.sx, -1, py);
    const int64_t kt = nb_addr(g, c, k, g.nz, g.sz, 1, false), kb = nb_addr(g, c, k, g.nz, g.sz, -1, false);
    const double m = pp[c] - pp[ie] * ax[c + 1] - pp[iw] * ax[c] - pp[jn] * ay[c + g.sx] - pp[js] * ay[c] -
                     pp[kt] * az[c + g.sz] - pp[kb] * az[c];
    mp[c] = m;
    acc[0] = r[c] * r[c];
    acc[1] = pp[c] * m;
  }
  __shared__ double lds[2 * SULI_LDS_RED];
  block_reduce<2>(acc, lds);
  if (grid_reduce_finish<2>(acc, rs)) {
    if (threadIdx.x == 0 && threadIdx.y == 0) {
      ctl->iters += 1;
      ctl->next = (fabs(ctl->alfa[0]) > ctl->thr && ctl->iters < ctl->maxit) ? 1 : 0;
    }
  }
}

// e -= alpha pp; r -= alpha mp; sum r.r (:205-219).
static __global__ __launch_bounds__(256) void k_cg2(Grid g, double *__restrict__ e, double *__restrict__ r,
                                             const double *__restrict__ pp, const double *__restrict__ mp,
                                             CGCtl *ctl, RedSlot rs) {
#pragma clang fp contract(off)
  if (!ctl->active) return;
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  const double alfapr = ctl->alfa[0] / ctl->alfa[1];
  double acc[1] = {0.0};
  if (i <= g.nx && j <= g.ny) {
    const int64_t c = g.idx(i, j, k);
    e[c] = e[c] - alfapr * pp[c];
    const double rv = r[c] - alfapr * mp[c];
    r[c] = rv;
    acc[0] = rv * rv;
  }
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1>(acc, lds);
  grid_reduce_finish<1>(acc, rs);
}

// pp = r + beta pp (:221-230); hands the loop condition to the next iteration.
static __global__ __launch_bounds__(256) void k_cg3(Grid g, const double *__restrict__ r,
                                             double *__restrict__ pp, CGCtl *ctl,
                                             unsigned int *__restrict__ done_counter) {
#pragma clang fp contract(off)
  if (!ctl->active) return;
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  const double betapr = ctl->beta[0] / ctl->alfa[0];
  if (i <= g.nx && j <= g.ny) {
    const int64_t c = g.idx(i, j, k);
    pp[c] = r[c] + betapr * pp[c];
  }
  // the last block to finish switches the loop flag for the next iteration
  __shared__ bool last;
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    __threadfence();
    const unsigned int nb = gridDim.x * gridDim.y * gridDim.z;
    last = atomicAdd(done_counter, 1u) == nb - 1;
    if (last) {
      *done_counter = 0u;
      ctl->active = ctl->next;
    }
  }
}

// prd1 = e/sqrt(D) over the whole extent (:268-274); ghosts take the mirrored
// interior value when at least one iteration mirrored e.
static __global__ void k_pfinal(Grid g, const double *__restrict__ e, const double *__restrict__ sD,
                         double *__restrict__ p, const CGCtl *ctl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y,
            k = blockIdx.z;
  if (i > g.nx + 1 || j > g.ny + 1 || k > g.nz + 1) return;
  const int64_t c = g.idx(i, j, k);
  int64_t s = c;
  if (ctl->iters > 0)
    s = g.idx(mirror_idx(i, g.nx, g.ccx0 == 0), mirror_idx(j, g.ny, g.ccy0 == 0), mirror_idx(k, g.nz, false));
  p[c] = e[s] / sD[s];
}
