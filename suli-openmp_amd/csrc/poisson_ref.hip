// Reference CG of the projection (9_grad.f90:10-278) in FP64: the loop
// kernels and their host driver.  See poisson_ref.hpp for the layout.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>

#include "ctx.hpp"

// e = prd1*sqrt(D) over the whole (0:nx+1,0:ny+1,0:nz+1) extent (:118-124).
__global__ void k_pe(Grid g, const double *__restrict__ p, const double *__restrict__ sD,
                     double *__restrict__ e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y,
            k = blockIdx.z;
  if (i > g.nx + 1 || j > g.ny + 1 || k > g.nz + 1) return;
  const int64_t c = g.idx(i, j, k);
  e[c] = p[c] * sD[c];
}



// Initial residual r = e - sum a_nb e_nb - q/sqrt(D) (:127-151), stored
// interior only; starts the loop if sum r^2 > thr.
__global__ __launch_bounds__(256) void k_pr0(Grid g, const double *__restrict__ e,
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
  block_partial<1>(acc, rs.partials);
}

// loop-control epilogues of the reduction finishes (run by one thread)
struct EpiCgR0 {  // after the initial residual: start the loop if sum r^2 > thr
  CGCtl *ctl;
  __device__ bool enabled() const { return true; }
  __device__ void operator()(const double *res) const {
    ctl->alfa[0] = res[0];
    ctl->iters = 0;
    ctl->active = (fabs(res[0]) > ctl->thr && 0 < ctl->maxit) ? 1 : 0;
    ctl->next = ctl->active;
  }
};
struct EpiCgAlfa {  // after r.r and pp.mp: count the iteration, decide the next one (:175)
  CGCtl *ctl;
  __device__ bool enabled() const { return ctl->active != 0; }
  __device__ void operator()(const double *) const {
    ctl->iters += 1;
    ctl->next = (fabs(ctl->alfa[0]) > ctl->thr && ctl->iters < ctl->maxit) ? 1 : 0;
  }
};
struct EpiCgBeta {  // after the new r.r: the direction update of a final iteration is not needed
  CGCtl *ctl;
  __device__ bool enabled() const { return ctl->active != 0; }
  __device__ void operator()(const double *) const { ctl->active = ctl->next; }
};

// neighbour address with the mirrored-ghost convention
__device__ __forceinline__ int64_t nb_addr(const Grid &g, int64_t c, int q, int n, int64_t st,
                                           int dir, bool periodic) {
  const int t = q + dir;
  const int m = mirror_idx(t, n, periodic);
  return c + (int64_t)(m - q) * st;
}

// mp = A pp; sums r.r and pp.mp (:192-203).
__global__ __launch_bounds__(256) void k_cg1(Grid g, const double *__restrict__ ax,
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
  block_partial<2>(acc, rs.partials);
}

// e -= alpha pp; r -= alpha mp; sum r.r (:205-219).
__global__ __launch_bounds__(256) void k_cg2(Grid g, double *__restrict__ e, double *__restrict__ r,
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
  block_partial<1>(acc, rs.partials);
}

// pp = r + beta pp (:221-230).
__global__ __launch_bounds__(256) void k_cg3(Grid g, const double *__restrict__ r, double *__restrict__ pp,
                                             const CGCtl *ctl) {
#pragma clang fp contract(off)
  if (!ctl->active) return;
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  const double betapr = ctl->beta[0] / ctl->alfa[0];
  if (i <= g.nx && j <= g.ny) {
    const int64_t c = g.idx(i, j, k);
    pp[c] = r[c] + betapr * pp[c];
  }
}

// prd1 = e/sqrt(D) over the whole extent (:268-274); ghosts take the mirrored
// interior value when at least one iteration mirrored e.
__global__ void k_pfinal(Grid g, const double *__restrict__ e, const double *__restrict__ sD,
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

int poisson_ref_solve(Ctx &c, double dt, suli_poisson_stats *stt) {
  const Grid &g = c.g;
  HIPCHK(hipEventRecord(c.solve_ev[0], c.st));
  { ProfScope ps_(c, "poisson_setup"); DISPATCH(c, Launch<TT>::pD(g, c.st, dt, c.u, c.v, c.w, c.rho, c.D, c.sD, c.q)); }
  { ProfScope ps_(c, "poisson_setup"); DISPATCH(c, Launch<TT>::pcoef(g, c.st, dt, c.rho, c.D, c.ax, c.ay, c.az)); }
  const dim3 b(64, 4, 1);
  const dim3 pg((g.nx + 2 + 63) / 64, (g.ny + 2 + 3) / 4, g.nz + 2);
  { ProfScope ps_(c, "poisson_setup"); hipLaunchKernelGGL(k_pe, pg, b, 0, c.st, g, c.prd1, c.sD, c.e); }
  std::memset(c.ctl_h, 0, sizeof(CGCtl));
  c.ctl_h->thr = c.P.tol_abs;
  c.ctl_h->maxit = c.P.max_it;
  HIPCHK(hipMemcpyAsync(c.ctl, c.ctl_h, sizeof(CGCtl), hipMemcpyHostToDevice, c.st));
  const dim3 cg = cell_grid(g);
  RedSlot r0{c.partials, (double *)((char *)c.ctl + offsetof(CGCtl, r0))};
  { ProfScope ps_(c, "cg_residual0"); hipLaunchKernelGGL(k_pr0, cg, b, 0, c.st, g, c.e, c.ax, c.ay, c.az, c.q, c.sD, c.r, c.pp, c.ctl, r0);
    launch_finish<1>(c.st, cg, r0, EpiCgR0{c.ctl}); }
  RedSlot r1{c.partials, (double *)((char *)c.ctl + offsetof(CGCtl, alfa))};
  RedSlot r2{c.partials, (double *)((char *)c.ctl + offsetof(CGCtl, beta))};
  int chunk = 8;
  for (;;) {
    for (int n = 0; n < chunk; n++) {
      { ProfScope ps_(c, "cg_spmv_dot"); hipLaunchKernelGGL(k_cg1, cg, b, 0, c.st, g, c.ax, c.ay, c.az, c.r, c.pp, c.mp, c.ctl, r1);
        launch_finish<2>(c.st, cg, r1, EpiCgAlfa{c.ctl}); }
      { ProfScope ps_(c, "cg_update"); hipLaunchKernelGGL(k_cg2, cg, b, 0, c.st, g, c.e, c.r, c.pp, c.mp, c.ctl, r2);
        launch_finish<1>(c.st, cg, r2, EpiCgBeta{c.ctl}); }
      { ProfScope ps_(c, "cg_direction"); hipLaunchKernelGGL(k_cg3, cg, b, 0, c.st, g, c.r, c.pp, c.ctl); }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c.ctl_h, c.ctl, sizeof(CGCtl), hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    if (!c.ctl_h->active) break;
    if (chunk < 64) chunk *= 2;
  }
  { ProfScope ps_(c, "poisson_final"); hipLaunchKernelGGL(k_pfinal, pg, b, 0, c.st, g, c.e, c.sD, c.prd1, c.ctl); }
  HIPCHK(hipGetLastError());
  double gpu_ms = 0.0;
  if (solve_elapsed_ms(c, &gpu_ms)) return 1;
  if (stt) {
    std::memset(stt, 0, sizeof(*stt));
    stt->iters = c.ctl_h->iters;
    stt->outer = 0;
    stt->converged = !(std::fabs(c.ctl_h->alfa[0]) > c.ctl_h->thr) ? 1 : 0;
    stt->res0 = c.ctl_h->r0[0];
    stt->res = c.ctl_h->alfa[0];
    stt->ms = gpu_ms;
  }
  return 0;
}

