// Mixed-precision pressure projection: flexible conjugate gradients (FCG(1),
// Notay 2000) in FP32 inside an FP64 iterative-refinement loop.
//
// The system is the reference's symmetrically scaled one (9_grad.f90:118-151):
// A_s e = b_s with A_s = D^-1/2 A D^-1/2, b_s = q/sqrt(D), p = e/sqrt(D),
// set up in FP64 by the same kernels as the reference CG.  Each refinement
// pass computes the true residual r = b_s - A_s x in FP64 (FP64 dot
// products), removes its component along the null space sqrt(D) of the
// singular Neumann operator, and solves A_s d = r approximately in FP32 with
// FCG; x += d in FP64.  The solve stops on the FP64 relative residual of the
// solvable part, ||P(b_s - A_s x)|| / ||P b_s|| <= tol, P the projection that
// removes the component along sqrt(D).  In a closed tank b_s has no such
// component (the wall fluxes are zero) and this is the plain relative
// residual; with a net flux through the walls (a prescribed inflow) the
// system has no solution and the solve converges to the least-squares one.
// The component r.n/|n| is the same for every iterate (A_s n = 0), so the
// projected norms come from the r.r, r.n, n.n of the residual kernel.
//
// Preconditioners: Jacobi (identity in the scaled system: the reference's own
// symmetric diagonal scaling), a geometric multigrid V-cycle on the unscaled
// operator (z = D^1/2 V(D^1/2 r)), or the learned U-Net (unet.hip).  Loop
// control lives in device memory: every kernel returns at once when the
// solve has converged, so the host enqueues iterations in chunks and reads
// the state back once per chunk.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "fcg.hpp"

// truncation of the FCG recurrence used with the learned preconditioner: the
// new direction is A-orthogonalised against the last SULI_FCG_M directions,
// which keeps the method convergent for a nonlinear preconditioner
#define SULI_FCG_M 4

// ---------------------------------------------------------------------------
// device control block
struct FcgCtl {
  double bb[1];      // ||b_s||^2
  double rout[3];    // outer: r.r, r.n, n.n (n = sqrt(D))
  double rin0[1];    // inner: initial r.r
  double k1[2];      // p.q, p.r
  double k2[2];      // r.r, r.q (next beta numerator for Jacobi)
  double zq[1];      // z.q_prev (preconditioned beta numerator)
  double thr_in;     // inner stop: r.r <= thr_in
  int maxit_in;
  int active;        // inner loop running
  int it;            // inner iterations of this pass
  int first;         // next K1 is the first of the pass (beta = 0)
  int done;          // the last outer residual met the tolerance (set with the pass's arming)
  int out;           // the last outer residual ends the solve: the output kernel writes
  // FCG(m) with the learned preconditioner: the last SULI_FCG_M directions
  float sig;                 // input scale of the finest level's staged V-cycle input (FP16 form)
  unsigned seq;              // (host copy only) the number of the publish that wrote it
  int nk;                    // directions held (0..SULI_FCG_M)
  double zqm[SULI_FCG_M];    // z.q_j of the current z
  double pqm[SULI_FCG_M];    // p_j.q_j of the held directions
};

#define CTL_PTR(c, member) ((double *)((char *)(c) + offsetof(FcgCtl, member)))

__device__ __forceinline__ int64_t nbx(const Grid &g, int64_t c, int q, int n, int64_t st, int dir) {
  const int t = q + dir;
  const int m = t < 1 ? 1 : (t > n ? n : t);
  return c + (int64_t)(m - q) * st;
}

// A_s v at cell (i,j,k) with the reference's mirrored Neumann ghosts
// (periodic walls wrap), coefficients of type C, vector of type V.
template <typename C, typename V, typename A>
__device__ __forceinline__ A apply_As(const Grid &g, int64_t c, int i, int j, int k, const C *__restrict__ ax,
                                      const C *__restrict__ ay, const C *__restrict__ az, const V *__restrict__ v) {
  const bool px = g.ccx0 == 0, py = g.ccy0 == 0;
  auto nb = [&](int q, int n, int64_t st, int dir, bool per) -> int64_t {
    int t = q + dir;
    int m = t < 1 ? (per ? n : 1) : (t > n ? (per ? 1 : n) : t);
    return c + (int64_t)(m - q) * st;
  };
  const int64_t ie = nb(i, g.nx, 1, 1, px), iw = nb(i, g.nx, 1, -1, px);
  const int64_t jn = nb(j, g.ny, g.sx, 1, py), js = nb(j, g.ny, g.sx, -1, py);
  const int64_t kt = nb(k, g.nz, g.sz, 1, false), kb = nb(k, g.nz, g.sz, -1, false);
  return (A)v[c] - (A)ax[c + 1] * (A)v[ie] - (A)ax[c] * (A)v[iw] - (A)ay[c + g.sx] * (A)v[jn] -
         (A)ay[c] * (A)v[js] - (A)az[c + g.sz] * (A)v[kt] - (A)az[c] * (A)v[kb];
}

#define CELL_IJK                                                                        \
  int bx_, by_, bz_;                                                                    \
  xcd_band_block(bx_, by_, bz_);                                                         \
  const int i = 1 + bx_ * SULI_BX + threadIdx.x, j = 1 + by_ * SULI_BY + threadIdx.y; \
  const int kz0 = 1 + bz_;

// ---- outer (FP64) ---------------------------------------------------------
// r = b - A_s x in FP64; sums r.r, r.n, n.n
__global__ __launch_bounds__(256) void k_res64(Grid g, const double *__restrict__ ax, const double *__restrict__ ay,
                                               const double *__restrict__ az, const double *__restrict__ x,
                                               const double *__restrict__ b, const double *__restrict__ sD,
                                               double *__restrict__ r, RedSlot rs) {
  CELL_IJK;
  double acc[3] = {0.0, 0.0, 0.0};
  if (i <= g.nx && j <= g.ny)
    for (int k = kz0; k <= g.nz; k += gridDim.z) {
      const int64_t c = g.idx(i, j, k);
      const double rv = b[c] - apply_As<double, double, double>(g, c, i, j, k, ax, ay, az, x);
      r[c] = rv;
      acc[0] += rv * rv;
      acc[1] += rv * sD[c];
      acc[2] += sD[c] * sD[c];
    }
  __shared__ double lds[3 * SULI_LDS_RED];
  block_reduce<3>(acc, lds);
  block_partial<3>(acc, rs.partials);
}

// the finest multigrid level's input staged from r (k_mg0_in's values, cell
// by cell): b = sqrt(D) r split by colour, red x = b/diag
// Storage of the finest multigrid level's colour-split iterate and
// right-hand side: FP32, or FP16 with the V-cycle's input scaled to a unit
// RMS residual (the scale from the last residual norm the device holds; the
// output kernel divides it out).  FP16 rounding (2^-11 relative) perturbs
// the preconditioner far below its own error; it cuts the bytes of the
// finest-level half sweeps from 12 to 9 per cell.  Values beyond the FP16
// range saturate (the V-cycle stays finite; FCG corrects a poor z).
__device__ __forceinline__ float mgc(_Float16 v) { return (float)v; }
template <typename XT> __device__ __forceinline__ XT to_x(float v);
template <> __device__ __forceinline__ float to_x<float>(float v) { return v; }
template <> __device__ __forceinline__ _Float16 to_x<_Float16>(float v) {
  return (_Float16)fminf(fmaxf(v, -65504.f), 65504.f);
}
// 1/RMS of a residual with squared norm rr over n cells (1 when rr is not
// a positive finite number)
__device__ __forceinline__ float unit_rms_scale(double rr, int64_t n) {
  return (rr > 0.0 && rr < 1e300) ? (float)(1.0 / sqrt(rr / (double)n)) : 1.f;
}

struct StageIn {
  Grid h;  // pair frame of the colour-split arrays
  const float *s, *dred;
  void *bred, *bblack, *xred;  // bred null: no staging
  int f16;                     // the three arrays hold FP16 of the scaled input (else FP32)
};

template <typename XT>
__device__ __forceinline__ void stage_cell(XT *bred, XT *bblack, XT *xred, const float *dred, int64_t qh, bool red,
                                           float bv) {
  if (red) {
    bred[qh] = to_x<XT>(bv);
    xred[qh] = to_x<XT>((bv + 0.f) / dred[qh]);
  } else {
    bblack[qh] = to_x<XT>(bv);
  }
}

// (sg: the input scale, 1 for FP32 storage)
__device__ __forceinline__ void stage_in(const StageIn &st, const Grid &g, int i, int j, int k, int64_t c,
                                         float rv, float sg) {
  const float bv = st.s[c] * rv * sg;
  const int64_t qh = st.h.idx((i - 1) >> 1, j, k);
  const bool red = ((i + j + k) & 1) == 0;
  if (st.f16)
    stage_cell((_Float16 *)st.bred, (_Float16 *)st.bblack, (_Float16 *)st.xred, st.dred, qh, red, bv);
  else
    stage_cell((float *)st.bred, (float *)st.bblack, (float *)st.xred, st.dred, qh, red, bv);
}
// the input scale of the V-cycle that the staging kernel feeds, kept for its
// output kernel (one thread writes it)
__device__ __forceinline__ float stage_scale(const StageIn &st, FcgCtl *ctl, const Grid &g, double rr) {
  const float sg = st.f16 ? unit_rms_scale(rr, (int64_t)g.nx * g.ny * g.nz) : 1.f;
  if (st.bred && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 && threadIdx.y == 0)
    ctl->sig = sg;
  return sg;
}

// r32 = r - (r.n/n.n) n; d = 0; sums r32.r32 and arms the inner loop; with
// st.bred, the first V-cycle's input from r32
__global__ __launch_bounds__(256) void k_r32(Grid g, const double *__restrict__ r, const double *__restrict__ sD,
                                             float *__restrict__ r32, float *__restrict__ d32, FcgCtl *ctl,
                                             double rel_in, int maxit, RedSlot rs, StageIn st) {
  CELL_IJK;
  const double mu = ctl->rout[1] / ctl->rout[2];
  const float sg = stage_scale(st, ctl, g, ctl->rout[0]);
  double acc[1] = {0.0};
  if (i <= g.nx && j <= g.ny)
    for (int k = kz0; k <= g.nz; k += gridDim.z) {
      const int64_t c = g.idx(i, j, k);
      const float rv = (float)(r[c] - mu * sD[c]);
      r32[c] = rv;
      d32[c] = 0.f;
      acc[0] += (double)rv * (double)rv;
      if (st.bred) stage_in(st, g, i, j, k, c, rv, sg);
    }
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1>(acc, lds);
  block_partial<1>(acc, rs.partials);
}

// loop-control epilogues of the reduction finishes (run by one thread)
// Arm the inner loop of a refinement pass: stop when r.r <= rin^2 r0.r0 with
// rin = max(rel_in, min(0.3 tol ||b|| / ||r_outer||, 1)) (no further than the
// overall tolerance needs), from the outer residual of this pass (rout) and
// ||b|| (bb) already on the device; when that residual meets the tolerance
// the pass does not run (done).
// One FP32 pass reaches a reduction of a few 1e-7 before its single
// precision limits it: a warm start needing at least 1e-4 times the usual
// reduction (an initial residual below ~0.3 of the tolerance's reciprocal
// scale) takes one pass aimed at the tolerance, with 3e-3 of the usual
// reduction as the floor.  On the dam break that leaves one pass per
// projection through step 700 (1.85 passes per solve at steps 506-525 with
// the thresholds 0.03 and 0.1, 1.0 with these; 769 -> 801 Mcell-steps/s).
#ifndef SULI_FCG_ONEPASS_WANT
#define SULI_FCG_ONEPASS_WANT 0.0001  // one pass when it needs at least this fraction of the usual reduction
#endif
#ifndef SULI_FCG_ONEPASS_FLOOR
#define SULI_FCG_ONEPASS_FLOOR 0.003  // and then aims for at most this fraction of it
#endif
// The learned preconditioner's passes are capped (kLearnedPassCap): a deep
// single pass would stall against the cap and hand the solve to multigrid,
// so its passes keep to at most 10x the usual reduction, and only when that
// ends the solve (a warm start within 3e-2 of the usual reduction).
#define SULI_FCG_ONEPASS_WANT_CAPPED 0.03
#define SULI_FCG_ONEPASS_FLOOR_CAPPED 0.1
// squared norms of the solvable parts of the outer residual and of b_s
__host__ __device__ inline void proj_norms(const double *rout, double bb_full, double &rr, double &bb) {
  const double nc = rout[2] > 0 ? rout[1] * rout[1] / rout[2] : 0.0;
  rr = fmax(rout[0] - nc, 0.0);
  bb = fmax(bb_full - nc, 0.0);
}
struct EpiR32 {
  FcgCtl *ctl;
  double rel_in, tol;
  int maxit;
  double want_frac, floor_frac;  // one-pass thresholds (SULI_FCG_ONEPASS_*)
  __device__ bool enabled() const { return true; }
  __device__ void operator()(const double *res) const {
    double rr, bb;
    proj_norms(ctl->rout, ctl->bb[0], rr, bb);
    const bool conv = !(rr > tol * tol * bb);
    const double want = rr > 0 ? 0.3 * tol * sqrt(bb) / sqrt(rr) : 1.0;
    // one pass of up to 1/floor_frac times the usual reduction when that
    // ends the solve (a warm start); cold starts take passes of rel_in,
    // which cost fewer iterations in FP32
    const double floor = want >= want_frac * rel_in ? floor_frac * rel_in : rel_in;
    const double rin = fmax(floor, fmin(want, 1.0));
    ctl->thr_in = res[0] * rin * rin;
    ctl->maxit_in = maxit;
    ctl->it = 0;
    ctl->first = 1;
    ctl->nk = 0;
    ctl->done = conv ? 1 : 0;
    ctl->active = (res[0] > 0 && !conv) ? 1 : 0;
  }
};
// the outer residual's verdict for the output kernel: the solve ends with it
// when it meets the tolerance (the host's test, proj_res) or when force is set
struct EpiOut {
  FcgCtl *ctl;
  double tol;
  int force;
  FcgCtl *pub;   // set: publish the control block to this host copy (epi_publish)
  unsigned seq;
  __device__ bool enabled() const { return true; }
  __device__ void operator()(const double *res) const {
    double rr, bb;
    proj_norms(res, ctl->bb[0], rr, bb);
    const double bn = sqrt(bb);
    const double rel = bn > 0 ? sqrt(rr) / bn : 0.0;
    ctl->out = (force || !(rel > tol)) ? 1 : 0;
  }
};
struct EpiActive {  // plain result, only while the inner loop runs
  FcgCtl *ctl;
  __device__ bool enabled() const { return ctl->active != 0; }
  __device__ void operator()(const double *) const {}
};
struct EpiK1 {
  FcgCtl *ctl;
  __device__ bool enabled() const { return ctl->active != 0; }
  __device__ void operator()(const double *) const { ctl->first = 0; }
};
struct EpiK2 {  // count the iteration and decide whether the next one runs
  FcgCtl *ctl;
  FcgCtl *pub;   // set on the last iteration of an enqueued chunk (epi_publish)
  unsigned seq;
  __device__ bool enabled() const { return ctl->active != 0; }
  __device__ void operator()(const double *res) const {
    ctl->it += 1;
    const double pq = ctl->k1[0];
    const bool stop = !(res[0] > ctl->thr_in) || ctl->it >= ctl->maxit_in || !(pq > 0.0) || !(res[0] == res[0]);
    ctl->active = stop ? 0 : 1;
  }
};

// The control block written straight into its pinned, coherent host copy
// by the finish kernel that ends an enqueued chunk (in place of a device to
// host copy, which the runtime runs as a kernel of its own): the host reads
// it after the stream synchronisation and checks seq.
__device__ __forceinline__ void publish_ctl(const FcgCtl *ctl, FcgCtl *pub, unsigned seq) {
  static_assert(sizeof(FcgCtl) % 8 == 0, "the control block is copied in 8-byte words");
  const uint64_t *s = (const uint64_t *)ctl;
  uint64_t *d = (uint64_t *)pub;
  for (int w = 0; w < (int)(sizeof(FcgCtl) / 8); w++) d[w] = s[w];
  pub->seq = seq;
  __threadfence_system();
}
__device__ __forceinline__ void epi_publish(const EpiK2 &e) {
  if (e.pub) publish_ctl(e.ctl, e.pub, e.seq);
}
__device__ __forceinline__ void epi_publish(const EpiOut &e) {
  if (e.pub) publish_ctl(e.ctl, e.pub, e.seq);
}

// x += d (FP64 += FP32)
__global__ __launch_bounds__(256) void k_xupd(Grid g, double *__restrict__ x, const float *__restrict__ d) {
  CELL_IJK;
  if (i > g.nx || j > g.ny) return;
  for (int k = kz0; k <= g.nz; k += gridDim.z) {
    const int64_t c = g.idx(i, j, k);
    x[c] += (double)d[c];
  }
}

// ---- inner FCG (FP32 vectors) --------------------------------------------------
// K1: beta from the previous reductions, p = z + beta p, q = A_s p; sums p.q, p.r
// z is r for Jacobi.  p at the neighbours is rebuilt on the fly from z and p_old.
__global__ __launch_bounds__(256) void k_fcg1(Grid g, const float *__restrict__ ax, const float *__restrict__ ay,
                                              const float *__restrict__ az, const float *__restrict__ z,
                                              const float *__restrict__ r, const float *__restrict__ pold,
                                              float *__restrict__ pnew, float *__restrict__ q, FcgCtl *ctl,
                                              int precond, RedSlot rs) {
  if (!ctl->active) return;
  CELL_IJK;
  // FCG(1): beta = -(z.q_prev)/(p_prev.q_prev); with z = r that is -(r.q_prev)/(p.q)
  float beta = 0.f;
  if (!ctl->first) beta = (float)(-(precond ? ctl->zq[0] : ctl->k2[1]) / ctl->k1[0]);
  double acc[2] = {0.0, 0.0};
  if (i <= g.nx && j <= g.ny)
    for (int k = kz0; k <= g.nz; k += gridDim.z) {
      const int64_t c = g.idx(i, j, k);
      const bool px = g.ccx0 == 0, py = g.ccy0 == 0;
      auto nb = [&](int qq, int n, int64_t st, int dir, bool per) -> int64_t {
        int t = qq + dir;
        int m = t < 1 ? (per ? n : 1) : (t > n ? (per ? 1 : n) : t);
        return c + (int64_t)(m - qq) * st;
      };
      const int64_t ie = nb(i, g.nx, 1, 1, px), iw = nb(i, g.nx, 1, -1, px);
      const int64_t jn = nb(j, g.ny, g.sx, 1, py), js = nb(j, g.ny, g.sx, -1, py);
      const int64_t kt = nb(k, g.nz, g.sz, 1, false), kb = nb(k, g.nz, g.sz, -1, false);
      auto P = [&](int64_t a) { return z[a] + beta * pold[a]; };
      const float pc = P(c);
      const float qv = pc - ax[c + 1] * P(ie) - ax[c] * P(iw) - ay[c + g.sx] * P(jn) - ay[c] * P(js) -
                       az[c + g.sz] * P(kt) - az[c] * P(kb);
      pnew[c] = pc;
      q[c] = qv;
      acc[0] += (double)pc * (double)qv;
      acc[1] += (double)pc * (double)r[c];
    }
  __shared__ double lds[2 * SULI_LDS_RED];
  block_reduce<2>(acc, lds);
  block_partial<2>(acc, rs.partials);
}

// K2: alpha = p.r/p.q; d += alpha p; r -= alpha q; sums r.r, r.q; loop control;
// with st.bred, the next V-cycle's input from the new r
__global__ __launch_bounds__(256) void k_fcg2(Grid g, const float *__restrict__ p, const float *__restrict__ q,
                                              float *__restrict__ r, float *__restrict__ d, FcgCtl *ctl,
                                              RedSlot rs, StageIn st) {
  if (!ctl->active) return;
  CELL_IJK;
  const double pq = ctl->k1[0];
  const float alpha = pq != 0.0 ? (float)(ctl->k1[1] / pq) : 0.f;
  // (the last residual norm on the device: the pass's initial one before its first update)
  const float sg = stage_scale(st, ctl, g, ctl->it == 0 ? ctl->rin0[0] : ctl->k2[0]);
  double acc[2] = {0.0, 0.0};
  if (i <= g.nx && j <= g.ny)
    for (int k = kz0; k <= g.nz; k += gridDim.z) {
      const int64_t c = g.idx(i, j, k);
      const float qv = q[c];
      d[c] += alpha * p[c];
      const float rv = r[c] - alpha * qv;
      r[c] = rv;
      acc[0] += (double)rv * (double)rv;
      acc[1] += (double)rv * (double)qv;
      if (st.bred) stage_in(st, g, i, j, k, c, rv, sg);
    }
  __shared__ double lds[2 * SULI_LDS_RED];
  block_reduce<2>(acc, lds);
  block_partial<2>(acc, rs.partials);
}

// z.q_prev for the preconditioned beta
__global__ __launch_bounds__(256) void k_zq(Grid g, const float *__restrict__ z, const float *__restrict__ q,
                                            FcgCtl *ctl, RedSlot rs) {
  if (!ctl->active) return;
  CELL_IJK;
  double acc[1] = {0.0};
  if (i <= g.nx && j <= g.ny && !ctl->first)
    for (int k = kz0; k <= g.nz; k += gridDim.z) {
      const int64_t c = g.idx(i, j, k);
      acc[0] += (double)z[c] * (double)q[c];
    }
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1>(acc, lds);
  block_partial<1>(acc, rs.partials);
}

// ---- FCG(m) steps of the learned preconditioner ---------------------------------
// LZ: s = A_s z; sums z.q_j over the held directions (slots j < nk)
__global__ __launch_bounds__(256) void k_fcgm_lz(Grid g, const float *__restrict__ ax, const float *__restrict__ ay,
                                                 const float *__restrict__ az, const float *__restrict__ z,
                                                 float *__restrict__ sv, const float *__restrict__ q0,
                                                 const float *__restrict__ q1, const float *__restrict__ q2,
                                                 const float *__restrict__ q3, FcgCtl *ctl, RedSlot rs) {
  if (!ctl->active) return;
  CELL_IJK;
  const int nk = ctl->nk;
  const float *qs[SULI_FCG_M] = {q0, q1, q2, q3};
  double acc[SULI_FCG_M] = {0.0, 0.0, 0.0, 0.0};
  if (i <= g.nx && j <= g.ny)
    for (int k = kz0; k <= g.nz; k += gridDim.z) {
      const int64_t c = g.idx(i, j, k);
      sv[c] = apply_As<float, float, float>(g, c, i, j, k, ax, ay, az, z);
      const double zc = z[c];
#pragma unroll
      for (int m = 0; m < SULI_FCG_M; m++)
        if (m < nk) acc[m] += zc * (double)qs[m][c];
    }
  __shared__ double lds[SULI_FCG_M * SULI_LDS_RED];
  block_reduce<SULI_FCG_M>(acc, lds);
  block_partial<SULI_FCG_M>(acc, rs.partials);
}

// LP: p = z - sum_j (z.q_j/p_j.q_j) p_j, q = s - sum_j (...) q_j, written
// into slot (pw, qw), which is the oldest held direction once all slots are
// in use (each element is read before it is overwritten); sums p.q, p.r
__global__ __launch_bounds__(256) void k_fcgm_lp(Grid g, const float *__restrict__ z, const float *__restrict__ sv,
                                                 const float *__restrict__ r, const float *p0, const float *p1,
                                                 const float *p2, const float *p3, const float *q0, const float *q1,
                                                 const float *q2, const float *q3, float *pw, float *qw, FcgCtl *ctl,
                                                 RedSlot rs) {
  if (!ctl->active) return;
  CELL_IJK;
  const int nk = ctl->nk;
  const float *ps[SULI_FCG_M] = {p0, p1, p2, p3};
  const float *qs[SULI_FCG_M] = {q0, q1, q2, q3};
  float beta[SULI_FCG_M];
#pragma unroll
  for (int m = 0; m < SULI_FCG_M; m++) beta[m] = m < nk ? (float)(ctl->zqm[m] / ctl->pqm[m]) : 0.f;
  double acc[2] = {0.0, 0.0};
  if (i <= g.nx && j <= g.ny)
    for (int k = kz0; k <= g.nz; k += gridDim.z) {
      const int64_t c = g.idx(i, j, k);
      float pv = z[c], qv = sv[c];
#pragma unroll
      for (int m = 0; m < SULI_FCG_M; m++)
        if (m < nk) {
          pv -= beta[m] * ps[m][c];
          qv -= beta[m] * qs[m][c];
        }
      pw[c] = pv;
      qw[c] = qv;
      acc[0] += (double)pv * (double)qv;
      acc[1] += (double)pv * (double)r[c];
    }
  __shared__ double lds[2 * SULI_LDS_RED];
  block_reduce<2>(acc, lds);
  block_partial<2>(acc, rs.partials);
}
struct EpiLP {
  FcgCtl *ctl;
  int slot;
  __device__ bool enabled() const { return ctl->active != 0; }
  __device__ void operator()(const double *res) const {
    ctl->k1[0] = res[0];
    ctl->k1[1] = res[1];
    ctl->pqm[slot] = res[0];
    if (ctl->nk < SULI_FCG_M) ctl->nk += 1;
  }
};

// ---- geometric multigrid on the unscaled operator ------------------------------
// Level l: cells 1..n, faces 1..n+1 in a frame of its own; face conductances
// cx (face i between cells i-1 and i), zero on the walls (Neumann); diagonal
// dg = sum of the six face conductances.

// level-0 conductances c_f = dt/(h^2 rho_f) from the FP64 set-up (a_f sqrt(D_P D_Q))
__global__ void k_mg_c0(Grid g, const double *__restrict__ ax, const double *__restrict__ ay,
                        const double *__restrict__ az, const double *__restrict__ D, float *__restrict__ cx,
                        float *__restrict__ cy, float *__restrict__ cz, float *__restrict__ dg,
                        float *__restrict__ ax32, float *__restrict__ ay32, float *__restrict__ az32,
                        const double *__restrict__ sD, float *__restrict__ sD32) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx + 1 || j > ny + 1 || k > nz + 1) return;
  const int64_t c = g.idx(i, j, k);
  if (j <= ny && k <= nz) {
    ax32[c] = (float)ax[c];
    cx[c] = (i == 1 || i == nx + 1) ? 0.f : (float)(ax[c] * sqrt(D[c] * D[c - 1]));
  }
  if (i <= nx && k <= nz) {
    ay32[c] = (float)ay[c];
    cy[c] = (j == 1 || j == ny + 1) ? 0.f : (float)(ay[c] * sqrt(D[c] * D[c - g.sx]));
  }
  if (i <= nx && j <= ny) {
    az32[c] = (float)az[c];
    cz[c] = (k == 1 || k == nz + 1) ? 0.f : (float)(az[c] * sqrt(D[c] * D[c - g.sz]));
  }
  if (i <= nx && j <= ny && k <= nz) sD32[c] = (float)sD[c];
}

// face conductances are FP32 on the coarse levels and BF16 on the finest one
// (half the bytes of the sweeps that dominate the V-cycle; the diagonal is
// the sum of the rounded faces, so the operator stays consistent): mgc and
// to_bf16 in common.hpp

// BF16 faces of the finest level and the diagonal as their sum
__global__ void k_mg_bf16(Grid g, const float *__restrict__ cx, const float *__restrict__ cy,
                          const float *__restrict__ cz, uint16_t *__restrict__ hx, uint16_t *__restrict__ hy,
                          uint16_t *__restrict__ hz) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  if (i > g.nx + 1 || j > g.ny + 1 || k > g.nz + 1) return;
  const int64_t c = g.idx(i, j, k);
  if (j <= g.ny && k <= g.nz) hx[c] = to_bf16(cx[c]);
  if (i <= g.nx && k <= g.nz) hy[c] = to_bf16(cy[c]);
  if (i <= g.nx && j <= g.ny) hz[c] = to_bf16(cz[c]);
}
// (and, when sD is given, its FP32 copy)
// (and, when given, the FP32 copy of sqrt(D) and the red cells' diagonal in
// the colour-split layout of the finest level)
__global__ void k_mg_diag_bf16(Grid g, const uint16_t *__restrict__ hx, const uint16_t *__restrict__ hy,
                               const uint16_t *__restrict__ hz, float *__restrict__ dg, const double *__restrict__ sD,
                               float *__restrict__ sD32, Grid h, float *__restrict__ dred) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny || k > g.nz) return;
  const int64_t c = g.idx(i, j, k);
  const float d = mgc(hx[c]) + mgc(hx[c + 1]) + mgc(hy[c]) + mgc(hy[c + g.sx]) + mgc(hz[c]) + mgc(hz[c + g.sz]);
  dg[c] = d;
  if (sD) sD32[c] = (float)sD[c];
  if (dred && ((i + j + k) & 1) == 0) dred[h.idx((i - 1) >> 1, j, k)] = d;
}
// coarse face conductance: half the sum of the four fine faces on it; and
// the coarse diagonal, the sum of the cell's six faces in the order x-, x+,
// y-, y+, z-, z+ (the sum the sweeps form), with its three upper faces
// formed again by the formula their own threads use
template <typename CT>
__global__ void k_mg_coarsen(Grid gf, Grid gc, const CT *__restrict__ fx, const CT *__restrict__ fy,
                             const CT *__restrict__ fz, float *__restrict__ cx, float *__restrict__ cy,
                             float *__restrict__ cz, float *__restrict__ dg) {
  const int I = 1 + blockIdx.x * SULI_BX + threadIdx.x, J = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            K = 1 + blockIdx.z;
  if (I > gc.nx + 1 || J > gc.ny + 1 || K > gc.nz + 1) return;
  const int64_t c = gc.idx(I, J, K);
  // the coarse faces x, y, z of coarse cell (a, b, d): its lower face in each direction
  auto fcx = [&](int a, int b, int d) {
    const int fi = 2 * a - 1, fj = 2 * b - 1, fk = 2 * d - 1;
    return 0.5f * (mgc(fx[gf.idx(fi, fj, fk)]) + mgc(fx[gf.idx(fi, fj + 1, fk)]) +
                   mgc(fx[gf.idx(fi, fj, fk + 1)]) + mgc(fx[gf.idx(fi, fj + 1, fk + 1)]));
  };
  auto fcy = [&](int a, int b, int d) {
    const int fi = 2 * a - 1, fj = 2 * b - 1, fk = 2 * d - 1;
    return 0.5f * (mgc(fy[gf.idx(fi, fj, fk)]) + mgc(fy[gf.idx(fi + 1, fj, fk)]) +
                   mgc(fy[gf.idx(fi, fj, fk + 1)]) + mgc(fy[gf.idx(fi + 1, fj, fk + 1)]));
  };
  auto fcz = [&](int a, int b, int d) {
    const int fi = 2 * a - 1, fj = 2 * b - 1, fk = 2 * d - 1;
    return 0.5f * (mgc(fz[gf.idx(fi, fj, fk)]) + mgc(fz[gf.idx(fi + 1, fj, fk)]) +
                   mgc(fz[gf.idx(fi, fj + 1, fk)]) + mgc(fz[gf.idx(fi + 1, fj + 1, fk)]));
  };
  float x0 = 0.f, y0 = 0.f, z0 = 0.f;
  if (J <= gc.ny && K <= gc.nz) cx[c] = x0 = fcx(I, J, K);
  if (I <= gc.nx && K <= gc.nz) cy[c] = y0 = fcy(I, J, K);
  if (I <= gc.nx && J <= gc.ny) cz[c] = z0 = fcz(I, J, K);
  if (dg && I <= gc.nx && J <= gc.ny && K <= gc.nz)
    dg[c] = x0 + fcx(I + 1, J, K) + y0 + fcy(I, J + 1, K) + z0 + fcz(I, J, K + 1);
}

template <typename CT>
__device__ __forceinline__ float mg_Ax(const Grid &g, int64_t c, const CT *__restrict__ cx,
                                       const CT *__restrict__ cy, const CT *__restrict__ cz,
                                       const float *__restrict__ x) {
  // sum_f c_f x_nb (wall faces have c = 0; their ghost neighbours hold 0)
  return mgc(cx[c + 1]) * x[c + 1] + mgc(cx[c]) * x[c - 1] + mgc(cy[c + g.sx]) * x[c + g.sx] +
         mgc(cy[c]) * x[c - g.sx] + mgc(cz[c + g.sz]) * x[c + g.sz] + mgc(cz[c]) * x[c - g.sz];
}

// smoothing sweeps (red and black half sweeps) on each side of the coarse correction
#ifndef SULI_MG_NU1
#define SULI_MG_NU1 SULI_MG_NU  // the same on level 1 when it has launches of its own
#endif
#ifndef SULI_MG_NU0
#define SULI_MG_NU0 2  // sweeps before and after the coarse correction on the finest level
#endif
#ifndef SULI_MG_NU
#define SULI_MG_NU 2  // sweeps before and after the coarse correction on the coarser levels
#endif

// The diagonal as the sum of the six face conductances, in k_mg_coarsen's order
// (the stored dg, bit for bit): the sweeps read the six faces anyway, so
// reading dg as well would only add bytes.
template <typename CT>
__device__ __forceinline__ float mg_dg(const Grid &g, int64_t c, const CT *__restrict__ cx,
                                       const CT *__restrict__ cy, const CT *__restrict__ cz) {
  return mgc(cx[c]) + mgc(cx[c + 1]) + mgc(cy[c]) + mgc(cy[c + g.sx]) + mgc(cz[c]) + mgc(cz[c + g.sz]);
}

// red-black Gauss-Seidel half sweep on the cells of one colour.  The 7-point
// stencil couples a cell only to cells of the other colour, so a half sweep
// reads only the other colour; with zero_x set that colour is taken as zero
// (the first sweep from x = 0), and its stale values are overwritten by the
// next half sweep before anything reads them.
template <typename CT>
__global__ __launch_bounds__(256) void k_mg_gs(Grid g, const CT *__restrict__ cx, const CT *__restrict__ cy,
                                               const CT *__restrict__ cz, const float *__restrict__ b,
                                               float *__restrict__ x, int color, int zero_x, const int *active) {
  if (!*active) return;
  const int ti = blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  if (j > g.ny || k > g.nz) return;
  const int i = 1 + 2 * ti + ((1 + j + k + color) & 1);
  if (i > g.nx) return;
  const int64_t c = g.idx(i, j, k);
  const float s = zero_x ? 0.f : mg_Ax(g, c, cx, cy, cz, x);
  x[c] = (b[c] + s) / mg_dg(g, c, cx, cy, cz);
}

// The coarse level's x after restriction: 0, or with the coarse FP32 faces
// given, the first red half sweep from x = 0 on its red cells (colour 0:
// i + j + k even), (b + 0) / diagonal as k_mg_gs computes it with zero_x.
// That saves the coarse level's first launch; its black cells stay 0 until
// the black half sweep that follows.
__device__ __forceinline__ float coarse_red_x(const Grid &gc, int64_t cc, int ijk, float bcv,
                                              const float *__restrict__ ccx, const float *__restrict__ ccy,
                                              const float *__restrict__ ccz) {
  if (!ccx || (ijk & 1)) return 0.f;
  const float s = 0.f;
  return (bcv + s) / mg_dg(gc, cc, ccx, ccy, ccz);
}

// coarse rhs = sum over the 8 children of the fine residual b - A x; coarse x = 0
// (or its first red half sweep, coarse_red_x)
template <typename CT>
__global__ __launch_bounds__(256) void k_mg_restrict(Grid gf, Grid gc, const CT *__restrict__ cx,
                                                     const CT *__restrict__ cy, const CT *__restrict__ cz,
                                                     const float *__restrict__ b, const float *__restrict__ x,
                                                     float *__restrict__ bc, float *__restrict__ xc,
                                                     const float *__restrict__ ccx, const float *__restrict__ ccy,
                                                     const float *__restrict__ ccz, const int *active) {
  if (!*active) return;
  const int I = 1 + blockIdx.x * SULI_BX + threadIdx.x, J = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            K = 1 + blockIdx.z;
  if (I > gc.nx || J > gc.ny) return;
  float s = 0.f;
  const int64_t c0 = gf.idx(2 * I - 1, 2 * J - 1, 2 * K - 1);
#pragma unroll
  for (int d = 0; d < 8; d++) {
    const int64_t c = c0 + (d & 1) + ((d >> 1) & 1) * (int64_t)gf.sx + (d >> 2) * (int64_t)gf.sz;
    s += b[c] - (mg_dg(gf, c, cx, cy, cz) * x[c] - mg_Ax(gf, c, cx, cy, cz, x));
  }
  const int64_t cc = gc.idx(I, J, K);
  bc[cc] = s;
  xc[cc] = coarse_red_x(gc, cc, I + J + K, s, ccx, ccy, ccz);
}

// ---- half sweeps fused with their neighbours in the V-cycle ----------------------
// A thread of these kernels owns the cell pair (2ti+1, 2ti+2) of row (j, k),
// one red and one black cell; its cell of the colour being updated sits
// where k_mg_gs puts it.

// First pre-smoothing half sweep of the finest level with the input scaling:
// b = s r on both cells of the pair (k_scale), x = b/dg on the red one (the
// zero_x sweep of k_mg_gs).
__global__ __launch_bounds__(256) void k_mg_gs_in(Grid g, const float *__restrict__ dg, const float *__restrict__ s,
                                                  const float *__restrict__ r, float *__restrict__ b,
                                                  float *__restrict__ x, const int *active) {
  if (!*active) return;
  const int ti = blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  const int i1 = 1 + 2 * ti;
  if (j > g.ny || k > g.nz || i1 > g.nx) return;
  const int64_t c1 = g.idx(i1, j, k);
  const bool two = i1 < g.nx;
  const float b1 = s[c1] * r[c1], b2 = two ? s[c1 + 1] * r[c1 + 1] : 0.f;
  b[c1] = b1;
  if (two) b[c1 + 1] = b2;
  if (((i1 + j + k) & 1) == 0)
    x[c1] = (b1 + 0.f) / dg[c1];
  else if (two)
    x[c1 + 1] = (b2 + 0.f) / dg[c1 + 1];
}

// First post-smoothing half sweep (black) with the prolongation: the red
// neighbours are read as x + x_coarse(parent), the values k_mg_prolong would
// have stored.  Neither colour's prolongated value is stored: this sweep
// overwrites the black cells and the next one the red cells without reading
// their old values.
template <typename CT>
__global__ __launch_bounds__(256) void k_mg_gs_prolong(Grid g, Grid gc, const CT *__restrict__ cx,
                                                       const CT *__restrict__ cy, const CT *__restrict__ cz,
                                                       const float *__restrict__ b, float *__restrict__ x,
                                                       const float *__restrict__ xc, const int *active) {
  if (!*active) return;
  const int ti = blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  if (j > g.ny || k > g.nz) return;
  const int i = 1 + 2 * ti + ((j + k) & 1);  // colour 1
  if (i > g.nx) return;
  const int64_t c = g.idx(i, j, k);
  const auto X = [&](int64_t cn, int in, int jn, int kn) {
    return x[cn] + xc[gc.idx((in + 1) >> 1, (jn + 1) >> 1, (kn + 1) >> 1)];
  };
  const float s = mgc(cx[c + 1]) * X(c + 1, i + 1, j, k) + mgc(cx[c]) * X(c - 1, i - 1, j, k) +
                  mgc(cy[c + g.sx]) * X(c + g.sx, i, j + 1, k) + mgc(cy[c]) * X(c - g.sx, i, j - 1, k) +
                  mgc(cz[c + g.sz]) * X(c + g.sz, i, j, k + 1) + mgc(cz[c]) * X(c - g.sz, i, j, k - 1);
  x[c] = (b[c] + s) / mg_dg(g, c, cx, cy, cz);
}

// Last post-smoothing half sweep (red) of the finest level with the output
// scaling z = s x on both cells of the pair (k_scale) and, when partials is
// set and *first is 0, the block sums of z.q (k_zq).  x itself is not
// stored: nothing reads it before the next V-cycle starts from x = 0.
template <typename CT>
__global__ __launch_bounds__(256) void k_mg_gs_out(Grid g, const CT *__restrict__ cx, const CT *__restrict__ cy,
                                                   const CT *__restrict__ cz, const float *__restrict__ b,
                                                   const float *__restrict__ x,
                                                   const float *__restrict__ s, float *__restrict__ z,
                                                   const float *__restrict__ q, const int *first,
                                                   double *partials, const int *active) {
  if (!*active) return;
  const int ti = blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  const int i1 = 1 + 2 * ti;
  double acc[1] = {0.0};
  const bool dot = partials && !*first;
  if (j <= g.ny && k <= g.nz && i1 <= g.nx) {
    const int i = i1 + ((1 + j + k) & 1);  // colour 0
    const int64_t c1 = g.idx(i1, j, k);
    const bool two = i1 < g.nx;
    if (i <= g.nx) {
      const int64_t c = c1 + (int64_t)(i - i1), cb = i == i1 ? c1 + 1 : c1;
      const float xr = (b[c] + mg_Ax(g, c, cx, cy, cz, x)) / mg_dg(g, c, cx, cy, cz);
      const float zr = s[c] * xr;
      z[c] = zr;
      if (dot) acc[0] += (double)zr * (double)q[c];
      if (two) {
        const float zb = s[cb] * x[cb];
        z[cb] = zb;
        if (dot) acc[0] += (double)zb * (double)q[cb];
      }
    } else {  // the pair's only cell is black
      const float zb = s[c1] * x[c1];
      z[c1] = zb;
      if (dot) acc[0] += (double)zb * (double)q[c1];
    }
  }
  if (!partials) return;
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1>(acc, lds);
  block_partial<1>(acc, partials);
}

// ---- the finest level with its iterate and right-hand side split by colour ---------
// sum_f c_f x_f of a half sweep as explicit fused multiply-adds in a fixed
// order, so that every kernel forming it (one or two cells per thread)
// rounds it the same way
__device__ __forceinline__ float nb_sum6(float c1, float x1, float c2, float x2, float c3, float x3, float c4,
                                         float x4, float c5, float x5, float c6, float x6) {
  return fmaf(c6, x6, fmaf(c5, x5, fmaf(c4, x4, fmaf(c3, x3, fmaf(c2, x2, c1 * x1)))));
}

// A half sweep reads the other colour's x and writes its own colour's x and
// reads its b: in the frame layout each of those touches every cache line of
// the array, twice the bytes the colour needs.  With the fused kernels the
// finest level keeps x, b (and the diagonal used by the first sweep) in two
// compact arrays, one per colour: the cell of colour col of pair p (cells
// 2p+1, 2p+2) of row (j, k) sits at h.idx(p, j, k) of array col, where h is a
// frame of (nx+1)/2 pairs with one ghost pair/row/plane on each side (all
// zero).  The faces stay in the frame layout.  Seen from a cell of pair p
// that is the second of its pair (sec = 1), the x-neighbours of the other
// colour are pairs p and p+1; for the first (sec = 0) pairs p-1 and p; the
// y- and z-neighbours are pair p of the neighbouring rows.

// first pre-smoothing half sweep from x = 0, with the input scaling b = s r
template <typename XT>
__global__ __launch_bounds__(256) void k_mg0_in(Grid g, Grid h, const float *__restrict__ dred,
                                                const float *__restrict__ s, const float *__restrict__ r,
                                                XT *__restrict__ bred, XT *__restrict__ bblack,
                                                XT *__restrict__ xred, const int *active) {
  if (!*active) return;
  int bx, by, bz;
  xcd_band_block(bx, by, bz);
  const int p = bx * SULI_BX + threadIdx.x, j = 1 + by * SULI_BY + threadIdx.y, k = 1 + bz;
  const int i1 = 1 + 2 * p;
  if (j > g.ny || k > g.nz || i1 > g.nx) return;
  const int64_t c1 = g.idx(i1, j, k), q = h.idx(p, j, k);
  const bool two = i1 < g.nx;
  const float b1 = s[c1] * r[c1], b2 = two ? s[c1 + 1] * r[c1 + 1] : 0.f;
  if (((i1 + j + k) & 1) == 0) {  // the first cell is red
    stage_cell(bred, bblack, xred, dred, q, true, b1);
    if (two) stage_cell(bred, bblack, xred, dred, q, false, b2);
  } else {
    stage_cell(bred, bblack, xred, dred, q, false, b1);
    if (two) stage_cell(bred, bblack, xred, dred, q, true, b2);
  }
}

// half sweep of colour `color` (x of the colour from b and the other colour's x)
template <typename XT>
__global__ __launch_bounds__(256) void k_mg0_gs(Grid g, Grid h, const uint16_t *__restrict__ cx,
                                                const uint16_t *__restrict__ cy, const uint16_t *__restrict__ cz,
                                                const XT *__restrict__ bcol, const XT *__restrict__ xo,
                                                XT *__restrict__ xcol, int color, const int *active) {
  if (!*active) return;
  int bx, by, bz;
  xcd_band_block(bx, by, bz);
  const int p = bx * SULI_BX + threadIdx.x, j = 1 + by * SULI_BY + threadIdx.y, k = 1 + bz;
  if (j > g.ny || k > g.nz) return;
  const int sec = (1 + j + k + color) & 1;
  const int i = 1 + 2 * p + sec;
  if (i > g.nx) return;
  const int64_t c = g.idx(i, j, k), q = h.idx(p, j, k);
  const float s = nb_sum6(mgc(cx[c + 1]), mgc(xo[q + sec]), mgc(cx[c]), mgc(xo[q - 1 + sec]), mgc(cy[c + g.sx]),
                          mgc(xo[q + h.sx]), mgc(cy[c]), mgc(xo[q - h.sx]), mgc(cz[c + g.sz]), mgc(xo[q + h.sz]),
                          mgc(cz[c]), mgc(xo[q - h.sz]));
  xcol[q] = to_x<XT>((mgc(bcol[q]) + s) / mg_dg(g, c, cx, cy, cz));
}

// The same half sweep with two cells of the colour per thread: those of
// pairs 2t and 2t+1 of a row, i0 = 2(2t)+1+sec and i0+2.  In a row sec is
// uniform, the pair index 2t is even and the face index of i0 is sec modulo
// 4 (the frame's x offset is 3 modulo 4, its pitches multiples of 4; checked
// on the host), so the two cells' x, b and other-colour neighbours in y and
// z come as aligned pairs, and each face array's values as one aligned
// group of four BF16 (x faces of i0..i0+3, or of i0-1..i0+2 and i0+3 when
// sec = 1; y and z faces at positions sec and sec+2).  Each cell's update is
// k_mg0_gs's, term for term (nb_sum6 fixes the rounding of both).
template <typename XT> struct Pair2;
template <> struct Pair2<float> { typedef float __attribute__((ext_vector_type(2))) t; };
template <> struct Pair2<_Float16> { typedef _Float16 __attribute__((ext_vector_type(2))) t; };
__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
// the four BF16 of the aligned group at a, element e (0..3)
struct Bf4 {
  uint2 w;
  __device__ __forceinline__ float at(int e) const {
    const uint32_t h = e >= 2 ? w.y : w.x;
    return (e & 1) ? bf_hi(h) : bf_lo(h);
  }
};
__device__ __forceinline__ Bf4 ld_bf4(const uint16_t *__restrict__ p, fidx a) {
  return Bf4{*(const uint2 *)(p + a)};
}

// The faces and other-colour values around the two cells i0 = 1+2p0+sec
// and i0+2 (pairs p0 = 2t and p0+1) of row (j, k) that a thread of the pair
// kernels updates; xo is the other colour's array, q0 = h.idx(p0, j, k).
struct PairNb {
  float ax[4];                            // x faces at c0, c0+1, c0+2, c0+3
  float fy[2], fyu[2], fz[2], fzu[2];     // y and z faces (lower, upper) of each cell
  float xw0, xe0, xe1;                    // other colour at i0-1, i0+1, i0+3
  float ys[2], yn[2], zs[2], zn[2];       // other colour at j-1, j+1, k-1, k+1 of each cell
  // the half-sweep update of cell n (0: i0, 1: i0+2) from its b, term for
  // term k_mg0_gs's
  __device__ __forceinline__ float update(int n, float b) const {
    const float xw = n ? xe0 : xw0, xe = n ? xe1 : xe0;
    const float s = nb_sum6(ax[2 * n + 1], xe, ax[2 * n], xw, fyu[n], yn[n], fy[n], ys[n], fzu[n], zn[n], fz[n],
                            zs[n]);
    return (b + s) / (ax[2 * n] + ax[2 * n + 1] + fy[n] + fyu[n] + fz[n] + fzu[n]);
  }
};
template <typename XT>
__device__ __forceinline__ PairNb load_pair_nb(const Grid &g, const Grid &h, const uint16_t *__restrict__ cx,
                                               const uint16_t *__restrict__ cy, const uint16_t *__restrict__ cz,
                                               const XT *__restrict__ xo, fidx c0, fidx q0, int sec) {
  typedef typename Pair2<XT>::t X2;
  PairNb r;
  const fidx a = c0 - sec;
  const X2 xm = *(const X2 *)(xo + q0);
  const float xe = mgc(xo[q0 - 1 + 3 * sec]);  // the x-neighbour outside the two pairs
  const X2 ysn = *(const X2 *)(xo + q0 - h.sx), ynn = *(const X2 *)(xo + q0 + h.sx);
  const X2 zsn = *(const X2 *)(xo + q0 - h.sz), znn = *(const X2 *)(xo + q0 + h.sz);
  const Bf4 fx = ld_bf4(cx, a), fy = ld_bf4(cy, a), fyu = ld_bf4(cy, a + g.sx), fz = ld_bf4(cz, a),
            fzu = ld_bf4(cz, a + g.sz);
  const float fx4 = sec ? mgc(cx[c0 + 3]) : 0.f;  // x face c0+3 (the next group when sec = 1)
  r.ax[0] = fx.at(sec);
  r.ax[1] = fx.at(sec + 1);
  r.ax[2] = fx.at(sec + 2);
  r.ax[3] = sec ? fx4 : fx.at(3);
  // i0-1, i0+1, i0+3 hold xo[q0-1+sec], xo[q0+sec], xo[q0+1+sec]
  r.xw0 = sec ? mgc(xm.x) : xe;
  r.xe0 = sec ? mgc(xm.y) : mgc(xm.x);
  r.xe1 = sec ? xe : mgc(xm.y);
#pragma unroll
  for (int n = 0; n < 2; n++) {
    r.fy[n] = fy.at(sec + 2 * n);
    r.fyu[n] = fyu.at(sec + 2 * n);
    r.fz[n] = fz.at(sec + 2 * n);
    r.fzu[n] = fzu.at(sec + 2 * n);
  }
  r.ys[0] = mgc(ysn.x), r.ys[1] = mgc(ysn.y), r.yn[0] = mgc(ynn.x), r.yn[1] = mgc(ynn.y);
  r.zs[0] = mgc(zsn.x), r.zs[1] = mgc(zsn.y), r.zn[0] = mgc(znn.x), r.zn[1] = mgc(znn.y);
  return r;
}

template <typename XT>
__global__ __launch_bounds__(256) void k_mg0_gs2(Grid g, Grid h, const uint16_t *__restrict__ cx,
                                                 const uint16_t *__restrict__ cy, const uint16_t *__restrict__ cz,
                                                 const XT *__restrict__ bcol, const XT *__restrict__ xo,
                                                 XT *__restrict__ xcol, int color, const int *active) {
  typedef typename Pair2<XT>::t X2;
  if (!*active) return;
  int bx, by, bz;
  xcd_band_block(bx, by, bz);
  const int p0 = 2 * (bx * SULI_BX + threadIdx.x), j = 1 + by * SULI_BY + threadIdx.y, k = 1 + bz;
  if (j > g.ny || k > g.nz) return;
  const int sec = (1 + j + k + color) & 1;
  const int i0 = 1 + 2 * p0 + sec;
  if (i0 > g.nx) return;
  const fidx c0 = g.idx(i0, j, k), q0 = h.idx(p0, j, k);
  const PairNb nb = load_pair_nb(g, h, cx, cy, cz, xo, c0, q0, sec);
  const X2 bb = *(const X2 *)(bcol + q0);
  const float x0 = nb.update(0, mgc(bb.x));
  if (i0 + 2 > g.nx) {  // the second cell is missing
    xcol[q0] = to_x<XT>(x0);
    return;
  }
  X2 out;
  out.x = to_x<XT>(x0);
  out.y = to_x<XT>(nb.update(1, mgc(bb.y)));
  *(X2 *)(xcol + q0) = out;
}

// k_mg0_out with two pairs per thread: the red cells i0 = i1+sec and i0+2 of
// cells i1 = 4t+1 .. i1+3 updated as in k_mg0_gs2, z = s x (/ *sig) of the
// four cells (an aligned float4 of the frame: the x offset 3 modulo 4) and
// their z.q; a row end with fewer than four cells takes the cells one by one
template <typename XT>
__global__ __launch_bounds__(256) void k_mg0_out2(Grid g, Grid h, const uint16_t *__restrict__ cx,
                                                  const uint16_t *__restrict__ cy, const uint16_t *__restrict__ cz,
                                                  const XT *__restrict__ bred, const XT *__restrict__ xblack,
                                                  const float *__restrict__ s, float *__restrict__ z,
                                                  const float *__restrict__ q, const int *first, double *partials,
                                                  const int *active, const float *sig) {
  typedef typename Pair2<XT>::t X2;
  if (!*active) return;
  int bx, by, bz;
  xcd_band_block(bx, by, bz);
  const int p0 = 2 * (bx * SULI_BX + threadIdx.x), j = 1 + by * SULI_BY + threadIdx.y, k = 1 + bz;
  const int i1 = 1 + 2 * p0;
  double acc[1] = {0.0};
  const bool dot = partials && !*first;
  const float isg = sig ? 1.f / *sig : 1.f;
  if (j <= g.ny && k <= g.nz && i1 <= g.nx) {
    const int sec = (1 + j + k) & 1;  // colour 0: the red cells are i1 + sec and i1 + 2 + sec
    const fidx c1 = g.idx(i1, j, k), q0 = h.idx(p0, j, k);
    // black values of the two pairs (also the x-neighbours the red updates read)
    if (i1 + 3 <= g.nx) {
      const PairNb nb = load_pair_nb(g, h, cx, cy, cz, xblack, c1 + sec, q0, sec);
      const X2 bb = *(const X2 *)(bred + q0);
      const X2 xb = *(const X2 *)(xblack + q0);
      const float xr0 = nb.update(0, mgc(bb.x)), xr1 = nb.update(1, mgc(bb.y));
      const float4 sv = *(const float4 *)(s + c1);
      // cells i1..i1+3: red at i1+sec and i1+2+sec, black at the others
      float4 zv;
      zv.x = sv.x * (sec ? mgc(xb.x) : xr0) * isg;
      zv.y = sv.y * (sec ? xr0 : mgc(xb.x)) * isg;
      zv.z = sv.z * (sec ? mgc(xb.y) : xr1) * isg;
      zv.w = sv.w * (sec ? xr1 : mgc(xb.y)) * isg;
      *(float4 *)(z + c1) = zv;
      if (dot) {
        const float4 qv = *(const float4 *)(q + c1);
        acc[0] += (double)zv.x * (double)qv.x + (double)zv.y * (double)qv.y + (double)zv.z * (double)qv.z +
                  (double)zv.w * (double)qv.w;
      }
    } else {  // fewer than four cells: each pair as k_mg0_out takes it
      for (int pp = 0; pp < 2; pp++) {
        const int ia = i1 + 2 * pp;
        if (ia > g.nx) break;
        const fidx ca = c1 + 2 * pp, qh = q0 + pp;
        const bool twoc = ia < g.nx;
        const int ir = ia + sec;
        if (ir <= g.nx) {
          const fidx c = ca + sec, cb = sec ? ca : ca + 1;
          const float sx = nb_sum6(mgc(cx[c + 1]), mgc(xblack[qh + sec]), mgc(cx[c]), mgc(xblack[qh - 1 + sec]),
                                   mgc(cy[c + g.sx]), mgc(xblack[qh + h.sx]), mgc(cy[c]), mgc(xblack[qh - h.sx]),
                                   mgc(cz[c + g.sz]), mgc(xblack[qh + h.sz]), mgc(cz[c]), mgc(xblack[qh - h.sz]));
          const float xr = (mgc(bred[qh]) + sx) / mg_dg(g, c, cx, cy, cz);
          const float zr = s[c] * xr * isg;
          z[c] = zr;
          if (dot) acc[0] += (double)zr * (double)q[c];
          if (twoc) {
            const float zb = s[cb] * mgc(xblack[qh]) * isg;
            z[cb] = zb;
            if (dot) acc[0] += (double)zb * (double)q[cb];
          }
        } else {
          const float zb = s[ca] * mgc(xblack[qh]) * isg;
          z[ca] = zb;
          if (dot) acc[0] += (double)zb * (double)q[ca];
        }
      }
    }
  }
  if (!partials) return;
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1>(acc, lds);
  block_partial<1>(acc, partials);
}

// first post-smoothing half sweep (black) with the prolongation folded into
// the red neighbours it reads (k_mg_gs_prolong)
template <typename XT>
__global__ __launch_bounds__(256) void k_mg0_prolong(Grid g, Grid h, Grid gc, const uint16_t *__restrict__ cx,
                                                     const uint16_t *__restrict__ cy,
                                                     const uint16_t *__restrict__ cz, const XT *__restrict__ bblack,
                                                     const XT *__restrict__ xred, XT *__restrict__ xblack,
                                                     const float *__restrict__ xc, const int *active) {
  if (!*active) return;
  int bx, by, bz;
  xcd_band_block(bx, by, bz);
  const int p = bx * SULI_BX + threadIdx.x, j = 1 + by * SULI_BY + threadIdx.y, k = 1 + bz;
  if (j > g.ny || k > g.nz) return;
  const int sec = (j + k) & 1;  // colour 1
  const int i = 1 + 2 * p + sec;
  if (i > g.nx) return;
  const fidx c = g.idx(i, j, k), q = h.idx(p, j, k);
  // The six neighbours' coarse parents are four cells: the cell's own parent
  // (that of one neighbour on each axis) and, on each axis, the parent across
  // the coarse face (the lower one from an odd index, the upper from an even)
  const fidx P = gc.idx((i + 1) >> 1, (j + 1) >> 1, (k + 1) >> 1);
  const float xp = xc[P], xpx = xc[P + ((i & 1) ? (fidx)-1 : (fidx)1)],
              xpy = xc[P + ((j & 1) ? (fidx)0 - gc.sx : (fidx)gc.sx)],
              xpz = xc[P + ((k & 1) ? (fidx)0 - gc.sz : gc.sz)];
  const float xe = mgc(xred[q + sec]) + ((i & 1) ? xp : xpx), xw = mgc(xred[q - 1 + sec]) + ((i & 1) ? xpx : xp);
  const float xn = mgc(xred[q + h.sx]) + ((j & 1) ? xp : xpy), xs = mgc(xred[q - h.sx]) + ((j & 1) ? xpy : xp);
  const float xt = mgc(xred[q + h.sz]) + ((k & 1) ? xp : xpz), xb = mgc(xred[q - h.sz]) + ((k & 1) ? xpz : xp);
  const float s = nb_sum6(mgc(cx[c + 1]), xe, mgc(cx[c]), xw, mgc(cy[c + g.sx]), xn, mgc(cy[c]), xs,
                          mgc(cz[c + g.sz]), xt, mgc(cz[c]), xb);
  xblack[q] = to_x<XT>((mgc(bblack[q]) + s) / mg_dg(g, c, cx, cy, cz));
}

// last post-smoothing half sweep (red) with the output scaling z = s x on both
// cells of the pair (frame layout) and the block sums of z.q (k_mg_gs_out);
// with sig set, z = s x / *sig
template <typename XT>
__global__ __launch_bounds__(256) void k_mg0_out(Grid g, Grid h, const uint16_t *__restrict__ cx,
                                                 const uint16_t *__restrict__ cy, const uint16_t *__restrict__ cz,
                                                 const XT *__restrict__ bred, const XT *__restrict__ xblack,
                                                 const float *__restrict__ s, float *__restrict__ z,
                                                 const float *__restrict__ q, const int *first, double *partials,
                                                 const int *active, const float *sig) {
  if (!*active) return;
  int bx, by, bz;
  xcd_band_block(bx, by, bz);
  const int p = bx * SULI_BX + threadIdx.x, j = 1 + by * SULI_BY + threadIdx.y, k = 1 + bz;
  const int i1 = 1 + 2 * p;
  double acc[1] = {0.0};
  const bool dot = partials && !*first;
  const float isg = sig ? 1.f / *sig : 1.f;  // the input scale divided out (FP16 storage)
  if (j <= g.ny && k <= g.nz && i1 <= g.nx) {
    const int sec = (1 + j + k) & 1;  // colour 0
    const int i = i1 + sec;
    const int64_t c1 = g.idx(i1, j, k), qh = h.idx(p, j, k);
    const bool two = i1 < g.nx;
    if (i <= g.nx) {
      const int64_t c = c1 + sec, cb = sec ? c1 : c1 + 1;
      const float sx = nb_sum6(mgc(cx[c + 1]), mgc(xblack[qh + sec]), mgc(cx[c]), mgc(xblack[qh - 1 + sec]),
                               mgc(cy[c + g.sx]), mgc(xblack[qh + h.sx]), mgc(cy[c]), mgc(xblack[qh - h.sx]),
                               mgc(cz[c + g.sz]), mgc(xblack[qh + h.sz]), mgc(cz[c]), mgc(xblack[qh - h.sz]));
      const float xr = (mgc(bred[qh]) + sx) / mg_dg(g, c, cx, cy, cz);
      const float zr = s[c] * xr * isg;
      z[c] = zr;
      if (dot) acc[0] += (double)zr * (double)q[c];
      if (two) {
        const float zb = s[cb] * mgc(xblack[qh]) * isg;
        z[cb] = zb;
        if (dot) acc[0] += (double)zb * (double)q[cb];
      }
    } else {  // the pair's only cell is black
      const float zb = s[c1] * mgc(xblack[qh]) * isg;
      z[c1] = zb;
      if (dot) acc[0] += (double)zb * (double)q[c1];
    }
  }
  if (!partials) return;
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1>(acc, lds);
  block_partial<1>(acc, partials);
}

// restriction of the residual to level 1 (k_mg_restrict on the split arrays)
template <typename XT>
__global__ __launch_bounds__(256) void k_mg0_restrict(Grid gf, Grid h, Grid gc, const uint16_t *__restrict__ cx,
                                                      const uint16_t *__restrict__ cy,
                                                      const uint16_t *__restrict__ cz, const XT *__restrict__ b0,
                                                      const XT *__restrict__ b1, const XT *__restrict__ x0,
                                                      const XT *__restrict__ x1, float *__restrict__ bc,
                                                      float *__restrict__ xc, const float *__restrict__ ccx,
                                                      const float *__restrict__ ccy, const float *__restrict__ ccz,
                                                      const int *active) {
  if (!*active) return;
  int bx, by, bz;
  xcd_band_block(bx, by, bz);
  const int I = 1 + bx * SULI_BX + threadIdx.x, J = 1 + by * SULI_BY + threadIdx.y, K = 1 + bz;
  if (I > gc.nx || J > gc.ny) return;
  float s = 0.f;
  const int64_t c0 = gf.idx(2 * I - 1, 2 * J - 1, 2 * K - 1), q0 = h.idx(I - 1, 2 * J - 1, 2 * K - 1);
#pragma unroll
  for (int d = 0; d < 8; d++) {
    const int di = d & 1, dj = (d >> 1) & 1, dk = d >> 2;
    const int64_t c = c0 + di + dj * (int64_t)gf.sx + dk * (int64_t)gf.sz;
    const int64_t q = q0 + dj * (int64_t)h.sx + dk * (int64_t)h.sz;
    const bool red = ((di + dj + dk + 2 * I - 1 + 2 * J - 1 + 2 * K - 1) & 1) == 0;
    const XT *xs = red ? x0 : x1, *xo = red ? x1 : x0;
    const float bv = mgc(red ? b0[q] : b1[q]);
    const float ax = mgc(cx[c + 1]) * mgc(xo[q + di]) + mgc(cx[c]) * mgc(xo[q - 1 + di]) +
                     mgc(cy[c + gf.sx]) * mgc(xo[q + h.sx]) + mgc(cy[c]) * mgc(xo[q - h.sx]) +
                     mgc(cz[c + gf.sz]) * mgc(xo[q + h.sz]) + mgc(cz[c]) * mgc(xo[q - h.sz]);
    s += bv - (mg_dg(gf, c, cx, cy, cz) * mgc(xs[q]) - ax);
  }
  const int64_t cc = gc.idx(I, J, K);
  bc[cc] = s;
  xc[cc] = coarse_red_x(gc, cc, I + J + K, s, ccx, ccy, ccz);
}

// x_fine += x_coarse(parent)
__global__ __launch_bounds__(256) void k_mg_prolong(Grid gf, Grid gc, const float *__restrict__ xc,
                                                    float *__restrict__ x, const int *active) {
  if (!*active) return;
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  if (i > gf.nx || j > gf.ny) return;
  x[gf.idx(i, j, k)] += xc[gc.idx((i + 1) >> 1, (j + 1) >> 1, (k + 1) >> 1)];
}

// ---- fused smoothing of the small multigrid levels ------------------------------
// On a level of a few thousand to tens of thousands of cells each half sweep
// is a latency-bound launch.  These kernels run all the half sweeps of one
// side of the V-cycle in one launch: a work-group holds its tile of T^3 cells
// and a halo of NH cells (one per half sweep) in LDS, repeats the half
// sweeps on a region that shrinks by one cell per half sweep, and writes only
// its tile.  The arithmetic of each cell update is k_mg_gs's, so the tile
// ends with the values of the separate launches.  The pre-smoothing kernel
// ends with the restriction of the residual to the next level, the
// post-smoothing kernel starts with the prolongation from it.
#define SULI_MG_FT 8                  // tile edge (cells); T^3 threads per work-group
#define SULI_MG_FNH (2 * SULI_MG_NU)  // half sweeps per side of the V-cycle

// sum_f c_f x_nb from the LDS tile (edge E), in mg_Ax's order
template <int E>
__device__ __forceinline__ float tile_Ax(int l, const float *fx, const float *fy, const float *fz, const float *sx) {
  return fx[l + 1] * sx[l + 1] + fx[l] * sx[l - 1] + fy[l + E] * sx[l + E] + fy[l] * sx[l - E] +
         fz[l + E * E] * sx[l + E * E] + fz[l] * sx[l - E * E];
}

// Load b, dg and the three face arrays over the tile and its halo (local
// index a <-> cell i0 + a per direction); cells outside the level's frame
// read as 0.  x is the prolongated iterate (post) or 0 (pre, x = 0 on entry).
template <int E, int NT>
__device__ __forceinline__ void mg_tile_load(const Grid &g, const Grid &gc, int i0, int j0, int k0,
                                             const float *__restrict__ cx, const float *__restrict__ cy,
                                             const float *__restrict__ cz, const float *__restrict__ dg,
                                             const float *__restrict__ b, const float *__restrict__ x,
                                             const float *__restrict__ xc, float *sx, float *sb, float *sd,
                                             float *fx, float *fy, float *fz) {
  // all loads of a thread are issued before its LDS stores (E^3 = NPT * NT)
  constexpr int NPT = E * E * E / NT;
  static_assert(NPT * NT == E * E * E, "LDS region must split evenly over the work-group");
  float vx[NPT], vb[NPT], vd[NPT], v1[NPT], v2[NPT], v3[NPT];
#pragma unroll
  for (int q = 0; q < NPT; q++) {
    const int t = threadIdx.x + q * NT;
    const int i = i0 + t % E, j = j0 + (t / E) % E, k = k0 + t / (E * E);
    vx[q] = vb[q] = vd[q] = v1[q] = v2[q] = v3[q] = 0.f;
    if (i >= 0 && i <= g.nx + 1 && j >= 0 && j <= g.ny + 1 && k >= 0 && k <= g.nz + 1) {
      const int64_t c = g.idx(i, j, k);
      vb[q] = b[c];
      vd[q] = dg[c];
      v1[q] = cx[c];
      v2[q] = cy[c];
      v3[q] = cz[c];
      if (x && i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz)
        vx[q] = x[c] + xc[gc.idx((i + 1) >> 1, (j + 1) >> 1, (k + 1) >> 1)];
    }
  }
#pragma unroll
  for (int q = 0; q < NPT; q++) {
    const int t = threadIdx.x + q * NT;
    sx[t] = vx[q];
    sb[t] = vb[q];
    sd[t] = vd[q];
    fx[t] = v1[q];
    fy[t] = v2[q];
    fz[t] = v3[q];
  }
  __syncthreads();
}

// One half sweep of colour `color` over the local box [lo, lo + w)^3 (w even).
template <int E>
__device__ __forceinline__ void mg_tile_gs(const Grid &g, int i0, int j0, int k0, int lo, int w, int color,
                                           bool zero_x, float *sx, const float *sb, const float *sd,
                                           const float *fx, const float *fy, const float *fz) {
  const int hw = w / 2;
  for (int t = threadIdx.x; t < hw * w * w; t += blockDim.x) {
    const int bb = lo + (t / hw) % w, cc = lo + t / (hw * w);
    const int j = j0 + bb, k = k0 + cc;
    const int a = lo + 2 * (t % hw) + ((color - (i0 + lo + j + k)) & 1);  // (i + j + k) & 1 == color
    const int i = i0 + a;
    if (i < 1 || i > g.nx || j < 1 || j > g.ny || k < 1 || k > g.nz) continue;
    const int l = a + E * (bb + E * cc);
    const float s = zero_x ? 0.f : tile_Ax<E>(l, fx, fy, fz, sx);
    sx[l] = (sb[l] + s) / sd[l];
  }
  __syncthreads();
}

// Pre-smoothing from x = 0 (half sweeps red, black, ...), then x of the tile
// to the level and the coarse rhs bc = sum of the 8 children's residuals
// (coarse x = 0), as k_mg_restrict computes them.
template <int T, int NH>
__global__ __launch_bounds__(T *T *T) void k_mg_pre_fused(Grid g, Grid gc, const float *__restrict__ cx,
                                                         const float *__restrict__ cy, const float *__restrict__ cz,
                                                         const float *__restrict__ dg, const float *__restrict__ b,
                                                         float *__restrict__ x, float *__restrict__ bc,
                                                         float *__restrict__ xc, const int *active) {
  if (!*active) return;
  constexpr int E = T + 2 * NH, N = E * E * E;
  __shared__ float sx[N], sb[N], sd[N], fx[N], fy[N], fz[N];
  __shared__ float rt[T * T * T];
  const int i0 = 1 + blockIdx.x * T - NH, j0 = 1 + blockIdx.y * T - NH, k0 = 1 + blockIdx.z * T - NH;
  mg_tile_load<E, T * T * T>(g, gc, i0, j0, k0, cx, cy, cz, dg, b, nullptr, nullptr, sx, sb, sd, fx, fy, fz);
  // half sweep h updates the box tile + (NH - h): after NH of them x is final on tile + 1
#pragma unroll
  for (int h = 0; h < NH; h++) mg_tile_gs<E>(g, i0, j0, k0, h, E - 2 * h, h & 1, h == 0, sx, sb, sd, fx, fy, fz);
  const int t = threadIdx.x;
  const int a = NH + t % T, bb = NH + (t / T) % T, cc = NH + t / (T * T);
  const int i = i0 + a, j = j0 + bb, k = k0 + cc;
  float r = 0.f;
  if (i <= g.nx && j <= g.ny && k <= g.nz) {
    const int l = a + E * (bb + E * cc);
    x[g.idx(i, j, k)] = sx[l];
    r = sb[l] - (sd[l] * sx[l] - tile_Ax<E>(l, fx, fy, fz, sx));
  }
  rt[t] = r;
  __syncthreads();
  constexpr int H = T / 2;
  if (t < H * H * H) {
    const int u = t % H, v = (t / H) % H, w = t / (H * H);
    const int I = 1 + blockIdx.x * H + u, J = 1 + blockIdx.y * H + v, K = 1 + blockIdx.z * H + w;
    if (I <= gc.nx && J <= gc.ny && K <= gc.nz) {
      float s = 0.f;
#pragma unroll
      for (int d = 0; d < 8; d++)
        s += rt[(2 * u + (d & 1)) + T * ((2 * v + ((d >> 1) & 1)) + T * (2 * w + (d >> 2)))];
      const int64_t c = gc.idx(I, J, K);
      bc[c] = s;
      xc[c] = 0.f;
    }
  }
}

// x_out = post-smoothing (half sweeps black, red, ...) of x + x_coarse(parent).
// x_out must not alias x: neighbouring work-groups read x over their halos.
template <int T, int NH>
__global__ __launch_bounds__(T *T *T) void k_mg_post_fused(Grid g, Grid gc, const float *__restrict__ cx,
                                                          const float *__restrict__ cy, const float *__restrict__ cz,
                                                          const float *__restrict__ dg, const float *__restrict__ b,
                                                          const float *__restrict__ x, const float *__restrict__ xc,
                                                          float *__restrict__ xo, const int *active) {
  if (!*active) return;
  constexpr int E = T + 2 * NH, N = E * E * E;
  __shared__ float sx[N], sb[N], sd[N], fx[N], fy[N], fz[N];
  const int i0 = 1 + blockIdx.x * T - NH, j0 = 1 + blockIdx.y * T - NH, k0 = 1 + blockIdx.z * T - NH;
  mg_tile_load<E, T * T * T>(g, gc, i0, j0, k0, cx, cy, cz, dg, b, x, xc, sx, sb, sd, fx, fy, fz);
  // half sweep h updates the box tile + (NH - 1 - h): after NH of them x is final on the tile
#pragma unroll
  for (int h = 0; h < NH; h++)
    mg_tile_gs<E>(g, i0, j0, k0, h + 1, E - 2 - 2 * h, (h & 1) ^ 1, false, sx, sb, sd, fx, fy, fz);
  const int t = threadIdx.x;
  const int a = NH + t % T, bb = NH + (t / T) % T, cc = NH + t / (T * T);
  const int i = i0 + a, j = j0 + bb, k = k0 + cc;
  if (i <= g.nx && j <= g.ny && k <= g.nz) xo[g.idx(i, j, k)] = sx[a + E * (bb + E * cc)];
}

// V-cycle over the small levels s..L-1 (each at most SULI_MG_SMALL cells) in
// one work-group, levels separated by barriers: the coarse end of the
// hierarchy costs one launch instead of a dozen latency-bound ones.  When the
// levels' frames fit (SULI_MG_SMALL_LDS floats for the six arrays of every
// level) they are copied to LDS first and only level s's x goes back.
#define SULI_MG_MAXLV 12
#ifndef SULI_MG_SMALL
#define SULI_MG_SMALL 1024
#endif
#define SULI_MG_SMALL_LDS 32768  // 128 KiB
#ifndef SULI_MG_SMALL_THREADS
#define SULI_MG_SMALL_THREADS 1024
#endif
#ifndef SULI_MG_MIN_CELLS
#define SULI_MG_MIN_CELLS 512  // no level is made below one of at most this many cells
#endif
struct MgSmall {
  Grid g[SULI_MG_MAXLV];
  float *cx[SULI_MG_MAXLV], *cy[SULI_MG_MAXLV], *cz[SULI_MG_MAXLV], *dg[SULI_MG_MAXLV];
  float *b[SULI_MG_MAXLV], *x[SULI_MG_MAXLV];
  int off[SULI_MG_MAXLV];  // LDS offset of level l's six frames (cx, cy, cz, dg, b, x)
  int s, L, nu, coarse_sweeps;
};


struct SmallLv {
  const float *cx, *cy, *cz, *dg;
  float *b, *x;
};

__device__ __forceinline__ void cell_of(const Grid &g, int t, int &i, int &j, int &k) {
  i = 1 + t % g.nx;
  j = 1 + (t / g.nx) % g.ny;
  k = 1 + t / (g.nx * g.ny);
}


// A thread's first cell pair on a small level: row position ti (cells 2ti+1
// and 2ti+2) in row (j, k), computed once per level; further pairs (levels of
// more than 2 * blockDim cells) are decomposed as they come.
struct SmallCell {
  int ti, j, k;
};
__device__ __forceinline__ SmallCell small_cell(const Grid &g) {
  const int hx = (g.nx + 1) >> 1, t = threadIdx.x;
  return SmallCell{t % hx, 1 + (t / hx) % g.ny, 1 + t / (hx * g.ny)};
}

// red-black half sweep of a small level (the cell of colour `color` in each
// pair, placed as in k_mg_gs)
__device__ __forceinline__ void small_gs(const Grid &g, const SmallLv &v, const SmallCell &sc, int color,
                                         bool zero_x) {
  const int hx = (g.nx + 1) >> 1, tot = hx * g.ny * g.nz;
  for (int t = threadIdx.x; t < tot; t += blockDim.x) {
    int ti = sc.ti, j = sc.j, k = sc.k;
    if (t != (int)threadIdx.x) {
      ti = t % hx;
      j = 1 + (t / hx) % g.ny;
      k = 1 + t / (hx * g.ny);
    }
    const int i = 1 + 2 * ti + ((1 + j + k + color) & 1);
    if (i > g.nx) continue;
    const int64_t c = g.idx(i, j, k);
    const float s = zero_x ? 0.f : mg_Ax(g, c, v.cx, v.cy, v.cz, v.x);
    v.x[c] = (v.b[c] + s) / v.dg[c];
  }
  __syncthreads();
}

// A cell of a small level held in registers for its sweeps: its frame
// offset, six face conductances, the reciprocal of its diagonal and its
// right-hand side, which do not change while the level is smoothed; a half
// sweep then reads only the neighbours' x (in k_mg_gs's order of terms).
struct RegCell {
  uint32_t c;
  float f[6], inv, b;
  bool ok;
};
__device__ __forceinline__ RegCell reg_cell(const Grid &g, const SmallLv &v, int t, int color) {
  RegCell r;
  const int hx = (g.nx + 1) >> 1, tot = hx * g.ny * g.nz;
  const int ti = t % hx, j = 1 + (t / hx) % g.ny, k = 1 + t / (hx * g.ny);
  const int i = 1 + 2 * ti + ((1 + j + k + color) & 1);
  r.ok = t < tot && i <= g.nx;
  r.c = r.ok ? g.idx(i, j, k) : g.idx(1, 1, 1);
  const uint32_t c = r.c;
  r.f[0] = v.cx[c + 1];
  r.f[1] = v.cx[c];
  r.f[2] = v.cy[c + g.sx];
  r.f[3] = v.cy[c];
  r.f[4] = v.cz[c + g.sz];
  r.f[5] = v.cz[c];
  r.inv = 1.f / v.dg[c];
  r.b = v.b[c];
  return r;
}
__device__ __forceinline__ void reg_update(const Grid &g, float *x, const RegCell &r, bool zero_x) {
  if (!r.ok) return;
  const uint32_t c = r.c;
  const float s = zero_x ? 0.f
                         : r.f[0] * x[c + 1] + r.f[1] * x[c - 1] + r.f[2] * x[c + g.sx] + r.f[3] * x[c - g.sx] +
                               r.f[4] * x[c + g.sz] + r.f[5] * x[c - g.sz];
  x[c] = (r.b + s) * r.inv;
}
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// nu red-black sweeps of a small level in LDS with register-held cells:
// one pair of cells per thread (the level has at most blockDim pairs)
__device__ __forceinline__ void small_sweeps_reg(const Grid &g, const SmallLv &v, int nu, bool from_zero,
                                                 bool black_first) {
  const RegCell r0 = reg_cell(g, v, threadIdx.x, 0), r1 = reg_cell(g, v, threadIdx.x, 1);
  for (int sw = 0; sw < nu; sw++) {
    if (!black_first) {
      reg_update(g, v.x, r0, from_zero && sw == 0);
      __syncthreads();
      reg_update(g, v.x, r1, false);
    } else {
      reg_update(g, v.x, r1, false);
      __syncthreads();
      reg_update(g, v.x, r0, false);
    }
    __syncthreads();
  }
}

template <bool LDS>
__device__ __forceinline__ SmallLv small_lv(const MgSmall &m, int l, float *smem) {
  if (LDS) {
    float *p = smem + m.off[l];
    const int64_t n = m.g[l].total;
    return SmallLv{p, p + n, p + 2 * n, p + 3 * n, p + 4 * n, p + 5 * n};
  }
  return SmallLv{m.cx[l], m.cy[l], m.cz[l], m.dg[l], m.b[l], m.x[l]};
}

template <bool LDS>
__global__ __launch_bounds__(1024) void k_mg_small(MgSmall m, const int *active) {
  if (!*active) return;
  __shared__ float smem[LDS ? SULI_MG_SMALL_LDS : 1];
  if (LDS) {
    // whole frames: ghost faces are 0 and ghost x must read as 0
    for (int l = m.s; l < m.L; l++) {
      const int n = (int)m.g[l].total;
      float *p = smem + m.off[l];
      for (int t = threadIdx.x; t < n; t += blockDim.x) {
        p[t] = m.cx[l][t];
        p[n + t] = m.cy[l][t];
        p[2 * n + t] = m.cz[l][t];
        p[3 * n + t] = m.dg[l][t];
        p[4 * n + t] = l == m.s ? m.b[l][t] : 0.f;
        p[5 * n + t] = 0.f;
      }
    }
    __syncthreads();
  }
  for (int l = m.s; l < m.L - 1; l++) {
    const SmallLv v = small_lv<LDS>(m, l, smem), vc = small_lv<LDS>(m, l + 1, smem);
    if (LDS && ((m.g[l].nx + 1) >> 1) * m.g[l].ny * m.g[l].nz <= (int)blockDim.x) {
      small_sweeps_reg(m.g[l], v, m.nu, true, false);
    } else {
      const SmallCell sc = small_cell(m.g[l]);
      for (int sw = 0; sw < m.nu; sw++) {
        small_gs(m.g[l], v, sc, 0, sw == 0);
        small_gs(m.g[l], v, sc, 1, false);
      }
    }
    const Grid &gf = m.g[l], &gc = m.g[l + 1];
    const int nc = gc.nx * gc.ny * gc.nz;
    for (int t = threadIdx.x; t < nc; t += blockDim.x) {
      int I, J, K;
      cell_of(gc, t, I, J, K);
      float s = 0.f;
      for (int d = 0; d < 8; d++) {
        const int64_t c = gf.idx(2 * I - 1 + (d & 1), 2 * J - 1 + ((d >> 1) & 1), 2 * K - 1 + (d >> 2));
        s += v.b[c] - (v.dg[c] * v.x[c] - mg_Ax(gf, c, v.cx, v.cy, v.cz, v.x));
      }
      vc.b[gc.idx(I, J, K)] = s;
    }
    __syncthreads();
  }
  {
    const int lc = m.L - 1;
    const SmallLv v = small_lv<LDS>(m, lc, smem);
    const Grid &g = m.g[lc];
    if (LDS && ((g.nx + 1) >> 1) * g.ny * g.nz <= 128) {
      // the coarsest level's sweeps on one wavefront, up to two pairs of
      // cells per lane held in registers: within a wavefront the half sweeps
      // need no work-group barrier (the lanes' LDS accesses are ordered by
      // wavefront-scope fences)
      if (threadIdx.x < 64) {
        RegCell r[2][2];
#pragma unroll
        for (int q = 0; q < 2; q++) {
          r[q][0] = reg_cell(g, v, threadIdx.x + 64 * q, 0);
          r[q][1] = reg_cell(g, v, threadIdx.x + 64 * q, 1);
        }
        for (int sw = 0; sw < m.coarse_sweeps; sw++)
#pragma unroll
          for (int color = 0; color < 2; color++) {
#pragma unroll
            for (int q = 0; q < 2; q++) reg_update(g, v.x, r[q][color], color == 0 && sw == 0);
            wave_lds_sync();
          }
      }
      __syncthreads();
    } else {
      const SmallCell sc = small_cell(g);
      for (int sw = 0; sw < m.coarse_sweeps; sw++) {
        small_gs(g, v, sc, 0, sw == 0);
        small_gs(g, v, sc, 1, false);
      }
    }
  }
  for (int l = m.L - 2; l >= m.s; l--) {
    const SmallLv v = small_lv<LDS>(m, l, smem), vc = small_lv<LDS>(m, l + 1, smem);
    const Grid &gf = m.g[l], &gc = m.g[l + 1];
    const int nf = gf.nx * gf.ny * gf.nz;
    for (int t = threadIdx.x; t < nf; t += blockDim.x) {
      int i, j, k;
      cell_of(gf, t, i, j, k);
      v.x[gf.idx(i, j, k)] += vc.x[gc.idx((i + 1) >> 1, (j + 1) >> 1, (k + 1) >> 1)];
    }
    __syncthreads();
    if (LDS && ((gf.nx + 1) >> 1) * gf.ny * gf.nz <= (int)blockDim.x) {
      small_sweeps_reg(gf, v, m.nu, false, true);
    } else {
      const SmallCell sc = small_cell(gf);
      for (int sw = 0; sw < m.nu; sw++) {
        small_gs(gf, v, sc, 1, false);
        small_gs(gf, v, sc, 0, false);
      }
    }
  }
  if (LDS) {
    const SmallLv v = small_lv<LDS>(m, m.s, smem);
    const Grid &g = m.g[m.s];
    const int n = g.nx * g.ny * g.nz;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
      int i, j, k;
      cell_of(g, t, i, j, k);
      const int64_t c = g.idx(i, j, k);
      m.x[m.s][c] = v.x[c];
    }
  }
}

// level-0 input b0 = sqrt(D) r and output z = sqrt(D) x0
__global__ __launch_bounds__(256) void k_scale(Grid g, const float *__restrict__ s, const float *__restrict__ a,
                                               float *__restrict__ o, const int *active) {
  if (!*active) return;
  CELL_IJK;
  if (i > g.nx || j > g.ny) return;
  for (int k = kz0; k <= g.nz; k += gridDim.z) {
    const int64_t c = g.idx(i, j, k);
    o[c] = s[c] * a[c];
  }
}

// final pressure p = x/sqrt(D) with mirrored ghosts (9_grad.f90:268-274);
// with dl, also the change of p in FP32, dl = p_new - p_old.  With cond, only
// when *cond is set (the residual just checked ends the solve), so that a
// solve writes p once.
__global__ void k_pout(Grid g, const double *__restrict__ x, const double *__restrict__ sD, double *__restrict__ p,
                       float *__restrict__ dl, const int *cond) {
  if (cond && !*cond) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y,
            k = blockIdx.z;
  if (i > g.nx + 1 || j > g.ny + 1 || k > g.nz + 1) return;
  auto m = [](int q, int n, bool per) { return q < 1 ? (per ? n : 1) : (q > n ? (per ? 1 : n) : q); };
  const int64_t s = g.idx(m(i, g.nx, g.ccx0 == 0), m(j, g.ny, g.ccy0 == 0), m(k, g.nz, false));
  const int64_t e = g.idx(i, j, k);
  const double pn = x[s] / sD[s];
  if (dl) dl[e] = (float)(pn - p[e]);
  p[e] = pn;
}

// initial guess p + sum_m w_m o_m from the pressure and n of its FP32
// offsets (n = 0: the pressure itself)
struct GuessOff {
  const float *o[3];
  double w[3];
  int n;
};

// initial iterate x = p_guess sqrt(D) and b_s = q qs/sqrt(D) with the block
// sums of b_s.b_s; q = -div u* formed here with k_pD's expression
template <typename T>
__global__ __launch_bounds__(256) void k_xbq(Grid g, const double *__restrict__ p, GuessOff go,
                                             const T *__restrict__ u, const T *__restrict__ v,
                                             const T *__restrict__ w, const double *__restrict__ sD, double qs,
                                             double *__restrict__ x, double *__restrict__ b, RedSlot rs) {
  CELL_IJK;
  double acc[1] = {0.0};
  if (i <= g.nx && j <= g.ny)
    for (int k = kz0; k <= g.nz; k += gridDim.z) {
      const fidx c = g.idx(i, j, k);
      const double sd = sD[c];
      double pg = p[c];
#pragma unroll
      for (int m = 0; m < 3; m++)
        if (m < go.n) pg += go.w[m] * (double)go.o[m][c];
      x[c] = pg * sd;
      const double bv = div_q<T>(g, u, v, w, c) * qs / sd;
      b[c] = bv;
      acc[0] += bv * bv;
    }
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1>(acc, lds);
  block_partial<1>(acc, rs.partials);
}

// ---------------------------------------------------------------------------
// host side

struct MgLevel {
  Grid g;
  float *cx = nullptr, *cy = nullptr, *cz = nullptr, *dg = nullptr, *b = nullptr, *x = nullptr;
  uint16_t *hx = nullptr, *hy = nullptr, *hz = nullptr;  // BF16 faces (finest level)
  float *xo = nullptr;  // post-smoothed x of a fused level (the fused kernel cannot update x in place)
};

struct PoissonFCG {
  FcgCtl *ctl = nullptr, *ctl_h = nullptr;
  FcgCtl *ctl_hd = nullptr;  // ctl_h as the device addresses it (the finish kernels publish into it)
  unsigned pub_seq = 0;      // publishes so far
  int64_t pub_hits = 0;      // control-block reads that found their publish
  int64_t pub_misses = 0;    // reads that found no fresh publish (and copied instead)
  double *b = nullptr, *x = nullptr, *r = nullptr;          // FP64 outer vectors (frame)
  float *r32 = nullptr, *d32 = nullptr, *p32 = nullptr, *p32b = nullptr, *q32 = nullptr, *z32 = nullptr;
  float *ax32 = nullptr, *ay32 = nullptr, *az32 = nullptr, *sD32 = nullptr;
  std::vector<MgLevel> lv;  // lv[0] shares the main frame
  // finest level split by colour (fused kernels): pair frame, x and b per colour, red diagonal
  Grid hg{};
  float *x0c = nullptr, *x1c = nullptr, *b0c = nullptr, *b1c = nullptr, *d0c = nullptr;
  // the same in FP16 (arrays of their own, so each form's ghost pairs stay zero)
  _Float16 *x0h = nullptr, *x1h = nullptr, *b0h = nullptr, *b1h = nullptr;
  bool setup_done = false;  // operator copies and hierarchy built by a solve
  // the density epoch (Ctx::rho_epoch) and the dt they were built for
  uint64_t op_epoch = 0;
  double op_dt = 0.0;
  int kz = 1;               // grid.z of the k-looped reduction kernels
  UNetPrecond *unet = nullptr;
  float *pm[SULI_FCG_M] = {}, *qm[SULI_FCG_M] = {};  // FCG(m) directions and their images
  // inner iterations of the passes of the last solve of each RK stage (0:
  // solves outside a stage): the two stages' solves differ
  int pass_iters[3][4] = {};
  int cur_iter = 0;  // inner iteration of the preconditioner application being enqueued
};

static int falloc(Ctx &c, float **p, int64_t n) {
  HIPCHK(hipMalloc((void **)p, (size_t)n * sizeof(float)));
  HIPCHK(hipMemsetAsync(*p, 0, (size_t)n * sizeof(float), c.st));
  c.allocs.push_back(*p);
  return 0;
}
static int dalloc64(Ctx &c, double **p, int64_t n) {
  HIPCHK(hipMalloc((void **)p, (size_t)n * sizeof(double)));
  HIPCHK(hipMemsetAsync(*p, 0, (size_t)n * sizeof(double), c.st));
  c.allocs.push_back(*p);
  return 0;
}

int fcg_create(Ctx &c) {
  PoissonFCG *f = new PoissonFCG();
  c.fcg = f;
  const Grid &g = c.g;
  const int64_t n = g.total;
  HIPCHK(hipMalloc((void **)&f->ctl, sizeof(FcgCtl)));
  c.allocs.push_back(f->ctl);
  HIPCHK(hipMemsetAsync(f->ctl, 0, sizeof(FcgCtl), c.st));
  HIPCHK(hipHostMalloc((void **)&f->ctl_h, sizeof(FcgCtl), hipHostMallocCoherent | hipHostMallocMapped));
  HIPCHK(hipHostGetDevicePointer((void **)&f->ctl_hd, f->ctl_h, 0));
  if (dalloc64(c, &f->b, n) || dalloc64(c, &f->x, n) || dalloc64(c, &f->r, n)) return 1;
  float **fl[] = {&f->r32, &f->d32, &f->p32, &f->p32b, &f->q32, &f->z32, &f->ax32, &f->ay32, &f->az32, &f->sD32};
  for (float **q : fl)
    if (falloc(c, q, n)) return 1;
  // k-loop depth of the reduction kernels: about 2048 blocks in total
  const int64_t plane_blocks = (int64_t)((g.nx + SULI_BX - 1) / SULI_BX) * ((g.ny + SULI_BY - 1) / SULI_BY);
  f->kz = (int)std::max<int64_t>(1, std::min<int64_t>(g.nz, 2048 / std::max<int64_t>(plane_blocks, 1)));
  // multigrid hierarchy
  MgLevel l0;
  l0.g = g;
  if (falloc(c, &l0.cx, n) || falloc(c, &l0.cy, n) || falloc(c, &l0.cz, n) || falloc(c, &l0.dg, n) ||
      falloc(c, &l0.b, n) || falloc(c, &l0.x, n))
    return 1;
  for (uint16_t **h : {&l0.hx, &l0.hy, &l0.hz}) {
    HIPCHK(hipMalloc((void **)h, (size_t)n * sizeof(uint16_t)));
    HIPCHK(hipMemsetAsync(*h, 0, (size_t)n * sizeof(uint16_t), c.st));
    c.allocs.push_back(*h);
  }
  f->lv.push_back(l0);
  {
    Grid h = g;
    h.nx = (g.nx + 1) / 2;
    h.sx = ((h.nx + 2 + 31) / 32) * 32;
    h.sy = g.ny + 2;
    h.sz = (fidx)h.sx * h.sy;
    h.off = 32;
    h.total = (int64_t)h.off + (int64_t)h.sz * (g.nz + 2) + 32;
    f->hg = h;
    if (falloc(c, &f->x0c, h.total) || falloc(c, &f->x1c, h.total) || falloc(c, &f->b0c, h.total) ||
        falloc(c, &f->b1c, h.total) || falloc(c, &f->d0c, h.total))
      return 1;
    for (_Float16 **a : {&f->x0h, &f->x1h, &f->b0h, &f->b1h}) {
      HIPCHK(hipMalloc((void **)a, (size_t)h.total * sizeof(_Float16)));
      HIPCHK(hipMemsetAsync(*a, 0, (size_t)h.total * sizeof(_Float16), c.st));
      c.allocs.push_back(*a);
    }
  }
  for (;;) {
    const Grid &gf = f->lv.back().g;
    if (gf.nx % 2 || gf.ny % 2 || gf.nz % 2 || gf.nx < 4 || gf.ny < 4 || gf.nz < 4) break;
    if ((int64_t)gf.nx * gf.ny * gf.nz <= SULI_MG_MIN_CELLS || f->lv.size() >= SULI_MG_MAXLV) break;
    MgLevel l;
    Grid gc = gf;
    gc.nx = gf.nx / 2; gc.ny = gf.ny / 2; gc.nz = gf.nz / 2;
    gc.sx = gc.nx + 3; gc.sy = gc.ny + 3; gc.sz = (fidx)gc.sx * gc.sy; gc.off = 0;
    gc.total = (int64_t)gc.sz * (gc.nz + 3) + 8;
    gc.dx = gf.dx * 2; gc.dy = gf.dy * 2; gc.dz = gf.dz * 2;
    l.g = gc;
    if (falloc(c, &l.cx, gc.total) || falloc(c, &l.cy, gc.total) || falloc(c, &l.cz, gc.total) ||
        falloc(c, &l.dg, gc.total) || falloc(c, &l.b, gc.total) || falloc(c, &l.x, gc.total) ||
        falloc(c, &l.xo, gc.total))
      return 1;
    f->lv.push_back(l);
  }
  if (unet_create(c, &f->unet)) return 1;
  return 0;
}

UNetPrecond *fcg_unet(Ctx &c) { return c.fcg ? c.fcg->unet : nullptr; }

void fcg_destroy(Ctx &c) {
  if (!c.fcg) return;
  unet_destroy(c, c.fcg->unet);
  if (c.fcg->ctl_h) (void)hipHostFree(c.fcg->ctl_h);
  delete c.fcg;
  c.fcg = nullptr;
}

static dim3 cells_kloop(const Grid &g, int kz) {
  return dim3((g.nx + SULI_BX - 1) / SULI_BX, (g.ny + SULI_BY - 1) / SULI_BY, kz);
}
static dim3 cells_full(const Grid &g) {
  return dim3((g.nx + SULI_BX - 1) / SULI_BX, (g.ny + SULI_BY - 1) / SULI_BY, g.nz);
}
static dim3 faces_full(const Grid &g) {
  return dim3((g.nx + 1 + SULI_BX - 1) / SULI_BX, (g.ny + 1 + SULI_BY - 1) / SULI_BY, g.nz + 1);
}
static dim3 half_cells(const Grid &g) {
  return dim3(((g.nx + 1) / 2 + SULI_BX - 1) / SULI_BX, (g.ny + SULI_BY - 1) / SULI_BY, g.nz);
}
// two cells of a colour per thread (pairs of the colour-split frame in twos)
static dim3 half2_cells(const Grid &g) {
  const int threads = ((g.nx + 1) / 2 + 1) / 2;
  return dim3((threads + SULI_BX - 1) / SULI_BX, (g.ny + SULI_BY - 1) / SULI_BY, g.nz);
}

// Build the FP32 operator copies and the multigrid hierarchy for this dt/rho.
// (level 0 and the FP32 operator copies here unless the fused set-up made them)
static int mg_setup(Ctx &c, bool level0) {
  PoissonFCG &f = *c.fcg;
  const dim3 b2(SULI_BX, SULI_BY, 1);
  {
    ProfScope ps(c, "mg_setup");
    MgLevel &l0 = f.lv[0];
    if (level0) {
      hipLaunchKernelGGL(k_mg_c0, faces_full(c.g), b2, 0, c.st, c.g, c.ax, c.ay, c.az, c.D, l0.cx, l0.cy, l0.cz,
                         l0.dg, f.ax32, f.ay32, f.az32, c.sD, f.sD32);
      hipLaunchKernelGGL(k_mg_bf16, faces_full(c.g), b2, 0, c.st, c.g, l0.cx, l0.cy, l0.cz, l0.hx, l0.hy, l0.hz);
      hipLaunchKernelGGL(k_mg_diag_bf16, cells_full(c.g), b2, 0, c.st, c.g, l0.hx, l0.hy, l0.hz, l0.dg,
                         (const double *)nullptr, (float *)nullptr, f.hg, f.d0c);
    }
    for (size_t q = 1; q < f.lv.size(); q++) {
      MgLevel &lf = f.lv[q - 1], &lc = f.lv[q];
      // the finest level coarsens from its BF16 faces (the operator its sweeps use)
      // (the diagonal of the coarse level with its faces, one launch)
      if (lf.hx)
        hipLaunchKernelGGL(k_mg_coarsen<uint16_t>, faces_full(lc.g), b2, 0, c.st, lf.g, lc.g,
                           (const uint16_t *)lf.hx, (const uint16_t *)lf.hy, (const uint16_t *)lf.hz, lc.cx, lc.cy,
                           lc.cz, lc.dg);
      else
        hipLaunchKernelGGL(k_mg_coarsen<float>, faces_full(lc.g), b2, 0, c.st, lf.g, lc.g, (const float *)lf.cx,
                           (const float *)lf.cy, (const float *)lf.cz, lc.cx, lc.cy, lc.cz, lc.dg);
    }
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// Whether the V-cycle of inner iteration `iter` of a pass holds the finest
// level's iterate and right-hand side in FP16 (when its input is staged by
// the kernel that produced the residual, which also scales it).  The
// rounding makes the preconditioner slightly nonlinear, which FCG(1)
// tolerates only down to a residual reduction of about the rounding: with
// BF16 (2^-8) the benchmark's cold starts took 14 -> 16 iterations.
#ifndef SULI_MG_F16_ITERS
#define SULI_MG_F16_ITERS 1000
#endif
static bool fine_f16(const Ctx &c, int iter) { return c.mg_fuse_fine == 2 && iter < SULI_MG_F16_ITERS; }

// z.q of the preconditioned FCG beta, computed by the V-cycle's last kernel
struct ZqOut {
  const float *q;
  const int *first;  // no sum on the first iteration of a pass
  RedSlot rs;
};

// One multigrid V-cycle: z = sqrt(D) V(sqrt(D) r), and z.q when zq is set.
// Levels with more than SULI_MG_SMALL cells run one launch per half sweep
// (levels of at most c.mg_fuse_max cells below the finest: one launch per
// side), the rest run in k_mg_small.  With c.mg_fuse_fine the input scaling,
// the prolongations and the output scaling with z.q ride on the neighbouring
// half sweeps.  With net_level = e (1 <= e < number of levels) the levels
// below e are not visited: the network solves level e's problem
// (unet_coarse_solve) in place of the coarse cycle.  While the network's
// training capture is armed at entry level e, level e's problem is appended
// to the set after its restriction (multigrid and learned solves alike; the
// levels above e then run their own launches even where they would fit the
// single-work-group kernel).
static void mg_vcycle(Ctx &c, const float *r, float *z, const int *active, const ZqOut *zq = nullptr,
                      int net_level = 0, bool staged = false, bool f16 = false) {
  PoissonFCG &f = *c.fcg;
  const dim3 b2(SULI_BX, SULI_BY, 1);
  const int nu1 = SULI_MG_NU, nu2 = SULI_MG_NU;
  const int L = (int)f.lv.size();
  const auto ncell = [&](int l) { return (int64_t)f.lv[l].g.nx * f.lv[l].g.ny * f.lv[l].g.nz; };
  int S = 0;
  while (S < L - 1 && ncell(S) > SULI_MG_SMALL) S++;
  const int cap_level = unet_capture_armed(f.unet) ? unet_entry(f.unet) : 0;
  if (net_level > 0) S = net_level;
  else if (cap_level > S) S = cap_level;  // level cap_level's right-hand side is built by launches
  const auto capture = [&](int l) {  // level l's right-hand side is complete
    if (cap_level > 0 && l == cap_level) {
      MgLevel &le = f.lv[l];
      unet_capture_level(c, f.unet, le.g, le.cx, le.cy, le.cz, le.dg, le.b, active, f.cur_iter);
    }
  };
  const bool ff = c.mg_fuse_fine && S >= 1;  // the finest level has launches of its own
  const bool xb = ff && staged && f16;  // its colour-split x and b in FP16, the input scaled
  // its half sweeps with two cells per thread (k_mg0_gs2; the frame alignment it assumes)
  const bool pair2 = ff && c.g.fast && (c.g.off + 1) % 4 == 0 && c.g.sx % 4 == 0 && f.hg.off % 2 == 0 &&
                     f.hg.sx % 2 == 0 && f.hg.sz % 2 == 0;
  // one red-black half sweep of level lf (BF16 faces where the level has them)
  auto gs = [&](MgLevel &lf, const char *name, int color, int zero_x) {
    ProfKernel pk(c, name);
    if (ff && &lf == &f.lv[0] && pair2) {  // two cells per thread
      if (xb)
        SULI_TIMED_LAUNCH(pk, k_mg0_gs2<_Float16>, half2_cells(lf.g), b2, 0, c.st, lf.g, f.hg,
                          (const uint16_t *)lf.hx, (const uint16_t *)lf.hy, (const uint16_t *)lf.hz,
                          (const _Float16 *)(color ? f.b1h : f.b0h), (const _Float16 *)(color ? f.x0h : f.x1h),
                          color ? f.x1h : f.x0h, color, active);
      else
        SULI_TIMED_LAUNCH(pk, k_mg0_gs2<float>, half2_cells(lf.g), b2, 0, c.st, lf.g, f.hg, (const uint16_t *)lf.hx,
                          (const uint16_t *)lf.hy, (const uint16_t *)lf.hz, (const float *)(color ? f.b1c : f.b0c),
                          (const float *)(color ? f.x0c : f.x1c), color ? f.x1c : f.x0c, color, active);
      return;
    }
    if (ff && &lf == &f.lv[0]) {  // finest level, colour-split arrays (never from x = 0 here)
      if (xb)
        SULI_TIMED_LAUNCH(pk, k_mg0_gs<_Float16>, half_cells(lf.g), b2, 0, c.st, lf.g, f.hg, (const uint16_t *)lf.hx,
                          (const uint16_t *)lf.hy, (const uint16_t *)lf.hz, (const _Float16 *)(color ? f.b1h : f.b0h),
                          (const _Float16 *)(color ? f.x0h : f.x1h), color ? f.x1h : f.x0h, color, active);
      else
        SULI_TIMED_LAUNCH(pk, k_mg0_gs<float>, half_cells(lf.g), b2, 0, c.st, lf.g, f.hg, (const uint16_t *)lf.hx,
                          (const uint16_t *)lf.hy, (const uint16_t *)lf.hz, (const float *)(color ? f.b1c : f.b0c),
                          (const float *)(color ? f.x0c : f.x1c), color ? f.x1c : f.x0c, color, active);
      return;
    }
    if (lf.hx)
      SULI_TIMED_LAUNCH(pk, k_mg_gs<uint16_t>, half_cells(lf.g), b2, 0, c.st, lf.g, (const uint16_t *)lf.hx,
                        (const uint16_t *)lf.hy, (const uint16_t *)lf.hz, (const float *)lf.b, lf.x, color,
                        zero_x, active);
    else
      SULI_TIMED_LAUNCH(pk, k_mg_gs<float>, half_cells(lf.g), b2, 0, c.st, lf.g, (const float *)lf.cx,
                        (const float *)lf.cy, (const float *)lf.cz, (const float *)lf.b, lf.x, color, zero_x,
                        active);
  };
  if (ff && staged) {
    // the finest level's input was staged from r by the kernel that produced r
  } else if (ff) {
    ProfScope ps(c, "mg_gs_l0_in");
    hipLaunchKernelGGL(k_mg0_in<float>, half_cells(c.g), b2, 0, c.st, c.g, f.hg, (const float *)f.d0c,
                       (const float *)f.sD32, r, f.b0c, f.b1c, f.x0c, active);
  } else {
    ProfScope ps(c, "mg_scale");
    hipLaunchKernelGGL(k_scale, cells_kloop(c.g, f.kz), b2, 0, c.st, c.g, f.sD32, r, f.lv[0].b, active);
  }
  // levels 1..S-1 of at most c.mg_fuse_max cells smooth in one launch per side
  const auto fused = [&](int l) { return l >= 1 && l < S && ncell(l) <= c.mg_fuse_max; };
  const dim3 bt(SULI_MG_FT * SULI_MG_FT * SULI_MG_FT);
  const auto tiles = [](const Grid &g) {
    return dim3((g.nx + SULI_MG_FT - 1) / SULI_MG_FT, (g.ny + SULI_MG_FT - 1) / SULI_MG_FT,
                (g.nz + SULI_MG_FT - 1) / SULI_MG_FT);
  };
  // levels 1..S-1 with launches of their own and FP32 faces take their first
  // red half sweep from the restriction that fills them (coarse_red_x;
  // with the fast stencil forms, bit for bit the separate launch)
  const auto red_in_restrict = [&](int l) {
    return c.g.fast && l >= 1 && l < S && !fused(l) && !f.lv[l].hx && f.lv[l].cx;
  };
  for (int l = 0; l < S; l++) {
    MgLevel &lf = f.lv[l], &lc = f.lv[l + 1];
    const bool rz = red_in_restrict(l + 1);
    const float *ccx = rz ? (const float *)lc.cx : nullptr, *ccy = rz ? (const float *)lc.cy : nullptr,
                *ccz = rz ? (const float *)lc.cz : nullptr;
    if (fused(l)) {
      ProfScope ps(c, "mg_pre_fused");
      hipLaunchKernelGGL((k_mg_pre_fused<SULI_MG_FT, SULI_MG_FNH>), tiles(lf.g), bt, 0, c.st, lf.g, lc.g,
                         (const float *)lf.cx, (const float *)lf.cy, (const float *)lf.cz, (const float *)lf.dg,
                         (const float *)lf.b, lf.x, lc.b, lc.x, active);
      capture(l + 1);
      continue;
    }
    const char *gsn = l == 0 ? "mg_gs_l0" : "mg_gs";
    for (int s = 0; s < (l == 0 ? SULI_MG_NU0 : (l == 1 ? SULI_MG_NU1 : nu1)); s++) {
      if (!(s == 0 && ((ff && l == 0) || red_in_restrict(l)))) gs(lf, gsn, 0, s == 0 ? 1 : 0);
      gs(lf, gsn, 1, 0);
    }
    ProfScope ps(c, l == 0 ? "mg_restrict_l0" : "mg_restrict");
    if (ff && l == 0 && xb)
      hipLaunchKernelGGL(k_mg0_restrict<_Float16>, cells_full(lc.g), b2, 0, c.st, lf.g, f.hg, lc.g,
                         (const uint16_t *)lf.hx, (const uint16_t *)lf.hy, (const uint16_t *)lf.hz,
                         (const _Float16 *)f.b0h, (const _Float16 *)f.b1h, (const _Float16 *)f.x0h,
                         (const _Float16 *)f.x1h, lc.b, lc.x, ccx, ccy, ccz, active);
    else if (ff && l == 0)
      hipLaunchKernelGGL(k_mg0_restrict<float>, cells_full(lc.g), b2, 0, c.st, lf.g, f.hg, lc.g,
                         (const uint16_t *)lf.hx, (const uint16_t *)lf.hy, (const uint16_t *)lf.hz,
                         (const float *)f.b0c, (const float *)f.b1c, (const float *)f.x0c, (const float *)f.x1c,
                         lc.b, lc.x, ccx, ccy, ccz, active);
    else if (lf.hx)
      hipLaunchKernelGGL((k_mg_restrict<uint16_t>), cells_full(lc.g), b2, 0, c.st, lf.g, lc.g, (const uint16_t *)lf.hx,
                         (const uint16_t *)lf.hy, (const uint16_t *)lf.hz, (const float *)lf.b,
                         (const float *)lf.x, lc.b, lc.x, ccx, ccy, ccz, active);
    else
      hipLaunchKernelGGL((k_mg_restrict<float>), cells_full(lc.g), b2, 0, c.st, lf.g, lc.g, (const float *)lf.cx,
                         (const float *)lf.cy, (const float *)lf.cz, (const float *)lf.b, (const float *)lf.x,
                         lc.b, lc.x, ccx, ccy, ccz, active);
    capture(l + 1);
  }
  if (net_level > 0) {
    MgLevel &le = f.lv[S];
    unet_coarse_solve(c, f.unet, le.g, le.cx, le.cy, le.cz, le.dg, le.b, le.x, active);
  } else {
    MgSmall m;
    std::memset(&m, 0, sizeof(m));
    for (int l = S; l < L && l - S < SULI_MG_MAXLV; l++) {
      const int q = l;
      m.g[q] = f.lv[l].g;
      m.cx[q] = f.lv[l].cx; m.cy[q] = f.lv[l].cy; m.cz[q] = f.lv[l].cz; m.dg[q] = f.lv[l].dg;
      m.b[q] = f.lv[l].b; m.x[q] = f.lv[l].x;
    }
    m.s = S;
    m.L = L;
    m.nu = nu1;
    m.coarse_sweeps = c.mg_coarse_sweeps;
    int64_t off = 0;
    for (int l = S; l < L && l - S < SULI_MG_MAXLV; l++) {
      m.off[l] = (int)off;
      off += 6 * f.lv[l].g.total;
    }
    ProfScope ps(c, "mg_small_levels");
    if (off <= SULI_MG_SMALL_LDS)
      hipLaunchKernelGGL(k_mg_small<true>, dim3(1), dim3(SULI_MG_SMALL_THREADS), 0, c.st, m, active);
    else
      hipLaunchKernelGGL(k_mg_small<false>, dim3(1), dim3(SULI_MG_SMALL_THREADS), 0, c.st, m, active);
  }
  for (int l = S - 1; l >= 0; l--) {
    MgLevel &lf = f.lv[l], &lc = f.lv[l + 1];
    const float *xcoarse = fused(l + 1) ? lc.xo : lc.x;  // the next level's final x
    if (fused(l)) {
      ProfScope ps(c, "mg_post_fused");
      hipLaunchKernelGGL((k_mg_post_fused<SULI_MG_FT, SULI_MG_FNH>), tiles(lf.g), bt, 0, c.st, lf.g, lc.g,
                         (const float *)lf.cx, (const float *)lf.cy, (const float *)lf.cz, (const float *)lf.dg,
                         (const float *)lf.b, (const float *)lf.x, xcoarse, lf.xo, active);
      continue;
    }
    const char *gsn = l == 0 ? "mg_gs_l0" : "mg_gs";
    const int nu = l == 0 ? SULI_MG_NU0 : (l == 1 ? SULI_MG_NU1 : nu2);
    if (c.mg_fuse_fine) {
      ProfScope ps(c, l == 0 ? "mg_gs_l0_prolong" : "mg_gs_prolong");
      // (one cell per thread: a two-cell form of this sweep, with its eleven
      // coarse-parent loads, measured 12.8 against 11.5 us)
      if (l == 0 && xb)  // (c.mg_fuse_fine and S >= 1: ff)
        hipLaunchKernelGGL(k_mg0_prolong<_Float16>, half_cells(lf.g), b2, 0, c.st, lf.g, f.hg, lc.g,
                           (const uint16_t *)lf.hx, (const uint16_t *)lf.hy, (const uint16_t *)lf.hz,
                           (const _Float16 *)f.b1h, (const _Float16 *)f.x0h, f.x1h, xcoarse, active);
      else if (l == 0)
        hipLaunchKernelGGL(k_mg0_prolong<float>, half_cells(lf.g), b2, 0, c.st, lf.g, f.hg, lc.g,
                           (const uint16_t *)lf.hx, (const uint16_t *)lf.hy, (const uint16_t *)lf.hz,
                           (const float *)f.b1c, (const float *)f.x0c, f.x1c, xcoarse, active);
      else if (lf.hx)
        hipLaunchKernelGGL((k_mg_gs_prolong<uint16_t>), half_cells(lf.g), b2, 0, c.st, lf.g, lc.g,
                           (const uint16_t *)lf.hx, (const uint16_t *)lf.hy, (const uint16_t *)lf.hz,
                           (const float *)lf.b, lf.x, xcoarse, active);
      else
        hipLaunchKernelGGL((k_mg_gs_prolong<float>), half_cells(lf.g), b2, 0, c.st, lf.g, lc.g,
                           (const float *)lf.cx, (const float *)lf.cy, (const float *)lf.cz,
                           (const float *)lf.b, lf.x, xcoarse, active);
    } else {
      ProfScope ps(c, l == 0 ? "mg_prolong_l0" : "mg_prolong");
      hipLaunchKernelGGL(k_mg_prolong, cells_full(lf.g), b2, 0, c.st, lf.g, lc.g, xcoarse, lf.x, active);
    }
    for (int s = 0; s < nu; s++) {
      if (!(c.mg_fuse_fine && s == 0)) gs(lf, gsn, 1, 0);
      if (!(ff && l == 0 && s == nu - 1)) gs(lf, gsn, 0, 0);
    }
  }
  if (ff) {
    MgLevel &l0 = f.lv[0];
    ProfScope ps(c, "mg_gs_l0_out");
    double *part = zq ? zq->rs.partials : nullptr;
    const float *q = zq ? zq->q : nullptr;
    const int *first = zq ? zq->first : active;
    const float *sig = (const float *)((const char *)f.ctl + offsetof(FcgCtl, sig));
    const dim3 go = pair2 ? half2_cells(c.g) : half_cells(c.g);
    if (pair2 && xb)
      hipLaunchKernelGGL(k_mg0_out2<_Float16>, go, b2, 0, c.st, c.g, f.hg, (const uint16_t *)l0.hx,
                         (const uint16_t *)l0.hy, (const uint16_t *)l0.hz, (const _Float16 *)f.b0h,
                         (const _Float16 *)f.x1h, (const float *)f.sD32, z, q, first, part, active, sig);
    else if (pair2)
      hipLaunchKernelGGL(k_mg0_out2<float>, go, b2, 0, c.st, c.g, f.hg, (const uint16_t *)l0.hx,
                         (const uint16_t *)l0.hy, (const uint16_t *)l0.hz, (const float *)f.b0c,
                         (const float *)f.x1c, (const float *)f.sD32, z, q, first, part, active, nullptr);
    else if (xb)
      hipLaunchKernelGGL(k_mg0_out<_Float16>, half_cells(c.g), b2, 0, c.st, c.g, f.hg, (const uint16_t *)l0.hx,
                         (const uint16_t *)l0.hy, (const uint16_t *)l0.hz, (const _Float16 *)f.b0h,
                         (const _Float16 *)f.x1h, (const float *)f.sD32, z, q, first, part, active, sig);
    else
      hipLaunchKernelGGL(k_mg0_out<float>, half_cells(c.g), b2, 0, c.st, c.g, f.hg, (const uint16_t *)l0.hx,
                         (const uint16_t *)l0.hy, (const uint16_t *)l0.hz, (const float *)f.b0c,
                         (const float *)f.x1c, (const float *)f.sD32, z, q, first, part, active, nullptr);
    if (zq) launch_finish<1>(c.st, go, zq->rs, EpiActive{f.ctl});
    return;
  }
  {
    ProfScope ps(c, "mg_scale");
    hipLaunchKernelGGL(k_scale, cells_kloop(c.g, f.kz), b2, 0, c.st, c.g, f.sD32, f.lv[0].x, z, active);
  }
  if (zq) {
    const dim3 gk = cells_kloop(c.g, f.kz);
    ProfScope ps(c, "fcg_zq");
    hipLaunchKernelGGL(k_zq, gk, b2, 0, c.st, c.g, z, zq->q, f.ctl, zq->rs);
    launch_finish<1>(c.st, gk, zq->rs, EpiActive{f.ctl});
  }
}

// Apply the preconditioner selected for this solve: z = M r
// (not staged: the FP32 form)
static void apply_precond(Ctx &c, int pc, const float *r, float *z, const int *active) {
  if (pc == SULI_PRECOND_LEARNED) {
    const int e = unet_entry(c.fcg->unet);
    if (e > 0) mg_vcycle(c, r, z, active, nullptr, e);
    else unet_apply(c, c.fcg->unet, r, z, active);
    return;
  }
  mg_vcycle(c, r, z, active);
}

// inner iterations a learned-preconditioner pass may take before the solve
// falls back to multigrid
static constexpr int kLearnedPassCap = 64;

// One FCG(m) iteration with the learned preconditioner; the new direction
// goes to slot `slot` of the ring of held directions.
// Wait for the stream and read the control block that the last finish kernel
// published (publish number seq) into the host copy; a copy that is not that
// publish is read from the device instead.
static int ctl_read(Ctx &c, PoissonFCG &f, unsigned seq) {
  HIPCHK(hipStreamSynchronize(c.st));
  if (f.ctl_h->seq == seq) {
    f.pub_hits++;
  } else {
    f.pub_misses++;
    HIPCHK(hipMemcpyAsync(f.ctl_h, f.ctl, sizeof(FcgCtl), hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
  }
  return 0;
}

static void fcgm_iteration(Ctx &c, PoissonFCG &f, int iter, dim3 gk, const int *active, FcgCtl *pub = nullptr,
                           unsigned seq = 0) {
  const Grid &g = c.g;
  const dim3 b2(SULI_BX, SULI_BY, 1);
  const int slot = iter % SULI_FCG_M;
  apply_precond(c, SULI_PRECOND_LEARNED, f.r32, f.z32, active);
  {
    ProfScope ps(c, "fcgm_spmv_dots");
    const RedSlot rz{c.partials, CTL_PTR(f.ctl, zqm)};
    hipLaunchKernelGGL(k_fcgm_lz, gk, b2, 0, c.st, g, f.ax32, f.ay32, f.az32, f.z32, f.q32, f.qm[0], f.qm[1],
                       f.qm[2], f.qm[3], f.ctl, rz);
    launch_finish<SULI_FCG_M>(c.st, gk, rz, EpiActive{f.ctl});
  }
  {
    ProfScope ps(c, "fcgm_orthogonalise");
    const RedSlot rp{c.partials, CTL_PTR(f.ctl, k1)};
    hipLaunchKernelGGL(k_fcgm_lp, gk, b2, 0, c.st, g, f.z32, f.q32, f.r32, f.pm[0], f.pm[1], f.pm[2], f.pm[3],
                       f.qm[0], f.qm[1], f.qm[2], f.qm[3], f.pm[slot], f.qm[slot], f.ctl, rp);
    launch_finish<2>(c.st, gk, rp, EpiLP{f.ctl, slot});
  }
  {
    ProfScope ps(c, "fcg_update");
    const RedSlot r2{c.partials, CTL_PTR(f.ctl, k2)};
    hipLaunchKernelGGL(k_fcg2, gk, b2, 0, c.st, g, f.pm[slot], f.qm[slot], f.r32, f.d32, f.ctl, r2, StageIn{});
    launch_finish<2>(c.st, gk, r2, EpiK2{f.ctl, pub, seq});
  }
}

// relative residual of the solvable part from the read-back control block;
// bnorm receives ||P b_s||
static double proj_res(const FcgCtl *h, double &bnorm) {
  double rr, bb;
  proj_norms(h->rout, h->bb[0], rr, bb);
  bnorm = std::sqrt(bb);
  return bnorm > 0 ? std::sqrt(rr) / bnorm : 0.0;
}

int fcg_solve(Ctx &c, double dt, suli_poisson_stats *stt) {
  PoissonFCG &f = *c.fcg;
  const Grid &g = c.g;
  HIPCHK(hipEventRecord(c.solve_ev[0], c.st));
  const dim3 b2(SULI_BX, SULI_BY, 1);
  const dim3 gk = cells_kloop(g, f.kz);
  const double tol = c.P.tol_rel;
  int pc = c.P.precond;
  // Initial guess.  The reference warm-starts every projection from the
  // current pressure (9_grad.f90:121).  Here a projection of the first or
  // last stage of a step starts from the current pressure plus its stage's
  // offset extrapolated in time.  The first- and last-stage pressures of the
  // steps interleave (g(n), h(n)), and each solve starts from the other
  // kind's last pressure: a first-stage solve from h(n-1), a last-stage solve
  // from g(n).  The offsets d(n) = g(n) - h(n-1) (first stage) and
  // h(n) - g(n) (last stage) change slowly, and with the last two of the
  // same kind known the solve starts from p + 2 d(n-1) - d(n-2) (from
  // p + d(n-1) with one): for the first stage h(n-1) + 2 (g(n-1) - h(n-2)) -
  // (g(n-2) - h(n-3)).  The offsets are kept in FP32 (rounded relative to
  // the change they hold, not to p), written by the solve's output kernel,
  // and the guess is formed inside the b_s pass.  On the dam break the
  // offset forms took 6 and 4.75 FCG iterations at step 500 where a
  // quadratic extrapolation of the last-stage pressures took 7 and 5.9 and
  // the current pressure 8 and 7.8; suli_set_guess(0) restores the
  // reference's choice.
#ifndef SULI_GUESS_OFFSETS
#define SULI_GUESS_OFFSETS 2  // offsets extrapolated: 1 constant, 2 linear, 3 quadratic
#endif
  const double base_dt = c.stage == 1 ? 2.0 * dt : dt;  // a_dt(1) = dt/2, a_dt(2) = dt
  const int sk = c.stage == 1 ? 0 : (c.stage == 2 ? 1 : -1);  // kind: first, last stage
  if (sk >= 0 && base_dt != c.hdt) {
    c.offn[0] = c.offn[1] = 0;
    c.hdt = base_dt;
  }
  // this solve's offset joins the history when it starts from the other kind's pressure
  const bool chain = sk >= 0 && c.last == 2 - sk;
  if (sk >= 0 && !chain) c.offn[sk] = 0;
  GuessOff go{};
  if (c.guess == 1 && chain) {
    static const double wts[4][3] = {{0, 0, 0}, {1, 0, 0}, {2, -1, 0}, {3, -3, 1}};
    go.n = std::min(c.offn[sk], SULI_GUESS_OFFSETS);
    for (int m = 0; m < 3; m++) {
      go.o[m] = c.off[sk][m];
      go.w[m] = wts[go.n][m];
    }
  }
  // the output kernel writes the offset into the oldest slot, which then
  // leaves the history
  float *dslot = nullptr;
  if (chain) {
    if (!c.off[sk][2] && falloc(c, &c.off[sk][2], g.total)) return 1;
    dslot = c.off[sk][2];
    c.offn[sk] = std::min(c.offn[sk], 2);
  }
  c.last = 0;  // until this solve has written its pressure
  // Operator.  c_f = dt/(h^2 rho_f), D = sum of c_f and the scaled faces
  // a_f = c_f/sqrt(D_P D_Q) depend on dt only through a common factor, so
  // while the density is unchanged (the stages of one step) the operator
  // built for dt_s serves a projection at dt: A(dt) p = q is
  // A(dt_s) p = q dt_s/dt, and with dt = 2 dt_s (the two stages of RK2) the
  // scaling is exact.  The faces, D, sqrt(D) and the multigrid hierarchy are
  // then kept and only b_s is formed.
#ifndef SULI_FCG_REUSE_OPERATOR
#define SULI_FCG_REUSE_OPERATOR 1
#endif
  const bool reuse = SULI_FCG_REUSE_OPERATOR && f.setup_done && f.op_epoch == c.rho_epoch && f.op_dt > 0.0;
  {
    ProfScope ps(c, "poisson_setup");
    const RedSlot rb{c.partials, CTL_PTR(f.ctl, bb)};
    if (!reuse) {
      // FP64 set-up shared with the reference CG (q formed below)
      // (the fused face pass reads mirrored cells itself: no ghost fill)
      DISPATCH(c, Launch<TT>::pD(g, c.st, dt, c.u, c.v, c.w, c.rho, c.D, c.sD, nullptr, !c.mg_fuse_fine));
      if (c.mg_fuse_fine) {
        // the faces in one pass with the finest multigrid level's BF16 faces
        MgLevel &l0 = f.lv[0];
        const FcgFaceOut fo{c.ax, c.ay, c.az, f.ax32, f.ay32, f.az32, l0.hx, l0.hy, l0.hz};
        DISPATCH(c, Launch<TT>::fcg_faces(g, c.st, dt, c.rho, c.D, fo));
        hipLaunchKernelGGL(k_mg_diag_bf16, cells_full(c.g), b2, 0, c.st, c.g, l0.hx, l0.hy, l0.hz, l0.dg,
                           (const double *)c.sD, f.sD32, f.hg, f.d0c);
      } else {
        DISPATCH(c, Launch<TT>::pcoef(g, c.st, dt, c.rho, c.D, c.ax, c.ay, c.az));
      }
    }
    const double qs = reuse ? f.op_dt / dt : 1.0;
    DISPATCH(c, hipLaunchKernelGGL((k_xbq<TT>), gk, b2, 0, c.st, g, (const double *)c.prd1, go, (const TT *)c.u,
                                   (const TT *)c.v, (const TT *)c.w, (const double *)c.sD, qs, f.x, f.b, rb));
    launch_finish<1>(c.st, gk, rb);
  }
  if (!reuse) {
    f.setup_done = false;
    if (mg_setup(c, !c.mg_fuse_fine)) return 1;
    f.setup_done = true;
    f.op_epoch = c.rho_epoch;
    f.op_dt = dt;
  }
  const int entry = unet_entry(f.unet);
  if (entry > 0 && entry >= (int)f.lv.size() && (pc == SULI_PRECOND_LEARNED || unet_capture_armed(f.unet))) {
    suli_set_error("learned preconditioner: entry level " + std::to_string(entry) + " but the multigrid hierarchy has " +
                   std::to_string(f.lv.size()) + " levels");
    return 1;
  }
  if (pc == SULI_PRECOND_LEARNED) {
    if (unet_prepare(c, f.unet)) return 1;
    for (int m = 0; m < SULI_FCG_M; m++)
      if (!f.pm[m] && (falloc(c, &f.pm[m], g.total) || falloc(c, &f.qm[m], g.total))) return 1;
  }
  if (entry > 0 && (pc == SULI_PRECOND_LEARNED || unet_capture_armed(f.unet)) && unet_level_density(c, f.unet))
    return 1;
  // FCG(1) with the network as a multigrid level's coarse solver (the
  // multigrid iteration with that level's solve replaced), when selected
  bool learned_m1 = pc == SULI_PRECOND_LEARNED && entry > 0 && unet_fcg_m(f.unet) == 1;
  int outer = 0, inner_total = 0, fallbacks = 0;
  double res0 = -1, res = 0, bnorm = 0;
  bool conv = false;
  const int max_outer = 12;
#ifndef SULI_FCG_REL_IN
#define SULI_FCG_REL_IN 1e-4  // inner (FP32) reduction a refinement pass aims for at most
#endif
  const double rel_in = SULI_FCG_REL_IN;
  const int maxit_in = c.P.max_it > 0 ? c.P.max_it : 10000;
  int *active = (int *)((char *)f.ctl + offsetof(FcgCtl, active));
  // Host round trips: the residual of a pass is read back before the pass is
  // enqueued only when the pass might not be needed.  When the previous pass
  // (or, for the first pass, the previous solve) leaves the residual clearly
  // above the tolerance, the residual, the arming of the next pass and its
  // first iterations go out together; the device then decides (EpiR32) and
  // a pass that turns out unnecessary is skipped by its own kernels.  When the
  // solve is expected to end, the output pressure is computed with the
  // residual check so that the last read-back finds it done.
  int *pits = f.pass_iters[c.stage == 1 || c.stage == 2 ? c.stage : 0];
  bool speculate = pc != SULI_PRECOND_LEARNED && pits[0] > 0;
  const dim3 pg((g.nx + 2 + 63) / 64, (g.ny + 2 + 3) / 4, g.nz + 2);
  bool pout_done = false;
  for (;;) {
    const bool spec = speculate && outer < max_outer;
    const unsigned seq_out = spec ? 0u : ++f.pub_seq;
    {
      ProfScope ps(c, "fcg_residual64");
      const RedSlot rr{c.partials, CTL_PTR(f.ctl, rout)};
      hipLaunchKernelGGL(k_res64, gk, b2, 0, c.st, g, c.ax, c.ay, c.az, f.x, f.b, c.sD, f.r, rr);
      launch_finish<3>(c.st, gk, rr, EpiOut{f.ctl, tol, outer >= max_outer ? 1 : 0, spec ? nullptr : f.ctl_hd, seq_out});
    }
    if (!spec) {
      {
        ProfScope ps(c, "poisson_final");  // written when this residual ends the solve
        hipLaunchKernelGGL(k_pout, pg, dim3(64, 4, 1), 0, c.st, g, f.x, c.sD, c.prd1, dslot,
                           (const int *)((char *)f.ctl + offsetof(FcgCtl, out)));
      }
      if (ctl_read(c, f, seq_out)) return 1;
      res = proj_res(f.ctl_h, bnorm);
      if (res0 < 0) res0 = res;
      if (f.ctl_h->out) {  // the device's verdict, the test above and outer >= max_outer
        conv = !(res > tol);
        pout_done = true;
        break;
      }
    }
    outer++;
    const bool use_pc = pc != SULI_PRECOND_JACOBI;
    // the V-cycles of the pass take their finest level's input from the
    // kernel that wrote r: the first from the pass's projection, the others
    // from K2
    const bool stage = use_pc && (pc != SULI_PRECOND_LEARNED || learned_m1) && c.mg_fuse_fine && f.x0c && f.b0c &&
                       f.b1c && f.d0c;
    // (in the form of the V-cycle that reads it)
    const StageIn sti_f = stage ? StageIn{f.hg, f.sD32, f.d0c, f.b0c, f.b1c, f.x0c, 0} : StageIn{};
    const StageIn sti_h = stage ? StageIn{f.hg, f.sD32, f.d0c, f.b0h, f.b1h, f.x0h, 1} : StageIn{};
    const auto sti = [&](int it) { return fine_f16(c, it) ? sti_h : sti_f; };
    const int pass_max = pc == SULI_PRECOND_LEARNED ? std::min(kLearnedPassCap, maxit_in - inner_total)
                                                    : maxit_in - inner_total;
    {
      ProfScope ps(c, "fcg_project");
      const RedSlot r3{c.partials, CTL_PTR(f.ctl, rin0)};
      hipLaunchKernelGGL(k_r32, gk, b2, 0, c.st, g, f.r, c.sD, f.r32, f.d32, f.ctl, rel_in, pass_max, r3, sti(0));
      const bool capped = pc == SULI_PRECOND_LEARNED;
      launch_finish<1>(c.st, gk, r3,
                       EpiR32{f.ctl, rel_in, tol, pass_max,
                              capped ? SULI_FCG_ONEPASS_WANT_CAPPED : SULI_FCG_ONEPASS_WANT,
                              capped ? SULI_FCG_ONEPASS_FLOOR_CAPPED : SULI_FCG_ONEPASS_FLOOR});
    }
    unet_capture(c, f.unet, f.r32, active, 0);
    // Preconditioned passes take few iterations and consecutive solves of a
    // run take about the same number: enqueue that many first, then two at a
    // time, so that few iterations are enqueued past the end (those still
    // cost their launches).  Jacobi passes take hundreds: chunks of 16..64.
    const int pidx = std::min(outer - 1, 3);
    int chunk = use_pc ? (pits[pidx] > 0 ? pits[pidx] : 4) : 16;
    int iter = 0;
    for (;;) {
      const unsigned seq_chunk = ++f.pub_seq;
      for (int n = 0; n < chunk; n++, iter++) {
        if (iter > 0) unet_capture(c, f.unet, f.r32, active, iter);
        f.cur_iter = iter;
        // the chunk's last finish publishes the control block
        FcgCtl *const pub = n == chunk - 1 ? f.ctl_hd : nullptr;
        if (pc == SULI_PRECOND_LEARNED && !learned_m1) {
          fcgm_iteration(c, f, iter, gk, active, pub, seq_chunk);
          continue;
        }
        const float *zin = f.r32;
        if (use_pc) {
          const ZqOut zo{f.q32, (const int *)((char *)f.ctl + offsetof(FcgCtl, first)),
                         RedSlot{c.partials, CTL_PTR(f.ctl, zq)}};
          mg_vcycle(c, f.r32, f.z32, active, &zo, pc == SULI_PRECOND_LEARNED ? entry : 0, stage, fine_f16(c, iter));
          zin = f.z32;
        }
        {
          const RedSlot r1{c.partials, CTL_PTR(f.ctl, k1)};
          {
            ProfKernel pk(c, "fcg_spmv_dot");
            SULI_TIMED_LAUNCH(pk, k_fcg1, gk, b2, 0, c.st, g, f.ax32, f.ay32, f.az32, zin, (const float *)f.r32,
                              (const float *)f.p32, f.p32b, f.q32, f.ctl, use_pc ? 1 : 0, r1);
          }
          launch_finish<2>(c.st, gk, r1, EpiK1{f.ctl});
        }
        std::swap(f.p32, f.p32b);
        {
          ProfScope ps(c, "fcg_update");
          const RedSlot r2{c.partials, CTL_PTR(f.ctl, k2)};
          hipLaunchKernelGGL(k_fcg2, gk, b2, 0, c.st, g, f.p32, f.q32, f.r32, f.d32, f.ctl, r2, sti(iter + 1));
          launch_finish<2>(c.st, gk, r2, EpiK2{f.ctl, pub, seq_chunk});
        }
      }
      HIPCHK(hipGetLastError());
      if (ctl_read(c, f, seq_chunk)) return 1;
      if (!f.ctl_h->active) break;
      if (use_pc) chunk = 2;
      else if (chunk < 64) chunk *= 2;
    }
    if (spec) {  // this pass's outer residual, read with its first chunk
      res = proj_res(f.ctl_h, bnorm);
      if (res0 < 0) res0 = res;
      if (f.ctl_h->done) {  // it met the tolerance: the pass did not run
        outer--;
        conv = true;
        break;
      }
    }
    inner_total += f.ctl_h->it;
    // the next pass is enqueued with this one's residual check when the FP32
    // reduction of this pass leaves the residual clearly above the tolerance
    {
      const double r0 = f.ctl_h->rin0[0], r1 = f.ctl_h->k2[0];
      const double pred = (r0 > 0 && r1 >= 0) ? res * std::sqrt(r1 / r0) : 0.0;
      speculate = pc != SULI_PRECOND_LEARNED && pred > 3.0 * tol;
    }
    if (use_pc) pits[pidx] = f.ctl_h->it;
    const double rr_end = f.ctl_h->k2[0];
    const bool finite = rr_end == rr_end && rr_end < 1e300;
    if (finite) {
      ProfScope ps(c, "fcg_xupdate");
      hipLaunchKernelGGL(k_xupd, gk, b2, 0, c.st, g, f.x, f.d32);
    }
    if (pc == SULI_PRECOND_LEARNED) {
      // stagnation, breakdown (p.q <= 0) or a non-finite residual: the rest of
      // this solve runs on multigrid
      const bool stalled = f.ctl_h->it >= pass_max && rr_end > f.ctl_h->thr_in;
      if (!finite || stalled || !(f.ctl_h->k1[0] > 0.0)) {
        pc = SULI_PRECOND_GMG;
        fallbacks++;
      }
    } else if (!finite) {
      char msg[320];
      snprintf(msg, sizeof msg,
               "fcg_solve: non-finite residual in the inner iteration (pass %d, %d iterations, r.r %g, p.q %g, "
               "p.r %g, r0.r0 %g, outer r.r %g, b.b %g)",
               outer, f.ctl_h->it, rr_end, f.ctl_h->k1[0], f.ctl_h->k1[1], f.ctl_h->rin0[0], f.ctl_h->rout[0],
               f.ctl_h->bb[0]);
      suli_set_error(msg);
      return 1;
    }
    if (inner_total >= maxit_in) {
      // iteration budget spent: report the residual after this pass
      const RedSlot rr{c.partials, CTL_PTR(f.ctl, rout)};
      hipLaunchKernelGGL(k_res64, gk, b2, 0, c.st, g, c.ax, c.ay, c.az, f.x, f.b, c.sD, f.r, rr);
      launch_finish<3>(c.st, gk, rr);
      HIPCHK(hipMemcpyAsync(f.ctl_h, f.ctl, sizeof(FcgCtl), hipMemcpyDeviceToHost, c.st));
      HIPCHK(hipStreamSynchronize(c.st));
      res = proj_res(f.ctl_h, bnorm);
      conv = !(res > tol);
      break;
    }
  }
  if (!pout_done) {
    ProfScope ps(c, "poisson_final");
    hipLaunchKernelGGL(k_pout, pg, dim3(64, 4, 1), 0, c.st, g, f.x, c.sD, c.prd1, dslot, (const int *)nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c.st));
  }
  HIPCHK(hipGetLastError());
  double gpu_ms = 0.0;
  if (solve_elapsed_ms(c, &gpu_ms)) return 1;
  if (dslot) {  // this solve's offset is the newest of its kind
    c.off[sk][2] = c.off[sk][1];
    c.off[sk][1] = c.off[sk][0];
    c.off[sk][0] = dslot;
    c.offn[sk] = std::min(c.offn[sk] + 1, 3);
  }
  c.last = sk + 1;
  if (stt) {
    std::memset(stt, 0, sizeof(*stt));
    stt->iters = inner_total;
    stt->outer = outer;
    stt->fallbacks = fallbacks;
    stt->converged = conv ? 1 : 0;
    stt->res0 = res0;
    stt->res = res;
    stt->ms = gpu_ms;
  }
  return 0;
}

int fcg_readbacks(Ctx &c, int64_t *published, int64_t *copied) {
  *published = c.fcg ? c.fcg->pub_hits : 0;
  *copied = c.fcg ? c.fcg->pub_misses : 0;
  return 0;
}

int fcg_precond_apply(Ctx &c, const float *r, float *z) {
  PoissonFCG &f = *c.fcg;
  if (!f.setup_done) {
    suli_set_error("suli_precond_apply: no FCG solve has set up the operator yet");
    return 1;
  }
  const int pc = c.P.precond;
  if (pc == SULI_PRECOND_LEARNED && unet_entry(f.unet) >= (int)f.lv.size()) {
    suli_set_error("suli_precond_apply: the network's entry level is not a level of the multigrid hierarchy");
    return 1;
  }
  if (pc == SULI_PRECOND_LEARNED && (unet_prepare(c, f.unet) || unet_level_density(c, f.unet))) return 1;
  // the loop-control flag the kernels test: set it for this one application
  FcgCtl *h = f.ctl_h;
  HIPCHK(hipMemcpyAsync(h, f.ctl, sizeof(FcgCtl), hipMemcpyDeviceToHost, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  const int saved = h->active;
  h->active = 1;
  int *active = (int *)((char *)f.ctl + offsetof(FcgCtl, active));
  HIPCHK(hipMemcpyAsync(active, &h->active, sizeof(int), hipMemcpyHostToDevice, c.st));
  if (pc == SULI_PRECOND_JACOBI)
    HIPCHK(hipMemcpyAsync(z, r, (size_t)c.g.total * sizeof(float), hipMemcpyDeviceToDevice, c.st));
  else
    apply_precond(c, pc, r, z, active);
  HIPCHK(hipGetLastError());
  h->active = saved;
  HIPCHK(hipMemcpyAsync(active, &h->active, sizeof(int), hipMemcpyHostToDevice, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  return 0;
}
