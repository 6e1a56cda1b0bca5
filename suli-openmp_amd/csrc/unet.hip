// Learned preconditioner for the projection's flexible CG: a 3-D
// convolutional U-Net applied to the scaled residual r_s and the
// log-density, z = rms(r) Net(r/rms(r), log rho).  Training and inference use
// this library's own kernels (unet_kernels.hpp): BF16 implicit-GEMM
// convolutions on MFMA with FP32 accumulation for forward, data gradient and
// weight gradient, FP32 master weights, Adam.
//
// Architecture with L levels (default 3) and C_l = min(16*2^l, 64) channels
// at level l (grid halved per level); all convolutions 3x3x3 with bias and
// ReLU except the output layer:
//   encoder  l = 0:        A0 = conv(x0: 2 -> C0)  S0 = conv(A0: C0 -> C0)
//            l = 1..L-2:   P_l = pool(S_{l-1})  A_l = conv(P_l)  S_l = conv(A_l)
//   bottom   l = L-1:      P = pool(S_{L-2})  A = conv(P: C_{L-2} -> C_{L-1})
//                          B = conv(A: C_{L-1} -> C_{L-2})
//   decoder  l = L-2..1:   U_l = up(prev) + S_l  D_l = conv(U_l: C_l -> C_{l-1})
//            l = 0:        U_0 = up(prev) + S_0  D_0 = conv(U_0: C0 -> C0)
//   output                 z = conv(D_0: C0 -> 1), no activation
// (pool: 2x2x2 average; up: nearest).  With L = 3 that is 16/32/64 channels
// and 181,233 parameters.  The first and last layers run as FP32 direct
// kernels, the others on MFMA.  Weight vector order: encoder (a, b per level),
// bottom (a, b), decoder from level L-2 down to 0, output; each layer's
// kernel W[27][c_out][c_in] then bias[c_out].
//
// Training loss on snapshots (r, rho) captured from the solver's own solves:
// the relative residual left after one preconditioned step with the optimal
// step length, L = 1 - (r.Az)^2 / (|r|^2 |Az|^2) = min_a |r - a A z|^2/|r|^2,
// with A the scaled projection operator of that density field (the
// coefficients depend only on rho: a_f = (1/rho_f)/sqrt(d_P d_Q)).  FCG is
// invariant to the scale of z, so the loss is too.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "fcg.hpp"
#include "unet_kernels.hpp"

namespace {

constexpr int kMaxLevels = 6;
constexpr int kMaxLayers = 3 * kMaxLevels;

struct LayerDesc {
  int ci, co;
};

inline dim3 g1(int64_t n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

// layer numbering of an L-level network
struct Arch {
  int L = 3;
  int C[kMaxLevels];
  int n = 0;
  LayerDesc lay[kMaxLayers];
  int ea(int l) const { return 2 * l; }
  int eb(int l) const { return 2 * l + 1; }
  int bot_a() const { return 2 * (L - 1); }
  int bot_b() const { return 2 * (L - 1) + 1; }
  int dec(int l) const { return 2 * L + (L - 2 - l); }
  int out() const { return 3 * L - 1; }
  void set(int levels) {
    L = levels;
    for (int l = 0; l < L; l++) C[l] = std::min(16 << l, 64);
    n = 3 * L;
    lay[ea(0)] = {2, C[0]};
    lay[eb(0)] = {C[0], C[0]};
    for (int l = 1; l <= L - 2; l++) {
      lay[ea(l)] = {C[l - 1], C[l]};
      lay[eb(l)] = {C[l], C[l]};
    }
    lay[bot_a()] = {C[L - 2], C[L - 1]};
    lay[bot_b()] = {C[L - 1], C[L - 2]};
    for (int l = L - 2; l >= 1; l--) lay[dec(l)] = {C[l], C[l - 1]};
    lay[dec(0)] = {C[0], C[0]};
    lay[out()] = {C[0], 1};
  }
  int64_t wcount(int q) const { return 27LL * lay[q].ci * lay[q].co; }
};

}  // namespace

struct UNetPrecond {
  int ready = 0;
  Arch a;
  int64_t nparam = 0;
  int64_t woff[kMaxLayers], boff[kMaxLayers];
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *wf[kMaxLayers] = {}, *wt[kMaxLayers] = {};
  // activations and their gradients for a level-0 grid of cap0 voxels
  Dims d0{0, 0, 0};
  int64_t cap0 = 0;
  float *x0 = nullptr, *z = nullptr, *dz = nullptr, *part = nullptr, *tmpf = nullptr;
  uint16_t *A[kMaxLevels] = {}, *S[kMaxLevels] = {}, *P[kMaxLevels] = {}, *U[kMaxLevels] = {}, *D[kMaxLevels] = {},
           *B = nullptr;
  uint16_t *gA[kMaxLevels] = {}, *gS[kMaxLevels] = {}, *gP[kMaxLevels] = {}, *gU[kMaxLevels] = {},
           *gD[kMaxLevels] = {}, *gB = nullptr;
  double *red = nullptr;  // reduction partials / results
  // training operator (compact grid) and scratch
  float *cx = nullptr, *cy = nullptr, *cz = nullptr, *s = nullptr, *t = nullptr, *rho1 = nullptr;
  // dataset captured from the solver (compact grid)
  Dims dsd{0, 0, 0};
  int ds_max = 0, ds_cap = 0, ds_every = 0;  // armed while ds_max > 0
  int *ds_cnt = nullptr;                       // samples held (device counter)
  float *ds_r = nullptr, *ds_rho = nullptr;
  // the network as the coarse solver of multigrid level `entry` (0: the
  // network is the whole preconditioner): block-mean density of that level
  // for the current solve, and the captured samples' face conductances
  int entry = 0;
  float *rho_e = nullptr;
  int64_t rho_e_cap = 0;
  double *rho_h[2] = {nullptr, nullptr};  // intermediate halvings (levels 1..e-1)
  int fcg_m = 4;            // directions FCG keeps for this preconditioner (1 or 4)
  int ds_entry = 0;         // entry level the held samples were captured at
  float *ds_f = nullptr;    // per sample: x, y, z faces of the level (compact)
  double lr1 = 0.0, lr2 = 0.0;  // log-density normalisation: (log rho - lr1)/lr2
  int step = 0;
};

// ---------------------------------------------------------------------------
// kernels specific to the preconditioner's interface with the solver

// x0 = [r/rms, (log rho - lr1)/lr2] from a frame-layout residual and density
template <typename T>
__global__ void k_prep_frame(Grid g, Dims d, const float *__restrict__ r, const T *__restrict__ rho,
                             const double *__restrict__ rr, double lr1, double lr2, float *__restrict__ x0,
                             const int *__restrict__ active) {
  if (active && !*active) return;
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  const int64_t c = g.idx(i + 1, j + 1, k + 1);
  const double rms = sqrt(rr[0] / (double)d.n());
  x0[2 * v] = rms > 0 ? (float)(r[c] / rms) : 0.f;
  x0[2 * v + 1] = (float)((log((double)rho[c]) - lr1) / lr2);
}

// the same from compact arrays (training)
__global__ void k_prep_compact(Dims d, const float *__restrict__ r, const float *__restrict__ rho,
                               const double *__restrict__ rr, double lr1, double lr2, float *__restrict__ x0) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const double rms = sqrt(rr[0] / (double)d.n());
  x0[2 * v] = rms > 0 ? (float)(r[v] / rms) : 0.f;
  x0[2 * v + 1] = (float)((log((double)rho[v]) - lr1) / lr2);
}

// sum of squares of a frame-layout FP32 vector over the cells -> partials
__global__ __launch_bounds__(256) void k_sumsq_frame(Grid g, const float *__restrict__ r, double *__restrict__ part,
                                                     const int *__restrict__ active) {
  if (active && !*active) return;
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y;
  double acc[1] = {0.0};
  if (i <= g.nx && j <= g.ny)
    for (int k = 1 + blockIdx.z; k <= g.nz; k += gridDim.z) {
      const double x = r[g.idx(i, j, k)];
      acc[0] += x * x;
    }
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1>(acc, lds);
  block_partial<1>(acc, part);
}

// z (frame) = rms * z_net (compact)
__global__ void k_scatter_frame(Grid g, Dims d, const float *__restrict__ zc, const double *__restrict__ rr,
                                float *__restrict__ z, const int *__restrict__ active) {
  if (active && !*active) return;
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  const float rms = (float)sqrt(rr[0] / (double)d.n());
  z[g.idx(i + 1, j + 1, k + 1)] = rms * zc[v];
}

// compact copies of the residual and density into sample *cnt of the set
template <typename T>
__global__ void k_capture(Grid g, Dims d, const float *__restrict__ r, const T *__restrict__ rho,
                          float *__restrict__ rs0, float *__restrict__ rhos0, const int *__restrict__ cnt, int max,
                          const int *__restrict__ active) {
  if (active && !*active) return;
  const int sidx = *cnt;
  if (sidx >= max) return;
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  float *rs = rs0 + (int64_t)sidx * d.n(), *rhos = rhos0 + (int64_t)sidx * d.n();
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  const int64_t c = g.idx(i + 1, j + 1, k + 1);
  rs[v] = r[c];
  rhos[v] = (float)rho[c];
}

__global__ void k_capture_bump(int *cnt, int max, const int *__restrict__ active) {
  if (active && !*active) return;
  if (threadIdx.x == 0 && *cnt < max) *cnt += 1;
}

// ---- the network as the coarse solver of a multigrid level -----------------
// Level e of the V-cycle (cells 1..n of its own frame, face conductances cx
// (face i between cells i-1 and i, zero on the walls), diagonal dg = their
// sum) holds the problem A x = b.  The network sees it Jacobi-scaled,
// b~ = b/sqrt(dg), with the block-mean density of the level's cells, and its
// output z~ maps back to y = z~/sqrt(dg); the correction is x = alpha y with
// the energy-optimal alpha = (y.b)/(y.A y).

// Compact block-mean density of level e, the mean of rho over the 2^e-cube
// of fine cells under each level-e cell, built as e halvings (each the mean
// of 8 children; the first reads the frame): every pass is one parallel
// sweep over its level.
template <typename T, typename TO>
__global__ void k_rho_halve_frame(Grid g, Dims d, const T *__restrict__ rho, TO *__restrict__ out) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  double s = 0.0;
  for (int c = 0; c < 8; c++) s += (double)rho[g.idx(1 + 2 * i + (c & 1), 1 + 2 * j + ((c >> 1) & 1), 1 + 2 * k + (c >> 2))];
  out[v] = (TO)(0.125 * s);
}
template <typename TO>
__global__ void k_rho_halve(Dims d, const double *__restrict__ in, TO *__restrict__ out) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  const int64_t fx = 2 * d.nx, fy = 2 * d.ny;
  double s = 0.0;
  for (int c = 0; c < 8; c++)
    s += in[(2 * i + (c & 1)) + fx * ((2 * j + ((c >> 1) & 1)) + fy * (int64_t)(2 * k + (c >> 2)))];
  out[v] = (TO)(0.125 * s);
}

// sum over the level's cells of b^2/dg -> partials (512-block grid-stride)
__global__ __launch_bounds__(256) void k_cs_sumsq(Grid ge, Dims d, const float *__restrict__ b,
                                                  const float *__restrict__ dg, double *__restrict__ part,
                                                  const int *__restrict__ active) {
  if (active && !*active) return;
  double acc[1] = {0.0};
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < d.n(); v += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
    const fidx c = ge.idx(i + 1, j + 1, k + 1);
    const double bv = b[c];
    acc[0] += bv * bv / (double)dg[c];
  }
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1>(acc, lds);
  block_partial<1>(acc, part);
}

// x0 = [b~/rms(b~), (log rho_e - lr1)/lr2]
__global__ void k_cs_prep(Grid ge, Dims d, const float *__restrict__ b, const float *__restrict__ dg,
                          const float *__restrict__ rho_e, const double *__restrict__ bb, double lr1, double lr2,
                          float *__restrict__ x0, const int *__restrict__ active) {
  if (active && !*active) return;
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  const fidx c = ge.idx(i + 1, j + 1, k + 1);
  const double rms = sqrt(bb[0] / (double)d.n());
  x0[2 * v] = rms > 0 ? (float)(b[c] / sqrt((double)dg[c]) / rms) : 0.f;
  x0[2 * v + 1] = (float)((log((double)rho_e[v]) - lr1) / lr2);
}

// y = z~/sqrt(dg) into the level's frame
__global__ void k_cs_scatter(Grid ge, Dims d, const float *__restrict__ zc, const float *__restrict__ dg,
                             float *__restrict__ x, const int *__restrict__ active) {
  if (active && !*active) return;
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  const fidx c = ge.idx(i + 1, j + 1, k + 1);
  x[c] = zc[v] / sqrtf(dg[c]);
}

// y.b and y.(A y) -> partials (2 per block); A y = dg y - sum_f c_f y_nb
__global__ __launch_bounds__(256) void k_cs_dots(Grid ge, Dims d, const float *__restrict__ cx,
                                                 const float *__restrict__ cy, const float *__restrict__ cz,
                                                 const float *__restrict__ dg, const float *__restrict__ b,
                                                 const float *__restrict__ y, double *__restrict__ part,
                                                 const int *__restrict__ active) {
  if (active && !*active) return;
  double acc[2] = {0.0, 0.0};
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < d.n(); v += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
    const fidx c = ge.idx(i + 1, j + 1, k + 1);
    const float nb = cx[c + 1] * y[c + 1] + cx[c] * y[c - 1] + cy[c + ge.sx] * y[c + ge.sx] + cy[c] * y[c - ge.sx] +
                     cz[c + ge.sz] * y[c + ge.sz] + cz[c] * y[c - ge.sz];
    const double yv = y[c], ay = (double)dg[c] * yv - (double)nb;
    acc[0] += yv * (double)b[c];
    acc[1] += yv * ay;
  }
  __shared__ double lds[2 * SULI_LDS_RED];
  block_reduce<2>(acc, lds);
  block_partial<2>(acc, part);
}

// x = alpha y, alpha = (y.b)/(y.A y) (0 when y.A y is not positive)
__global__ void k_cs_scale(Grid ge, Dims d, float *__restrict__ x, const double *__restrict__ dots,
                           const int *__restrict__ active) {
  if (active && !*active) return;
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const double yay = dots[1];
  const float alpha = yay > 0.0 ? (float)(dots[0] / yay) : 0.f;
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  const fidx c = ge.idx(i + 1, j + 1, k + 1);
  x[c] = alpha * x[c];
}

// training sample of a level's problem: b~ = b/sqrt(dg), rho_e and the face
// conductances in compact order (x faces i + (nx+1)(j + ny k), i = 0..nx;
// y and z alike), sample *cnt of the set
__global__ void k_capture_level(Grid ge, Dims d, const float *__restrict__ b, const float *__restrict__ dg,
                                const float *__restrict__ cx, const float *__restrict__ cy,
                                const float *__restrict__ cz, const float *__restrict__ rho_e,
                                float *__restrict__ rs0, float *__restrict__ rhos0, float *__restrict__ fs0,
                                const int *__restrict__ cnt, int max, const int *__restrict__ active) {
  if (active && !*active) return;
  const int sidx = *cnt;
  if (sidx >= max) return;
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int64_t nfx = (int64_t)(nx + 1) * ny * nz, nfy = (int64_t)nx * (ny + 1) * nz,
                nfz = (int64_t)nx * ny * (nz + 1);
  float *rs = rs0 + (int64_t)sidx * d.n(), *rhos = rhos0 + (int64_t)sidx * d.n();
  float *fx = fs0 + (int64_t)sidx * (nfx + nfy + nfz), *fy = fx + nfx, *fz = fy + nfy;
  const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((int64_t)nx * ny));
  const fidx c = ge.idx(i + 1, j + 1, k + 1);
  rs[v] = b[c] / sqrtf(dg[c]);
  rhos[v] = rho_e[v];
  fx[i + (int64_t)(nx + 1) * (j + (int64_t)ny * k)] = cx[c];
  if (i == nx - 1) fx[nx + (int64_t)(nx + 1) * (j + (int64_t)ny * k)] = cx[c + 1];
  fy[i + (int64_t)nx * (j + (int64_t)(ny + 1) * k)] = cy[c];
  if (j == ny - 1) fy[i + (int64_t)nx * (ny + (int64_t)(ny + 1) * k)] = cy[c + ge.sx];
  fz[i + (int64_t)nx * (j + (int64_t)ny * k)] = cz[c];
  if (k == nz - 1) fz[i + (int64_t)nx * (j + (int64_t)ny * nz)] = cz[c + ge.sz];
}

// Inverse face densities of the scaled operator on the compact grid: face
// (i) between cells i-1 and i, i = 0..n, with the end faces copying their
// neighbour (8_derivadas.f90:318-323); stored cx[(i) + (nx+1)*(j + ny*k)].
__global__ void k_train_faces(Dims d, float mx, float my, float mz, const float *__restrict__ rho,
                              float *__restrict__ cx, float *__restrict__ cy, float *__restrict__ cz) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  auto R = [&](int i, int j, int k) { return rho[i + (int64_t)nx * (j + (int64_t)ny * k)]; };
  const int64_t nfx = (int64_t)(nx + 1) * ny * nz, nfy = (int64_t)nx * (ny + 1) * nz, nfz = (int64_t)nx * ny * (nz + 1);
  if (e < nfx) {
    const int i = (int)(e % (nx + 1)), j = (int)((e / (nx + 1)) % ny), k = (int)(e / ((int64_t)(nx + 1) * ny));
    const int ic = i < 1 ? 1 : (i > nx - 1 ? nx - 1 : i);
    cx[e] = 2.f * mx / (R(ic, j, k) + R(ic - 1, j, k));
  } else if (e < nfx + nfy) {
    const int64_t f = e - nfx;
    const int i = (int)(f % nx), j = (int)((f / nx) % (ny + 1)), k = (int)(f / ((int64_t)nx * (ny + 1)));
    const int jc = j < 1 ? 1 : (j > ny - 1 ? ny - 1 : j);
    cy[f] = 2.f * my / (R(i, jc, k) + R(i, jc - 1, k));
  } else if (e < nfx + nfy + nfz) {
    const int64_t f = e - nfx - nfy;
    const int i = (int)(f % nx), j = (int)((f / nx) % ny), k = (int)(f / ((int64_t)nx * ny));
    const int kc = k < 1 ? 1 : (k > nz - 1 ? nz - 1 : k);
    cz[f] = 2.f * mz / (R(i, j, kc) + R(i, j, kc - 1));
  }
}

// s = A_s z on the compact grid (Neumann: the wall face couples a cell to itself)
__device__ __forceinline__ float train_diag(const Dims &d, const float *cx, const float *cy, const float *cz, int i,
                                            int j, int k) {
  const int nx = d.nx, ny = d.ny;
  return cx[i + (int64_t)(nx + 1) * (j + (int64_t)ny * k)] + cx[i + 1 + (int64_t)(nx + 1) * (j + (int64_t)ny * k)] +
         cy[i + (int64_t)nx * (j + (int64_t)(ny + 1) * k)] + cy[i + (int64_t)nx * (j + 1 + (int64_t)(ny + 1) * k)] +
         cz[i + (int64_t)nx * (j + (int64_t)ny * k)] + cz[i + (int64_t)nx * (j + (int64_t)ny * (k + 1))];
}
__global__ void k_train_As(Dims d, const float *__restrict__ cx, const float *__restrict__ cy,
                           const float *__restrict__ cz, const float *__restrict__ z, float *__restrict__ s) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int i = (int)(v % nx), j = (int)((v / nx) % ny), k = (int)(v / ((int64_t)nx * ny));
  const float dP = train_diag(d, cx, cy, cz, i, j, k);
  float acc = z[v];
  auto term = [&](float c, int ii, int jj, int kk) {
    const bool in = ii >= 0 && ii < nx && jj >= 0 && jj < ny && kk >= 0 && kk < nz;
    const int a = in ? ii : i, b = in ? jj : j, cc = in ? kk : k;
    const float dQ = in ? train_diag(d, cx, cy, cz, a, b, cc) : dP;
    acc -= c / sqrtf(dP * dQ) * z[a + (int64_t)nx * (b + (int64_t)ny * cc)];
  };
  term(cx[i + 1 + (int64_t)(nx + 1) * (j + (int64_t)ny * k)], i + 1, j, k);
  term(cx[i + (int64_t)(nx + 1) * (j + (int64_t)ny * k)], i - 1, j, k);
  term(cy[i + (int64_t)nx * (j + 1 + (int64_t)(ny + 1) * k)], i, j + 1, k);
  term(cy[i + (int64_t)nx * (j + (int64_t)(ny + 1) * k)], i, j - 1, k);
  term(cz[i + (int64_t)nx * (j + (int64_t)ny * (k + 1))], i, j, k + 1);
  term(cz[i + (int64_t)nx * (j + (int64_t)ny * k)], i, j, k - 1);
  s[v] = acc;
}

// sums r.s, s.s, r.r -> partials (3 per block)
__global__ __launch_bounds__(256) void k_train_dots(Dims d, const float *__restrict__ r, const float *__restrict__ s,
                                                    double *__restrict__ part) {
  double acc[3] = {0.0, 0.0, 0.0};
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < d.n(); v += (int64_t)gridDim.x * blockDim.x) {
    const double a = r[v], b = s[v];
    acc[0] += a * b;
    acc[1] += b * b;
    acc[2] += a * a;
  }
  __shared__ double lds[3 * SULI_LDS_RED];
  block_reduce<3>(acc, lds);
  block_partial<3>(acc, part);
}

// dL/ds for L = 1 - (r.s)^2/(|r|^2 |s|^2): t = -2 rs/(rr ss) r + 2 rs^2/(rr ss^2) s
__global__ void k_train_dLds(Dims d, const float *__restrict__ r, const float *__restrict__ s,
                             const double *__restrict__ dots, float *__restrict__ t) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const double rs = dots[0], ss = dots[1], rr = dots[2];
  if (!(rr > 0.0) || !(ss > 0.0)) {  // a zero residual or output: no gradient
    t[v] = 0.f;
    return;
  }
  const double a = -2.0 * rs / (rr * ss), b = 2.0 * rs * rs / (rr * ss * ss);
  t[v] = (float)(a * r[v] + b * s[v]);
}

// ---------------------------------------------------------------------------
// host side

namespace {

constexpr int kRedPart = 1 << 16;  // reduction partials; results follow
constexpr int kResNorm = kRedPart, kResDots = kRedPart + 8, kResCs = kRedPart + 16, kResCsDots = kRedPart + 24;

inline Dims level_dims(Dims d0, int l) { return Dims{d0.nx >> l, d0.ny >> l, d0.nz >> l}; }

// Grids of fewer voxels than this have too few 64x4x2 tiles to fill the
// chip: their convolutions take one block per 16-voxel row group (taps split
// over the block's wavefronts) and the FP32 first and last layers one
// 64-thread block per 64 voxels.
constexpr int64_t kSmallGrid = (int64_t)1 << 18;

template <int CIN, int COUT, int FLAGS>
void conv_t(hipStream_t st, Dims d, const uint16_t *in, const uint16_t *w, const float *bias, const uint16_t *gate,
            uint16_t *out, const int *active) {
  if (CIN == 16 && COUT == 16 && d.n() >= kSmallGrid) {
    const dim3 grid((d.nx + C16_TX - 1) / C16_TX, (d.ny + C16_TY - 1) / C16_TY, (d.nz + C16_TZ - 1) / C16_TZ);
    hipLaunchKernelGGL((k_conv16<FLAGS>), grid, dim3(256), 0, st, d, in, w, bias, gate, out, active);
  } else if (d.n() < kSmallGrid) {  // coarse grids: split the taps over the block's wavefronts
    hipLaunchKernelGGL((k_conv_mfma_split<CIN, COUT, FLAGS>), dim3((d.nx + 15) / 16, d.ny, d.nz), dim3(256), 0, st, d,
                       in, w, bias, gate, out, active);
  } else {
    hipLaunchKernelGGL((k_conv_mfma<CIN, COUT, FLAGS>), dim3((d.nx + 63) / 64, d.ny, d.nz), dim3(256), 0, st, d, in,
                       w, bias, gate, out, active);
  }
}

// run-time dispatch over the channel pairs of the architecture
enum ConvKind { FWD = CONV_BIAS | CONV_RELU, DGRAD = 0, DGRAD_MASK = CONV_MASK };
template <int FLAGS>
void conv_ci_co(int ci, int co, hipStream_t st, Dims d, const uint16_t *in, const uint16_t *w, const float *bias,
                const uint16_t *gate, uint16_t *out, const int *active) {
#define SULI_CONV_CASE(A, B) \
  if (ci == A && co == B) return conv_t<A, B, FLAGS>(st, d, in, w, bias, gate, out, active);
  SULI_CONV_CASE(16, 16) SULI_CONV_CASE(16, 32) SULI_CONV_CASE(32, 16) SULI_CONV_CASE(32, 32)
  SULI_CONV_CASE(32, 64) SULI_CONV_CASE(64, 32) SULI_CONV_CASE(64, 64)
#undef SULI_CONV_CASE
}
void conv(ConvKind kind, int ci, int co, hipStream_t st, Dims d, const uint16_t *in, const uint16_t *w,
          const float *bias, const uint16_t *gate, uint16_t *out, const int *active = nullptr) {
  if (kind == FWD) conv_ci_co<FWD>(ci, co, st, d, in, w, bias, gate, out, active);
  else if (kind == DGRAD) conv_ci_co<DGRAD>(ci, co, st, d, in, w, bias, gate, out, active);
  else conv_ci_co<DGRAD_MASK>(ci, co, st, d, in, w, bias, gate, out, active);
}

// chunking of the weight-gradient reductions: at most 64 chunks of whole
// 32-voxel groups
inline int wchunk(int64_t n) { return (int)(((n + 63) / 64 + 31) / 32 * 32); }
inline int wnchunk(int64_t n) { return (int)((n + wchunk(n) - 1) / wchunk(n)); }

// dW, db of layer q from the pre-activation gradient dy (c_out channels) and
// the layer input x (c_in channels), both BF16 at dims d
void wgrad(UNetPrecond &u, hipStream_t st, Dims d, const uint16_t *dy, const uint16_t *x, int q) {
  const int64_t n = d.n();
  const int chunk = wchunk(n), nch = wnchunk(n);
  const int ci = u.a.lay[q].ci, co = u.a.lay[q].co;
  const dim3 grid(27, nch);
#define SULI_WG_CASE(A, B)                                                                                    \
  if (ci == A && co == B) hipLaunchKernelGGL((k_wgrad_mfma<A, B>), grid, dim3(256), 0, st, d, dy, x, u.part, chunk);
  SULI_WG_CASE(16, 16) SULI_WG_CASE(16, 32) SULI_WG_CASE(32, 16) SULI_WG_CASE(32, 32)
  SULI_WG_CASE(32, 64) SULI_WG_CASE(64, 32) SULI_WG_CASE(64, 64)
#undef SULI_WG_CASE
  const int64_t len = u.a.wcount(q);
  hipLaunchKernelGGL(k_sum_chunks, g1(len), dim3(256), 0, st, u.part, nch, len, u.g + u.woff[q]);
  hipLaunchKernelGGL((k_bias_grad<uint16_t>), dim3(nch), dim3(256), 0, st, d, co, dy, u.part, chunk);
  hipLaunchKernelGGL(k_sum_chunks, g1(co), dim3(256), 0, st, u.part, nch, (int64_t)co, u.g + u.boff[q]);
}

void finish_sum(hipStream_t st, double *part, int nb, double *res, int nv) {
  if (nv == 1)
    hipLaunchKernelGGL((k_finish<1, false, EpiNone>), dim3(1), dim3(SULI_FINISH_THREADS), 0, st, part, nb, res,
                       EpiNone());
  else
    hipLaunchKernelGGL((k_finish<3, false, EpiNone>), dim3(1), dim3(SULI_FINISH_THREADS), 0, st, part, nb, res,
                       EpiNone());
}

int alloc_f(Ctx &c, float **p, int64_t n) {
  HIPCHK(hipMalloc((void **)p, (size_t)std::max<int64_t>(n, 1) * sizeof(float)));
  HIPCHK(hipMemsetAsync(*p, 0, (size_t)std::max<int64_t>(n, 1) * sizeof(float), c.st));
  c.allocs.push_back(*p);
  return 0;
}
int alloc_h(Ctx &c, uint16_t **p, int64_t n) {
  HIPCHK(hipMalloc((void **)p, (size_t)std::max<int64_t>(n, 1) * sizeof(uint16_t)));
  HIPCHK(hipMemsetAsync(*p, 0, (size_t)std::max<int64_t>(n, 1) * sizeof(uint16_t), c.st));
  c.allocs.push_back(*p);
  return 0;
}

bool dims_ok(const Grid &g, int L) {
  const int m = 1 << (L - 1);
  return g.nx % m == 0 && g.ny % m == 0 && g.nz % m == 0;
}

// activations, gradients and training scratch for a level-0 grid d (its
// dimensions divisible by 2^(L-1)); grown on demand, never shrunk
int alloc_acts(Ctx &c, UNetPrecond &u, Dims d) {
  const int64_t n0 = d.n();
  if (n0 <= u.cap0) {
    u.d0 = d;
    return 0;
  }
  const Arch &a = u.a;
  const int L = a.L;
  if (alloc_f(c, &u.x0, 2 * n0) || alloc_f(c, &u.z, n0) || alloc_f(c, &u.dz, n0) || alloc_f(c, &u.tmpf, n0)) return 1;
  for (int l = 0; l < L; l++) {
    const int64_t nl = n0 >> (3 * l);
    const int Cl = a.C[l];
    const int Cd = l == 0 ? a.C[0] : a.C[l - 1];  // decoder output channels at level l
    if (alloc_h(c, &u.A[l], Cl * nl) || alloc_h(c, &u.gA[l], Cl * nl)) return 1;
    if (l >= 1 && (alloc_h(c, &u.P[l], a.C[l - 1] * nl) || alloc_h(c, &u.gP[l], a.C[l - 1] * nl))) return 1;
    if (l <= L - 2 && (alloc_h(c, &u.S[l], Cl * nl) || alloc_h(c, &u.gS[l], Cl * nl) || alloc_h(c, &u.U[l], Cl * nl) ||
                       alloc_h(c, &u.gU[l], Cl * nl) || alloc_h(c, &u.D[l], Cd * nl) || alloc_h(c, &u.gD[l], Cd * nl)))
      return 1;
  }
  const int64_t nb = n0 >> (3 * (L - 1));
  if (alloc_h(c, &u.B, a.C[L - 2] * nb) || alloc_h(c, &u.gB, a.C[L - 2] * nb)) return 1;
  // face arrays: (n+1) faces along one direction, at most 2 n0 entries
  if (alloc_f(c, &u.cx, 2 * n0) || alloc_f(c, &u.cy, 2 * n0) || alloc_f(c, &u.cz, 2 * n0) || alloc_f(c, &u.s, n0) ||
      alloc_f(c, &u.t, n0) || alloc_f(c, &u.rho1, n0))
    return 1;
  u.cap0 = n0;
  u.d0 = d;
  return 0;
}

// forward pass from x0 (already prepared); z_net to u.z
void forward(Ctx &c, UNetPrecond &u, const int *active) {
  hipStream_t st = c.st;
  const Arch &a = u.a;
  const int L = a.L;
  const float *W = u.w;
  const Dims d0 = u.d0;
  const int q0 = a.ea(0);
  const dim3 tiles((d0.nx + C16_TX - 1) / C16_TX, (d0.ny + C16_TY - 1) / C16_TY, (d0.nz + C16_TZ - 1) / C16_TZ);
  const bool small = d0.n() < kSmallGrid;
  if (small)
    hipLaunchKernelGGL(k_conv_in, g1(d0.n(), 64), dim3(64), 0, st, d0, u.x0, W + u.woff[q0], W + u.boff[q0], u.A[0],
                       active);
  else
    hipLaunchKernelGGL(k_conv_in_tile, tiles, dim3(256), 0, st, d0, u.x0, W + u.woff[q0], W + u.boff[q0], u.A[0],
                       active);
  auto fc = [&](int q, Dims d, const uint16_t *in, uint16_t *out) {
    conv(FWD, a.lay[q].ci, a.lay[q].co, st, d, in, u.wf[q], W + u.boff[q], nullptr, out, active);
  };
  fc(a.eb(0), d0, u.A[0], u.S[0]);
  for (int l = 1; l < L; l++) {
    const Dims dl = level_dims(d0, l), du = level_dims(d0, l - 1);
    const int Cin = a.C[l - 1];
    hipLaunchKernelGGL(k_pool, g1(dl.n() * Cin / 8), dim3(256), 0, st, du, Cin, u.S[l - 1], u.P[l], active);
    if (l < L - 1) {
      fc(a.ea(l), dl, u.P[l], u.A[l]);
      fc(a.eb(l), dl, u.A[l], u.S[l]);
    } else {
      fc(a.bot_a(), dl, u.P[l], u.A[l]);
      fc(a.bot_b(), dl, u.A[l], u.B);
    }
  }
  const uint16_t *prev = u.B;
  for (int l = L - 2; l >= 0; l--) {
    const Dims dl = level_dims(d0, l);
    hipLaunchKernelGGL(k_up_add, g1(dl.n() * a.C[l] / 8), dim3(256), 0, st, dl, a.C[l], prev, u.S[l], u.U[l], active);
    fc(a.dec(l), dl, u.U[l], u.D[l]);
    prev = u.D[l];
  }
  const int qo = a.out();
  if (small)
    hipLaunchKernelGGL(k_conv_out, g1(d0.n(), 64), dim3(64), 0, st, d0, u.D[0], W + u.woff[qo], W + u.boff[qo], u.z,
                       active);
  else
    hipLaunchKernelGGL(k_conv_out_tile, tiles, dim3(256), 0, st, d0, u.D[0], W + u.woff[qo], W + u.boff[qo], u.z,
                       active);
}

// backward pass from u.dz; gradients into u.g (overwritten)
void backward(Ctx &c, UNetPrecond &u) {
  hipStream_t st = c.st;
  const Arch &a = u.a;
  const int L = a.L;
  const Dims d0 = u.d0;
  const int64_t n0 = d0.n();
  const int chunk = wchunk(n0), nch = wnchunk(n0);
  auto dg = [&](ConvKind kind, int q, Dims d, const uint16_t *dy, const uint16_t *gate, uint16_t *out) {
    // data gradient of layer q: the flipped, transposed kernel maps c_out -> c_in
    conv(kind, a.lay[q].co, a.lay[q].ci, st, d, dy, u.wt[q], nullptr, gate, out);
  };
  // output layer: z = conv(D0) + b
  const int qo = a.out();
  hipLaunchKernelGGL((k_wgrad_small<16, 1, float, uint16_t>), dim3(27, nch), dim3(256), 0, st, d0, u.dz, u.D[0],
                     u.part, chunk);
  hipLaunchKernelGGL(k_sum_chunks, g1(a.wcount(qo)), dim3(256), 0, st, u.part, nch, a.wcount(qo), u.g + u.woff[qo]);
  hipLaunchKernelGGL((k_bias_grad<float>), dim3(nch), dim3(256), 0, st, d0, 1, u.dz, u.part, chunk);
  hipLaunchKernelGGL(k_sum_chunks, g1(1), dim3(256), 0, st, u.part, nch, (int64_t)1, u.g + u.boff[qo]);
  hipLaunchKernelGGL(k_conv_out_bwd, g1(n0), dim3(256), 0, st, d0, u.dz, u.w + u.woff[qo], u.D[0], u.gD[0]);
  // decoder, top down: D_l = relu(conv(U_l)), U_l = up(prev) + S_l
  for (int l = 0; l <= L - 2; l++) {
    const Dims dl = level_dims(d0, l), dn = level_dims(d0, l + 1);
    const int q = a.dec(l);
    wgrad(u, st, dl, u.gD[l], u.U[l], q);
    dg(DGRAD, q, dl, u.gD[l], nullptr, u.gU[l]);
    const bool bottom = l + 1 == L - 1;
    hipLaunchKernelGGL(k_down_sum, g1(dn.n() * a.C[l] / 8), dim3(256), 0, st, dl, a.C[l], u.gU[l],
                       bottom ? u.B : u.D[l + 1], bottom ? u.gB : u.gD[l + 1]);
  }
  // bottom: B = relu(conv(A)), A = relu(conv(P))
  {
    const int l = L - 1;
    const Dims dl = level_dims(d0, l);
    wgrad(u, st, dl, u.gB, u.A[l], a.bot_b());
    dg(DGRAD_MASK, a.bot_b(), dl, u.gB, u.A[l], u.gA[l]);
    wgrad(u, st, dl, u.gA[l], u.P[l], a.bot_a());
    dg(DGRAD, a.bot_a(), dl, u.gA[l], nullptr, u.gP[l]);
  }
  // encoder, bottom up: S_l feeds the pool of level l+1 and the skip of U_l
  for (int l = L - 2; l >= 0; l--) {
    const Dims dl = level_dims(d0, l);
    hipLaunchKernelGGL(k_add_upavg_mask, g1(dl.n() * a.C[l] / 8), dim3(256), 0, st, dl, a.C[l], u.gU[l], u.gP[l + 1],
                       u.S[l], u.gS[l]);
    wgrad(u, st, dl, u.gS[l], u.A[l], a.eb(l));
    dg(DGRAD_MASK, a.eb(l), dl, u.gS[l], u.A[l], u.gA[l]);
    if (l > 0) {
      wgrad(u, st, dl, u.gA[l], u.P[l], a.ea(l));
      dg(DGRAD, a.ea(l), dl, u.gA[l], nullptr, u.gP[l]);
    }
  }
  // first layer: A0 = relu(conv(x0) + b), FP32 input
  const int q0 = a.ea(0);
  hipLaunchKernelGGL((k_wgrad_small<2, 16, uint16_t, float>), dim3(27, nch), dim3(256), 0, st, d0, u.gA[0], u.x0,
                     u.part, chunk);
  hipLaunchKernelGGL(k_sum_chunks, g1(a.wcount(q0)), dim3(256), 0, st, u.part, nch, a.wcount(q0), u.g + u.woff[q0]);
  hipLaunchKernelGGL((k_bias_grad<uint16_t>), dim3(nch), dim3(256), 0, st, d0, 16, u.gA[0], u.part, chunk);
  hipLaunchKernelGGL(k_sum_chunks, g1(16), dim3(256), 0, st, u.part, nch, (int64_t)16, u.g + u.boff[q0]);
}

void refresh_bf16(Ctx &c, UNetPrecond &u) {
  const Arch &a = u.a;
  for (int q = 0; q < a.n; q++)
    if (u.wf[q])
      hipLaunchKernelGGL(k_weights_bf16, g1(a.wcount(q)), dim3(256), 0, c.st, u.w + u.woff[q], a.lay[q].co,
                         a.lay[q].ci, u.wf[q], u.wt[q]);
}

inline int64_t nfaces(Dims d) {
  return (int64_t)(d.nx + 1) * d.ny * d.nz + (int64_t)d.nx * (d.ny + 1) * d.nz + (int64_t)d.nx * d.ny * (d.nz + 1);
}

// Loss terms and dL/dz for a compact (r, rho) pair at dims u.d0: dots
// (r.s, s.s, r.r) with s = A_s z land at red[kResDots..+2], dL/dz in u.dz.
// A_s is the scaled operator of the face conductances `faces` (x, y, z
// faces in compact order, a multigrid level's captured operator) or, when
// faces is null, of the projection on rho at the context's grid spacing.
void loss_and_dz(Ctx &c, UNetPrecond &u, const float *r, const float *rho, const float *faces = nullptr) {
  hipStream_t st = c.st;
  const Dims d = u.d0;
  const int64_t n = d.n();
  const float *cx = u.cx, *cy = u.cy, *cz = u.cz;
  if (faces) {
    cx = faces;
    cy = cx + (int64_t)(d.nx + 1) * d.ny * d.nz;
    cz = cy + (int64_t)d.nx * (d.ny + 1) * d.nz;
  } else {
    const float mx = (float)(1.0 / (c.g.dx * c.g.dx)), my = (float)(1.0 / (c.g.dy * c.g.dy)),
                mz = (float)(1.0 / (c.g.dz * c.g.dz));
    hipLaunchKernelGGL(k_train_faces, g1(nfaces(d)), dim3(256), 0, st, d, mx, my, mz, rho, u.cx, u.cy, u.cz);
  }
  // |r|^2 for the input normalisation
  hipLaunchKernelGGL(k_train_dots, dim3(512), dim3(256), 0, st, d, r, r, u.red);
  finish_sum(st, u.red, 512, u.red + kResNorm, 3);
  hipLaunchKernelGGL(k_prep_compact, g1(n), dim3(256), 0, st, d, r, rho, u.red + kResNorm + 2, u.lr1, u.lr2, u.x0);
  forward(c, u, nullptr);
  hipLaunchKernelGGL(k_train_As, g1(n), dim3(256), 0, st, d, cx, cy, cz, u.z, u.s);
  hipLaunchKernelGGL(k_train_dots, dim3(512), dim3(256), 0, st, d, r, u.s, u.red);
  finish_sum(st, u.red, 512, u.red + kResDots, 3);
  hipLaunchKernelGGL(k_train_dLds, g1(n), dim3(256), 0, st, d, r, u.s, u.red + kResDots, u.t);
  hipLaunchKernelGGL(k_train_As, g1(n), dim3(256), 0, st, d, cx, cy, cz, u.t, u.dz);
}

// (re)build the parameter storage for an L-level network
int configure(Ctx &c, UNetPrecond &u, int L) {
  u.a.set(L);
  int64_t off = 0;
  for (int q = 0; q < u.a.n; q++) {
    u.woff[q] = off;
    off += u.a.wcount(q);
    u.boff[q] = off;
    off += u.a.lay[q].co;
  }
  u.nparam = off;
  if (alloc_f(c, &u.w, off) || alloc_f(c, &u.g, off) || alloc_f(c, &u.m, off) || alloc_f(c, &u.v, off)) return 1;
  for (int q = 0; q < u.a.n; q++) {
    u.wf[q] = u.wt[q] = nullptr;
    if (q == u.a.ea(0) || q == u.a.out()) continue;  // FP32 direct kernels
    if (alloc_h(c, &u.wf[q], u.a.wcount(q)) || alloc_h(c, &u.wt[q], u.a.wcount(q))) return 1;
  }
  // activations are rebuilt for the new architecture on the next use
  u.cap0 = 0;
  u.ready = 0;
  u.step = 0;
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
int unet_create(Ctx &c, UNetPrecond **out) {
  UNetPrecond *u = new UNetPrecond();
  *out = u;
  if (configure(c, *u, 3)) return 1;
  HIPCHK(hipMalloc((void **)&u->red, (kRedPart + 64) * sizeof(double)));
  c.allocs.push_back(u->red);
  u->lr1 = 0.5 * (std::log(c.ph.rho_f1) + std::log(c.ph.rho_f2));
  u->lr2 = std::max(0.5 * std::fabs(std::log(c.ph.rho_f2) - std::log(c.ph.rho_f1)), 1e-6);
  // weight-gradient partials: up to 64 chunks x 27 taps x the largest layer (64x64)
  if (alloc_f(c, &u->part, 64LL * 27 * 64 * 64)) return 1;
  return 0;
}

void unet_destroy(Ctx &c, UNetPrecond *u) {
  (void)c;
  delete u;
}

bool unet_ready(const UNetPrecond *u) { return u && u->ready; }

// grid of the network: the fine grid (entry 0) or multigrid level `entry`
static Dims entry_dims(const Grid &g, int e) { return Dims{g.nx >> e, g.ny >> e, g.nz >> e}; }
static bool entry_fits(const Grid &g, int e, int L) {
  const int m = 1 << (e + L - 1);
  return g.nx % m == 0 && g.ny % m == 0 && g.nz % m == 0;
}
static std::string entry_msg(const UNetPrecond &u) {
  return "nx, ny and nz must be multiples of 2^(entry+levels-1) = " + std::to_string(1 << (u.entry + u.a.L - 1)) +
         " (entry level " + std::to_string(u.entry) + ", " + std::to_string(u.a.L) + " levels)";
}

int unet_prepare(Ctx &c, UNetPrecond *u) {
  if (!u || !u->ready) {
    suli_set_error("learned preconditioner selected but no weights are loaded (suli_unet_init or "
                   "suli_unet_set_weights)");
    return 1;
  }
  if (!entry_fits(c.g, u->entry, u->a.L)) {
    suli_set_error("learned preconditioner: " + entry_msg(*u));
    return 1;
  }
  return alloc_acts(c, *u, entry_dims(c.g, u->entry));
}

int unet_entry(const UNetPrecond *u) { return u ? u->entry : 0; }

int unet_fcg_m(const UNetPrecond *u) { return u ? u->fcg_m : 4; }

bool unet_capture_armed(const UNetPrecond *u) { return u && u->ds_max > 0; }

int unet_level_density(Ctx &c, UNetPrecond *u) {
  if (!u || u->entry <= 0) return 0;
  const int e = u->entry;
  const Dims d = entry_dims(c.g, e);
  if (d.n() > u->rho_e_cap) {
    // (level-1 sized intermediates serve every entry level of this grid)
    const int64_t n1 = entry_dims(c.g, 1).n();
    if (alloc_f(c, &u->rho_e, d.n())) return 1;
    for (double *&h : u->rho_h) {
      HIPCHK(hipMalloc((void **)&h, (size_t)n1 * sizeof(double)));
      c.allocs.push_back(h);
    }
    u->rho_e_cap = d.n();
  }
  // level 1 from the frame, then halvings; the last pass writes rho_e (FP32)
  const Dims d1 = entry_dims(c.g, 1);
  if (e == 1) {
    if (c.prec == SULI_PREC_FP64)
      hipLaunchKernelGGL((k_rho_halve_frame<double, float>), g1(d1.n()), dim3(256), 0, c.st, c.g, d1,
                         (const double *)c.rho, u->rho_e);
    else
      hipLaunchKernelGGL((k_rho_halve_frame<float, float>), g1(d1.n()), dim3(256), 0, c.st, c.g, d1,
                         (const float *)c.rho, u->rho_e);
  } else {
    if (c.prec == SULI_PREC_FP64)
      hipLaunchKernelGGL((k_rho_halve_frame<double, double>), g1(d1.n()), dim3(256), 0, c.st, c.g, d1,
                         (const double *)c.rho, u->rho_h[0]);
    else
      hipLaunchKernelGGL((k_rho_halve_frame<float, double>), g1(d1.n()), dim3(256), 0, c.st, c.g, d1,
                         (const float *)c.rho, u->rho_h[0]);
    for (int l = 2; l <= e; l++) {
      const Dims dl = entry_dims(c.g, l);
      const double *in = u->rho_h[(l - 2) % 2];  // level l-1
      if (l == e)
        hipLaunchKernelGGL((k_rho_halve<float>), g1(dl.n()), dim3(256), 0, c.st, dl, in, u->rho_e);
      else
        hipLaunchKernelGGL((k_rho_halve<double>), g1(dl.n()), dim3(256), 0, c.st, dl, in, u->rho_h[(l - 1) % 2]);
    }
  }
  HIPCHK(hipGetLastError());
  return 0;
}

void unet_coarse_solve(Ctx &c, UNetPrecond *u, const Grid &ge, const float *cx, const float *cy, const float *cz,
                       const float *dg, const float *b, float *x, const int *active) {
  const Dims d{ge.nx, ge.ny, ge.nz};
  const int64_t n = d.n();
  ProfScope ps(c, "unet_coarse_solve");
  hipLaunchKernelGGL(k_cs_sumsq, dim3(512), dim3(256), 0, c.st, ge, d, b, dg, u->red, active);
  finish_sum(c.st, u->red, 512, u->red + kResCs, 1);
  hipLaunchKernelGGL(k_cs_prep, g1(n), dim3(256), 0, c.st, ge, d, b, dg, (const float *)u->rho_e,
                     (const double *)(u->red + kResCs), u->lr1, u->lr2, u->x0, active);
  forward(c, *u, active);
  hipLaunchKernelGGL(k_cs_scatter, g1(n), dim3(256), 0, c.st, ge, d, (const float *)u->z, dg, x, active);
  hipLaunchKernelGGL(k_cs_dots, dim3(512), dim3(256), 0, c.st, ge, d, cx, cy, cz, dg, b, (const float *)x, u->red,
                     active);
  hipLaunchKernelGGL((k_finish<2, false, EpiNone>), dim3(1), dim3(SULI_FINISH_THREADS), 0, c.st, u->red, 512,
                     u->red + kResCsDots, EpiNone());
  hipLaunchKernelGGL(k_cs_scale, g1(n), dim3(256), 0, c.st, ge, d, x, (const double *)(u->red + kResCsDots),
                     active);
}

void unet_capture_level(Ctx &c, UNetPrecond *u, const Grid &ge, const float *cx, const float *cy, const float *cz,
                        const float *dg, const float *b, const int *active, int iter) {
  if (!u || u->ds_max <= 0 || u->ds_entry <= 0 || !u->rho_e) return;
  if (iter > 0 && (u->ds_every <= 0 || iter % u->ds_every != 0)) return;
  const Dims d{ge.nx, ge.ny, ge.nz};
  if (d.nx != u->dsd.nx || d.ny != u->dsd.ny || d.nz != u->dsd.nz) return;
  hipLaunchKernelGGL(k_capture_level, g1(d.n()), dim3(256), 0, c.st, ge, d, b, dg, cx, cy, cz,
                     (const float *)u->rho_e, u->ds_r, u->ds_rho, u->ds_f, u->ds_cnt, u->ds_max, active);
  hipLaunchKernelGGL(k_capture_bump, dim3(1), dim3(64), 0, c.st, u->ds_cnt, u->ds_max, active);
}

void unet_apply(Ctx &c, UNetPrecond *u, const float *r, float *z, const int *active) {
  const Grid &g = c.g;
  const Dims d{g.nx, g.ny, g.nz};
  const int64_t n = d.n();
  ProfScope ps(c, "unet_apply");
  const int64_t plane = (int64_t)((g.nx + SULI_BX - 1) / SULI_BX) * ((g.ny + SULI_BY - 1) / SULI_BY);
  const int kz = (int)std::max<int64_t>(1, std::min<int64_t>(g.nz, 2048 / std::max<int64_t>(plane, 1)));
  const dim3 gk((g.nx + SULI_BX - 1) / SULI_BX, (g.ny + SULI_BY - 1) / SULI_BY, kz);
  hipLaunchKernelGGL(k_sumsq_frame, gk, dim3(SULI_BX, SULI_BY, 1), 0, c.st, g, r, u->red, active);
  finish_sum(c.st, u->red, (int)(gk.x * gk.y * gk.z), u->red + kResNorm, 1);
  if (c.prec == SULI_PREC_FP64)
    hipLaunchKernelGGL((k_prep_frame<double>), g1(n), dim3(256), 0, c.st, g, d, r, (const double *)c.rho,
                       u->red + kResNorm, u->lr1, u->lr2, u->x0, active);
  else
    hipLaunchKernelGGL((k_prep_frame<float>), g1(n), dim3(256), 0, c.st, g, d, r, (const float *)c.rho,
                       u->red + kResNorm, u->lr1, u->lr2, u->x0, active);
  forward(c, *u, active);
  hipLaunchKernelGGL(k_scatter_frame, g1(n), dim3(256), 0, c.st, g, d, u->z, u->red + kResNorm, z, active);
}

void unet_capture(Ctx &c, UNetPrecond *u, const float *r, const int *active, int iter) {
  if (!u || u->ds_max <= 0 || u->ds_entry != 0) return;
  if (iter > 0 && (u->ds_every <= 0 || iter % u->ds_every != 0)) return;
  const Grid &g = c.g;
  const Dims d{g.nx, g.ny, g.nz};
  const int64_t n = d.n();
  if (c.prec == SULI_PREC_FP64)
    hipLaunchKernelGGL((k_capture<double>), g1(n), dim3(256), 0, c.st, g, d, r, (const double *)c.rho, u->ds_r,
                       u->ds_rho, u->ds_cnt, u->ds_max, active);
  else
    hipLaunchKernelGGL((k_capture<float>), g1(n), dim3(256), 0, c.st, g, d, r, (const float *)c.rho, u->ds_r,
                       u->ds_rho, u->ds_cnt, u->ds_max, active);
  hipLaunchKernelGGL(k_capture_bump, dim3(1), dim3(64), 0, c.st, u->ds_cnt, u->ds_max, active);
}

// samples held (synchronises the stream)
static int dataset_count(Ctx &c, UNetPrecond &u, int *n) {
  *n = 0;
  if (!u.ds_cnt) return 0;
  HIPCHK(hipMemcpyAsync(n, u.ds_cnt, sizeof(int), hipMemcpyDeviceToHost, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  return 0;
}

// ---------------------------------------------------------------------------
// C ABI

#define NEEDU()                                              \
  do {                                                       \
    if (!g_ctx || !g_ctx->fcg) {                             \
      suli_set_error("library not initialised (suli_init)"); \
      return 2;                                              \
    }                                                        \
  } while (0)

extern "C" int suli_unet_configure(int levels) {
  NEEDU();
  Ctx &c = *g_ctx;
  if (levels < 2 || levels > kMaxLevels) {
    suli_set_error("suli_unet_configure: levels must be in 2.." + std::to_string(kMaxLevels));
    return 1;
  }
  return configure(c, *fcg_unet(c), levels);
}

extern "C" int suli_unet_levels(void) {
  if (!g_ctx || !g_ctx->fcg) return -1;
  return fcg_unet(*g_ctx)->a.L;
}

extern "C" int64_t suli_unet_nparams(void) {
  if (!g_ctx || !g_ctx->fcg) return -1;
  return fcg_unet(*g_ctx)->nparam;
}

extern "C" int suli_unet_init(uint64_t seed) {
  NEEDU();
  Ctx &c = *g_ctx;
  UNetPrecond &u = *fcg_unet(c);
  std::vector<float> h((size_t)u.nparam, 0.f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < u.a.n; l++) {
    // He-uniform for the ReLU layers; the linear output layer starts small
    const double fan_in = 27.0 * u.a.lay[l].ci;
    const double a = (l == u.a.out() ? 0.1 : 1.0) * std::sqrt(6.0 / fan_in);
    std::uniform_real_distribution<double> U(-a, a);
    for (int64_t q = 0; q < u.a.wcount(l); q++) h[(size_t)(u.woff[l] + q)] = (float)U(rng);
  }
  HIPCHK(hipMemcpyAsync(u.w, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, c.st));
  HIPCHK(hipMemsetAsync(u.m, 0, h.size() * sizeof(float), c.st));
  HIPCHK(hipMemsetAsync(u.v, 0, h.size() * sizeof(float), c.st));
  refresh_bf16(c, u);
  HIPCHK(hipStreamSynchronize(c.st));
  u.step = 0;
  u.ready = 1;
  return 0;
}

extern "C" int suli_unet_set_weights(const float *w, int64_t n) {
  NEEDU();
  Ctx &c = *g_ctx;
  UNetPrecond &u = *fcg_unet(c);
  if (n != u.nparam) {
    suli_set_error("suli_unet_set_weights: expected " + std::to_string(u.nparam) + " parameters, got " +
                   std::to_string(n));
    return 1;
  }
  HIPCHK(hipMemcpyAsync(u.w, w, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c.st));
  HIPCHK(hipMemsetAsync(u.m, 0, (size_t)n * sizeof(float), c.st));
  HIPCHK(hipMemsetAsync(u.v, 0, (size_t)n * sizeof(float), c.st));
  refresh_bf16(c, u);
  HIPCHK(hipStreamSynchronize(c.st));
  u.step = 0;
  u.ready = 1;
  return 0;
}

extern "C" int suli_unet_get_weights(float *w, int64_t n) {
  NEEDU();
  Ctx &c = *g_ctx;
  UNetPrecond &u = *fcg_unet(c);
  if (n != u.nparam) {
    suli_set_error("suli_unet_get_weights: expected " + std::to_string(u.nparam) + " parameters, got " +
                   std::to_string(n));
    return 1;
  }
  HIPCHK(hipMemcpyAsync(w, u.w, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  return 0;
}

extern "C" int suli_unet_capture(int max_samples, int every) {
  NEEDU();
  Ctx &c = *g_ctx;
  UNetPrecond &u = *fcg_unet(c);
  if (!entry_fits(c.g, u.entry, u.a.L)) {
    suli_set_error("suli_unet_capture: " + entry_msg(u));
    return 1;
  }
  if (max_samples < 0 || every < 0) {
    suli_set_error("suli_unet_capture: max_samples and every must be >= 0");
    return 1;
  }
  if (max_samples == 0) {  // stop capturing, keep the set
    u.ds_max = 0;
    return 0;
  }
  const Dims d = entry_dims(c.g, u.entry);
  const int64_t n = d.n(), nf = u.entry > 0 ? nfaces(d) : 0;
  if (!u.ds_cnt) {
    HIPCHK(hipMalloc((void **)&u.ds_cnt, sizeof(int)));
    c.allocs.push_back(u.ds_cnt);
    HIPCHK(hipMemsetAsync(u.ds_cnt, 0, sizeof(int), c.st));
  }
  if (max_samples > u.ds_cap) {  // grow, keeping the samples held
    int held = 0;
    if (dataset_count(c, u, &held)) return 1;
    held = std::min(held, u.ds_cap);
    float *a = nullptr, *b = nullptr, *f = nullptr;
    if (alloc_f(c, &a, (int64_t)max_samples * n) || alloc_f(c, &b, (int64_t)max_samples * n)) return 1;
    if (nf > 0 && alloc_f(c, &f, (int64_t)max_samples * nf)) return 1;
    if (held > 0) {
      HIPCHK(hipMemcpyAsync(a, u.ds_r, (size_t)held * n * sizeof(float), hipMemcpyDeviceToDevice, c.st));
      HIPCHK(hipMemcpyAsync(b, u.ds_rho, (size_t)held * n * sizeof(float), hipMemcpyDeviceToDevice, c.st));
      if (nf > 0) HIPCHK(hipMemcpyAsync(f, u.ds_f, (size_t)held * nf * sizeof(float), hipMemcpyDeviceToDevice, c.st));
      HIPCHK(hipStreamSynchronize(c.st));
    }
    // the old buffers stay in the context's allocation list until suli_finalize
    u.ds_r = a;
    u.ds_rho = b;
    u.ds_f = f;
    u.ds_cap = max_samples;
  }
  u.dsd = d;
  u.ds_entry = u.entry;
  u.ds_max = max_samples;
  u.ds_every = every;
  return 0;
}

extern "C" int suli_unet_set_entry(int level) {
  NEEDU();
  Ctx &c = *g_ctx;
  UNetPrecond &u = *fcg_unet(c);
  if (level < 0 || level > 4) {
    suli_set_error("suli_unet_set_entry: level must be in 0..4");
    return 1;
  }
  if (level != u.entry) {  // the held samples belong to the old level: drop them
    u.entry = level;
    u.ds_max = 0;
    u.ds_cap = 0;
    if (u.ds_cnt) HIPCHK(hipMemsetAsync(u.ds_cnt, 0, sizeof(int), c.st));
    HIPCHK(hipStreamSynchronize(c.st));
  }
  return 0;
}

extern "C" int suli_unet_set_fcg_m(int m) {
  NEEDU();
  if (m != 1 && m != 4) {
    suli_set_error("suli_unet_set_fcg_m: m must be 1 or 4");
    return 1;
  }
  fcg_unet(*g_ctx)->fcg_m = m;
  return 0;
}

extern "C" int suli_unet_entry(void) {
  if (!g_ctx || !g_ctx->fcg) return -1;
  return fcg_unet(*g_ctx)->entry;
}

extern "C" int suli_unet_clear(void) {
  NEEDU();
  Ctx &c = *g_ctx;
  UNetPrecond &u = *fcg_unet(c);
  u.ds_max = 0;
  if (u.ds_cnt) HIPCHK(hipMemsetAsync(u.ds_cnt, 0, sizeof(int), c.st));
  return 0;
}

extern "C" int suli_unet_dataset_size(void) {
  if (!g_ctx || !g_ctx->fcg) return -1;
  int n = 0;
  if (dataset_count(*g_ctx, *fcg_unet(*g_ctx), &n)) return -1;
  return n;
}

extern "C" int suli_unet_loss_grad(const double *r, const double *rho, double *loss, float *grad, float *z_out) {
  NEEDU();
  Ctx &c = *g_ctx;
  UNetPrecond &u = *fcg_unet(c);
  if (!u.ready) {
    suli_set_error("suli_unet_loss_grad: no weights loaded (suli_unet_init or suli_unet_set_weights)");
    return 1;
  }
  if (!dims_ok(c.g, u.a.L)) {
    suli_set_error("suli_unet_loss_grad: nx, ny and nz must be multiples of 2^(levels-1)");
    return 1;
  }
  const Dims d{c.g.nx, c.g.ny, c.g.nz};
  if (alloc_acts(c, u, d)) return 1;
  const int64_t n = (int64_t)d.nx * d.ny * d.nz;
  std::vector<float> hr((size_t)n), hrho((size_t)n);
  for (int64_t q = 0; q < n; q++) {
    hr[(size_t)q] = (float)r[q];
    hrho[(size_t)q] = (float)rho[q];
  }
  HIPCHK(hipMemcpyAsync(u.tmpf, hr.data(), n * sizeof(float), hipMemcpyHostToDevice, c.st));
  HIPCHK(hipMemcpyAsync(u.rho1, hrho.data(), n * sizeof(float), hipMemcpyHostToDevice, c.st));
  loss_and_dz(c, u, u.tmpf, u.rho1);
  backward(c, u);
  double dots[3];
  HIPCHK(hipMemcpyAsync(dots, u.red + kResDots, 3 * sizeof(double), hipMemcpyDeviceToHost, c.st));
  if (grad) HIPCHK(hipMemcpyAsync(grad, u.g, (size_t)u.nparam * sizeof(float), hipMemcpyDeviceToHost, c.st));
  if (z_out) HIPCHK(hipMemcpyAsync(z_out, u.z, n * sizeof(float), hipMemcpyDeviceToHost, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  HIPCHK(hipGetLastError());
  if (loss) *loss = 1.0 - dots[0] * dots[0] / (dots[2] * dots[1]);
  return 0;
}

extern "C" int suli_unet_dataset_get(int idx, float *r, float *rho, float *faces) {
  NEEDU();
  Ctx &c = *g_ctx;
  UNetPrecond &u = *fcg_unet(c);
  int nset = 0;
  if (dataset_count(c, u, &nset)) return 1;
  if (idx < 0 || idx >= nset) {
    suli_set_error("suli_unet_dataset_get: sample " + std::to_string(idx) + " of " + std::to_string(nset));
    return 1;
  }
  const int64_t n = u.dsd.n(), nf = u.ds_entry > 0 ? nfaces(u.dsd) : 0;
  if (r) HIPCHK(hipMemcpyAsync(r, u.ds_r + idx * n, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c.st));
  if (rho) HIPCHK(hipMemcpyAsync(rho, u.ds_rho + idx * n, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c.st));
  if (faces && nf > 0)
    HIPCHK(hipMemcpyAsync(faces, u.ds_f + idx * nf, (size_t)nf * sizeof(float), hipMemcpyDeviceToHost, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  return 0;
}

extern "C" int suli_unet_dataset_dims(int *nx, int *ny, int *nz, int *entry) {
  NEEDU();
  UNetPrecond &u = *fcg_unet(*g_ctx);
  *nx = u.dsd.nx;
  *ny = u.dsd.ny;
  *nz = u.dsd.nz;
  *entry = u.ds_entry;
  return 0;
}

extern "C" int suli_unet_sample_loss_grad(int idx, double *loss, float *grad, float *z_out) {
  NEEDU();
  Ctx &c = *g_ctx;
  UNetPrecond &u = *fcg_unet(c);
  if (!u.ready) {
    suli_set_error("suli_unet_sample_loss_grad: no weights loaded (suli_unet_init or suli_unet_set_weights)");
    return 1;
  }
  int nset = 0;
  if (dataset_count(c, u, &nset)) return 1;
  if (idx < 0 || idx >= nset) {
    suli_set_error("suli_unet_sample_loss_grad: sample " + std::to_string(idx) + " of " + std::to_string(nset));
    return 1;
  }
  if (alloc_acts(c, u, u.dsd)) return 1;
  const int64_t n = u.dsd.n(), nf = u.ds_entry > 0 ? nfaces(u.dsd) : 0;
  loss_and_dz(c, u, u.ds_r + idx * n, u.ds_rho + idx * n, nf > 0 ? u.ds_f + idx * nf : nullptr);
  backward(c, u);
  double dots[3];
  HIPCHK(hipMemcpyAsync(dots, u.red + kResDots, 3 * sizeof(double), hipMemcpyDeviceToHost, c.st));
  if (grad) HIPCHK(hipMemcpyAsync(grad, u.g, (size_t)u.nparam * sizeof(float), hipMemcpyDeviceToHost, c.st));
  if (z_out) HIPCHK(hipMemcpyAsync(z_out, u.z, n * sizeof(float), hipMemcpyDeviceToHost, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  HIPCHK(hipGetLastError());
  if (loss) *loss = 1.0 - dots[0] * dots[0] / (dots[2] * dots[1]);
  return 0;
}

extern "C" int suli_unet_train(int iters, double lr, uint64_t seed, double *loss) {
  NEEDU();
  Ctx &c = *g_ctx;
  UNetPrecond &u = *fcg_unet(c);
  if (!u.ready) {
    suli_set_error("suli_unet_train: no weights loaded (suli_unet_init or suli_unet_set_weights)");
    return 1;
  }
  int nset = 0;
  if (dataset_count(c, u, &nset)) return 1;
  if (nset <= 0) {
    suli_set_error("suli_unet_train: the training set is empty (suli_unet_capture, then solves)");
    return 1;
  }
  if (iters < 0 || !(lr > 0)) {
    suli_set_error("suli_unet_train: iters must be >= 0 and lr > 0");
    return 1;
  }
  if (alloc_acts(c, u, u.dsd)) return 1;
  const int64_t n = (int64_t)u.dsd.nx * u.dsd.ny * u.dsd.nz, nf = u.ds_entry > 0 ? nfaces(u.dsd) : 0;
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  // the last min(iters, 32) losses are read back for the report
  const int nrep = std::min(iters, 32);
  std::vector<double> rep((size_t)nrep * 3, 0.0);
  struct DevBuf {  // freed on every return path
    double *p = nullptr;
    ~DevBuf() {
      if (p) (void)hipFree(p);
    }
  } rep_buf;
  HIPCHK(hipMalloc((void **)&rep_buf.p, std::max<size_t>(1, rep.size()) * sizeof(double)));
  double *rep_d = rep_buf.p;
  for (int it = 0; it < iters; it++) {
    const int s = (int)(rng() % (uint64_t)nset);
    loss_and_dz(c, u, u.ds_r + (int64_t)s * n, u.ds_rho + (int64_t)s * n, nf > 0 ? u.ds_f + (int64_t)s * nf : nullptr);
    backward(c, u);
    u.step++;
    const double c1 = 1.0 - std::pow((double)b1, u.step), c2 = 1.0 - std::pow((double)b2, u.step);
    hipLaunchKernelGGL(k_adam, dim3(256), dim3(256), 0, c.st, u.w, u.g, u.m, u.v, u.nparam, (float)lr, b1, b2, eps,
                       (float)c1, (float)c2, 1.f);
    refresh_bf16(c, u);
    if (it >= iters - nrep)
      HIPCHK(hipMemcpyAsync(rep_d + 3 * (it - (iters - nrep)), u.red + kResDots, 3 * sizeof(double),
                            hipMemcpyDeviceToDevice, c.st));
  }
  if (nrep > 0) HIPCHK(hipMemcpyAsync(rep.data(), rep_d, rep.size() * sizeof(double), hipMemcpyDeviceToHost, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  HIPCHK(hipGetLastError());
  double m = 0;
  for (int q = 0; q < nrep; q++) m += 1.0 - rep[3 * q] * rep[3 * q] / (rep[3 * q + 2] * rep[3 * q + 1]);
  if (loss) *loss = nrep > 0 ? m / nrep : 0.0;
  return 0;
}
