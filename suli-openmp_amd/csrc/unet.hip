// Learned preconditioner for the projection's flexible CG: a 3-level 3-D
// convolutional U-Net (16/32/64 channels, 181k parameters) applied to the
// scaled residual r_s and the log-density, z = ||r|| Net(r/||r||, log rho).
// Training and inference use this library's own kernels (unet_kernels.hpp):
// BF16 implicit-GEMM convolutions on MFMA with FP32 accumulation for
// forward, data gradient and weight gradient, FP32 master weights, Adam.
//
// Architecture (channels; c1a and c6 are FP32 direct kernels):
//   x0(2) -c1a-> a1(16) -c1b-> s1(16) -pool-> p1(16) -c2a-> a2(32) -c2b-> s2(32)
//   -pool-> p2(32) -c3a-> a3(64) -c3b-> b3(32) -up,+s2-> u2(32) -c4-> a4(16)
//   -up,+s1-> u1(16) -c5-> a5(16) -c6-> z(1)
// 3x3x3 convolutions with bias and ReLU except c6 (linear).
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

struct LayerDesc {
  int ci, co;
};
// c1a c1b c2a c2b c3a c3b c4 c5 c6
const LayerDesc kLayers[9] = {{2, 16}, {16, 16}, {16, 32}, {32, 32}, {32, 64}, {64, 32}, {32, 16}, {16, 16}, {16, 1}};

inline int64_t wcount(int l) { return 27LL * kLayers[l].ci * kLayers[l].co; }

inline dim3 g1(int64_t n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

}  // namespace

struct UNetPrecond {
  int ready = 0;
  int64_t nparam = 0;
  int64_t woff[9], boff[9];
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *wf[9] = {}, *wt[9] = {};
  // activations / gradients at the allocated size
  Dims d1{0, 0, 0};
  int64_t cap1 = 0;  // voxels allocated at level 1
  float *x0 = nullptr, *z = nullptr, *dz = nullptr, *part = nullptr, *tmpf = nullptr;
  uint16_t *a1, *s1, *p1, *a2, *s2, *p2, *a3, *b3, *u2, *a4, *u1, *a5;
  uint16_t *ga, *gb, *gc, *gd;  // gradient scratch (level-1 sized, 32 channels)
  double *red = nullptr;        // reduction partials / results
  // training operator (compact grid) and scratch
  float *cx = nullptr, *cy = nullptr, *cz = nullptr, *s = nullptr, *t = nullptr, *rho1 = nullptr;
  // dataset captured from the solver (compact grid)
  Dims dsd{0, 0, 0};
  int ds_n = 0, ds_max = 0, ds_cap = 0;
  float *ds_r = nullptr, *ds_rho = nullptr;
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

// compact copies of the residual and density for the dataset
template <typename T>
__global__ void k_capture(Grid g, Dims d, const float *__restrict__ r, const T *__restrict__ rho,
                          float *__restrict__ rs, float *__restrict__ rhos) {
  const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (v >= d.n()) return;
  const int i = (int)(v % d.nx), j = (int)((v / d.nx) % d.ny), k = (int)(v / ((int64_t)d.nx * d.ny));
  const int64_t c = g.idx(i + 1, j + 1, k + 1);
  rs[v] = r[c];
  rhos[v] = (float)rho[c];
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
  const double a = -2.0 * rs / (rr * ss), b = 2.0 * rs * rs / (rr * ss * ss);
  t[v] = (float)(a * r[v] + b * s[v]);
}

// ---------------------------------------------------------------------------
// host side

namespace {

constexpr int kRedPart = 1 << 16;  // reduction partials; results follow
constexpr int kResNorm = kRedPart, kResDots = kRedPart + 8;

template <int CIN, int COUT, int FLAGS>
void conv(hipStream_t st, Dims d, const uint16_t *in, const uint16_t *w, const float *bias, const uint16_t *gate,
          uint16_t *out, const int *active = nullptr) {
  hipLaunchKernelGGL((k_conv_mfma<CIN, COUT, FLAGS>), dim3((d.nx + 63) / 64, d.ny, d.nz), dim3(256), 0, st, d, in, w,
                     bias, gate, out, active);
}

// chunking of the weight-gradient reductions: at most 64 chunks of whole
// 32-voxel groups
inline int wchunk(int64_t n) { return (int)(((n + 63) / 64 + 31) / 32 * 32); }
inline int wnchunk(int64_t n) { return (int)((n + wchunk(n) - 1) / wchunk(n)); }

template <int CIN, int COUT>
void wgrad(UNetPrecond &u, hipStream_t st, Dims d, const uint16_t *dy, const uint16_t *x, int layer) {
  const int64_t n = (int64_t)d.nx * d.ny * d.nz;
  const int chunk = wchunk(n), nch = wnchunk(n);
  hipLaunchKernelGGL((k_wgrad_mfma<CIN, COUT>), dim3(27, nch), dim3(256), 0, st, d, dy, x, u.part, chunk);
  const int64_t len = wcount(layer);
  hipLaunchKernelGGL(k_sum_chunks, g1(len), dim3(256), 0, st, u.part, nch, len, u.g + u.woff[layer]);
  hipLaunchKernelGGL((k_bias_grad<uint16_t>), dim3(nch), dim3(256), 0, st, d, COUT, dy, u.part, chunk);
  hipLaunchKernelGGL(k_sum_chunks, g1(COUT), dim3(256), 0, st, u.part, nch, (int64_t)COUT, u.g + u.boff[layer]);
}

void finish_sum(hipStream_t st, double *part, int nb, double *res, int nv) {
  if (nv == 1) hipLaunchKernelGGL((k_finish<1, false, EpiNone>), dim3(1), dim3(256), 0, st, part, nb, res, EpiNone());
  else hipLaunchKernelGGL((k_finish<3, false, EpiNone>), dim3(1), dim3(256), 0, st, part, nb, res, EpiNone());
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

// activations and training scratch for a level-1 grid d (dimensions
// multiples of 4); grown on demand, never shrunk
int alloc_acts(Ctx &c, UNetPrecond &u, Dims d) {
  const int64_t n1 = (int64_t)d.nx * d.ny * d.nz;
  if (n1 <= u.cap1) {
    u.d1 = d;
    return 0;
  }
  const int64_t n2 = n1 / 8, n3 = n1 / 64;
  if (alloc_f(c, &u.x0, 2 * n1) || alloc_f(c, &u.z, n1) || alloc_f(c, &u.dz, n1) || alloc_f(c, &u.tmpf, n1)) return 1;
  uint16_t **l1[] = {&u.a1, &u.s1, &u.u1, &u.a5};
  for (auto q : l1)
    if (alloc_h(c, q, 16 * n1)) return 1;
  uint16_t **l2[] = {&u.p1, &u.a4};
  for (auto q : l2)
    if (alloc_h(c, q, 16 * n2)) return 1;
  uint16_t **l2b[] = {&u.a2, &u.s2, &u.u2};
  for (auto q : l2b)
    if (alloc_h(c, q, 32 * n2)) return 1;
  if (alloc_h(c, &u.p2, 32 * n3) || alloc_h(c, &u.a3, 64 * n3) || alloc_h(c, &u.b3, 32 * n3)) return 1;
  // gradient scratch: the largest gradient tensor is 16 channels at level 1
  uint16_t **gs[] = {&u.ga, &u.gb, &u.gc, &u.gd};
  for (auto q : gs)
    if (alloc_h(c, q, 16 * n1)) return 1;
  // face arrays: (n+1) faces along one direction, at most 2 n1 entries
  if (alloc_f(c, &u.cx, 2 * n1) || alloc_f(c, &u.cy, 2 * n1) || alloc_f(c, &u.cz, 2 * n1) || alloc_f(c, &u.s, n1) ||
      alloc_f(c, &u.t, n1) || alloc_f(c, &u.rho1, n1))
    return 1;
  u.cap1 = n1;
  u.d1 = d;
  return 0;
}

// forward pass from x0 (already prepared); z_net to u.z
void forward(Ctx &c, UNetPrecond &u, const int *active) {
  hipStream_t st = c.st;
  const Dims d1 = u.d1, d2{d1.nx / 2, d1.ny / 2, d1.nz / 2}, d3{d1.nx / 4, d1.ny / 4, d1.nz / 4};
  const int64_t n1 = (int64_t)d1.nx * d1.ny * d1.nz, n2 = n1 / 8, n3 = n1 / 64;
  const float *W = u.w;
  hipLaunchKernelGGL(k_conv_in, g1(n1), dim3(256), 0, st, d1, u.x0, W + u.woff[0], W + u.boff[0], u.a1, active);
  conv<16, 16, CONV_BIAS | CONV_RELU>(st, d1, u.a1, u.wf[1], W + u.boff[1], nullptr, u.s1, active);
  hipLaunchKernelGGL(k_pool, g1(n2 * 16), dim3(256), 0, st, d1, 16, u.s1, u.p1, active);
  conv<16, 32, CONV_BIAS | CONV_RELU>(st, d2, u.p1, u.wf[2], W + u.boff[2], nullptr, u.a2, active);
  conv<32, 32, CONV_BIAS | CONV_RELU>(st, d2, u.a2, u.wf[3], W + u.boff[3], nullptr, u.s2, active);
  hipLaunchKernelGGL(k_pool, g1(n3 * 32), dim3(256), 0, st, d2, 32, u.s2, u.p2, active);
  conv<32, 64, CONV_BIAS | CONV_RELU>(st, d3, u.p2, u.wf[4], W + u.boff[4], nullptr, u.a3, active);
  conv<64, 32, CONV_BIAS | CONV_RELU>(st, d3, u.a3, u.wf[5], W + u.boff[5], nullptr, u.b3, active);
  hipLaunchKernelGGL(k_up_add, g1(n2 * 32), dim3(256), 0, st, d2, 32, u.b3, u.s2, u.u2, active);
  conv<32, 16, CONV_BIAS | CONV_RELU>(st, d2, u.u2, u.wf[6], W + u.boff[6], nullptr, u.a4, active);
  hipLaunchKernelGGL(k_up_add, g1(n1 * 16), dim3(256), 0, st, d1, 16, u.a4, u.s1, u.u1, active);
  conv<16, 16, CONV_BIAS | CONV_RELU>(st, d1, u.u1, u.wf[7], W + u.boff[7], nullptr, u.a5, active);
  hipLaunchKernelGGL(k_conv_out, g1(n1), dim3(256), 0, st, d1, u.a5, W + u.woff[8], W + u.boff[8], u.z, active);
}

// backward pass from u.dz; gradients into u.g (overwritten)
void backward(Ctx &c, UNetPrecond &u) {
  hipStream_t st = c.st;
  const Dims d1 = u.d1, d2{d1.nx / 2, d1.ny / 2, d1.nz / 2}, d3{d1.nx / 4, d1.ny / 4, d1.nz / 4};
  const int64_t n1 = (int64_t)d1.nx * d1.ny * d1.nz, n2 = n1 / 8, n3 = n1 / 64;
  const int chunk = wchunk(n1), nch = wnchunk(n1);
  // c6: z = conv(a5) + b
  hipLaunchKernelGGL((k_wgrad_small<16, 1, float, uint16_t>), dim3(27, nch), dim3(256), 0, st, d1, u.dz, u.a5, u.part,
                     chunk);
  hipLaunchKernelGGL(k_sum_chunks, g1(wcount(8)), dim3(256), 0, st, u.part, nch, wcount(8), u.g + u.woff[8]);
  hipLaunchKernelGGL((k_bias_grad<float>), dim3(nch), dim3(256), 0, st, d1, 1, u.dz, u.part, chunk);
  hipLaunchKernelGGL(k_sum_chunks, g1(1), dim3(256), 0, st, u.part, nch, (int64_t)1, u.g + u.boff[8]);
  uint16_t *da5 = u.ga;
  hipLaunchKernelGGL(k_conv_out_bwd, g1(n1), dim3(256), 0, st, d1, u.dz, u.w + u.woff[8], u.a5, da5);
  // c5: a5 = relu(conv(u1))
  wgrad<16, 16>(u, st, d1, da5, u.u1, 7);
  uint16_t *du1 = u.gb;
  conv<16, 16, 0>(st, d1, da5, u.wt[7], nullptr, nullptr, du1);
  // u1 = up(a4) + s1
  uint16_t *da4 = u.gc;
  hipLaunchKernelGGL(k_down_sum, g1(n2 * 16), dim3(256), 0, st, d1, 16, du1, u.a4, da4);
  // c4: a4 = relu(conv(u2))
  wgrad<32, 16>(u, st, d2, da4, u.u2, 6);
  uint16_t *du2 = u.gd;
  conv<16, 32, 0>(st, d2, da4, u.wt[6], nullptr, nullptr, du2);
  // u2 = up(b3) + s2
  uint16_t *db3 = u.ga;
  hipLaunchKernelGGL(k_down_sum, g1(n3 * 32), dim3(256), 0, st, d2, 32, du2, u.b3, db3);
  // c3b: b3 = relu(conv(a3))
  wgrad<64, 32>(u, st, d3, db3, u.a3, 5);
  uint16_t *da3 = u.gc;
  conv<32, 64, CONV_MASK>(st, d3, db3, u.wt[5], nullptr, u.a3, da3);
  // c3a: a3 = relu(conv(p2))
  wgrad<32, 64>(u, st, d3, da3, u.p2, 4);
  uint16_t *dp2 = u.ga;
  conv<64, 32, 0>(st, d3, da3, u.wt[4], nullptr, nullptr, dp2);
  // p2 = pool(s2) and s2 feeds u2: ds2 = (du2 + up(dp2)/8) where s2 > 0
  uint16_t *ds2 = u.gc;
  hipLaunchKernelGGL(k_add_upavg_mask, g1(n2 * 32), dim3(256), 0, st, d2, 32, du2, dp2, u.s2, ds2);
  // c2b: s2 = relu(conv(a2))
  wgrad<32, 32>(u, st, d2, ds2, u.a2, 3);
  uint16_t *da2 = u.gd;
  conv<32, 32, CONV_MASK>(st, d2, ds2, u.wt[3], nullptr, u.a2, da2);
  // c2a: a2 = relu(conv(p1))
  wgrad<16, 32>(u, st, d2, da2, u.p1, 2);
  uint16_t *dp1 = u.ga;
  conv<32, 16, 0>(st, d2, da2, u.wt[2], nullptr, nullptr, dp1);
  // p1 = pool(s1) and s1 feeds u1: ds1 = (du1 + up(dp1)/8) where s1 > 0
  uint16_t *ds1 = u.gc;
  hipLaunchKernelGGL(k_add_upavg_mask, g1(n1 * 16), dim3(256), 0, st, d1, 16, du1, dp1, u.s1, ds1);
  // c1b: s1 = relu(conv(a1))
  wgrad<16, 16>(u, st, d1, ds1, u.a1, 1);
  uint16_t *da1 = u.gd;
  conv<16, 16, CONV_MASK>(st, d1, ds1, u.wt[1], nullptr, u.a1, da1);
  // c1a: a1 = relu(conv(x0) + b), FP32 input
  hipLaunchKernelGGL((k_wgrad_small<2, 16, uint16_t, float>), dim3(27, nch), dim3(256), 0, st, d1, da1, u.x0, u.part,
                     chunk);
  hipLaunchKernelGGL(k_sum_chunks, g1(wcount(0)), dim3(256), 0, st, u.part, nch, wcount(0), u.g + u.woff[0]);
  hipLaunchKernelGGL((k_bias_grad<uint16_t>), dim3(nch), dim3(256), 0, st, d1, 16, da1, u.part, chunk);
  hipLaunchKernelGGL(k_sum_chunks, g1(16), dim3(256), 0, st, u.part, nch, (int64_t)16, u.g + u.boff[0]);
}

void refresh_bf16(Ctx &c, UNetPrecond &u) {
  for (int l = 1; l <= 7; l++)
    hipLaunchKernelGGL(k_weights_bf16, g1(wcount(l)), dim3(256), 0, c.st, u.w + u.woff[l], kLayers[l].co,
                       kLayers[l].ci, u.wf[l], u.wt[l]);
}

// Loss terms and dL/dz for a compact (r, rho) pair at dims u.d1: dots
// (r.s, s.s, r.r) with s = A_s z land at red[kResDots..+2], dL/dz in u.dz.
void loss_and_dz(Ctx &c, UNetPrecond &u, const float *r, const float *rho) {
  hipStream_t st = c.st;
  const Dims d = u.d1;
  const int64_t n = (int64_t)d.nx * d.ny * d.nz;
  const int64_t nf = (int64_t)(d.nx + 1) * d.ny * d.nz + (int64_t)d.nx * (d.ny + 1) * d.nz +
                     (int64_t)d.nx * d.ny * (d.nz + 1);
  const float mx = (float)(1.0 / (c.g.dx * c.g.dx)), my = (float)(1.0 / (c.g.dy * c.g.dy)),
              mz = (float)(1.0 / (c.g.dz * c.g.dz));
  hipLaunchKernelGGL(k_train_faces, g1(nf), dim3(256), 0, st, d, mx, my, mz, rho, u.cx, u.cy, u.cz);
  // |r|^2 for the input normalisation
  hipLaunchKernelGGL(k_train_dots, dim3(512), dim3(256), 0, st, d, r, r, u.red);
  finish_sum(st, u.red, 512, u.red + kResNorm, 3);
  hipLaunchKernelGGL(k_prep_compact, g1(n), dim3(256), 0, st, d, r, rho, u.red + kResNorm + 2, u.lr1, u.lr2, u.x0);
  forward(c, u, nullptr);
  hipLaunchKernelGGL(k_train_As, g1(n), dim3(256), 0, st, d, u.cx, u.cy, u.cz, u.z, u.s);
  hipLaunchKernelGGL(k_train_dots, dim3(512), dim3(256), 0, st, d, r, u.s, u.red);
  finish_sum(st, u.red, 512, u.red + kResDots, 3);
  hipLaunchKernelGGL(k_train_dLds, g1(n), dim3(256), 0, st, d, r, u.s, u.red + kResDots, u.t);
  hipLaunchKernelGGL(k_train_As, g1(n), dim3(256), 0, st, d, u.cx, u.cy, u.cz, u.t, u.dz);
}

bool dims_ok(const Grid &g) { return g.nx % 4 == 0 && g.ny % 4 == 0 && g.nz % 4 == 0; }

}  // namespace

// ---------------------------------------------------------------------------
int unet_create(Ctx &c, UNetPrecond **out) {
  UNetPrecond *u = new UNetPrecond();
  *out = u;
  int64_t off = 0;
  for (int l = 0; l < 9; l++) {
    u->woff[l] = off;
    off += wcount(l);
    u->boff[l] = off;
    off += kLayers[l].co;
  }
  u->nparam = off;
  if (alloc_f(c, &u->w, off) || alloc_f(c, &u->g, off) || alloc_f(c, &u->m, off) || alloc_f(c, &u->v, off)) return 1;
  for (int l = 1; l <= 7; l++)
    if (alloc_h(c, &u->wf[l], wcount(l)) || alloc_h(c, &u->wt[l], wcount(l))) return 1;
  HIPCHK(hipMalloc((void **)&u->red, (kRedPart + 64) * sizeof(double)));
  c.allocs.push_back(u->red);
  u->lr1 = 0.5 * (std::log(c.ph.rho_f1) + std::log(c.ph.rho_f2));
  u->lr2 = std::max(0.5 * std::fabs(std::log(c.ph.rho_f2) - std::log(c.ph.rho_f1)), 1e-6);
  // weight-gradient partials: up to 64 chunks x 27 taps x the largest layer (64x32)
  if (alloc_f(c, &u->part, 64LL * 27 * 64 * 32)) return 1;
  return 0;
}

void unet_destroy(Ctx &c, UNetPrecond *u) {
  (void)c;
  delete u;
}

bool unet_ready(const UNetPrecond *u) { return u && u->ready; }

int unet_prepare(Ctx &c, UNetPrecond *u) {
  if (!u || !u->ready) {
    suli_set_error("learned preconditioner selected but no weights are loaded (suli_unet_init or "
                   "suli_unet_set_weights)");
    return 1;
  }
  if (!dims_ok(c.g)) {
    suli_set_error("learned preconditioner: nx, ny and nz must be multiples of 4");
    return 1;
  }
  return alloc_acts(c, *u, Dims{c.g.nx, c.g.ny, c.g.nz});
}

void unet_apply(Ctx &c, UNetPrecond *u, const float *r, float *z, const int *active) {
  const Grid &g = c.g;
  const Dims d{g.nx, g.ny, g.nz};
  const int64_t n = (int64_t)d.nx * d.ny * d.nz;
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

void unet_capture(Ctx &c, UNetPrecond *u, const float *r) {
  if (!u || u->ds_n >= u->ds_max) return;
  const Grid &g = c.g;
  const Dims d{g.nx, g.ny, g.nz};
  const int64_t n = (int64_t)d.nx * d.ny * d.nz;
  if ((int64_t)u->dsd.nx * u->dsd.ny * u->dsd.nz != n) return;
  float *rs = u->ds_r + (int64_t)u->ds_n * n, *rh = u->ds_rho + (int64_t)u->ds_n * n;
  if (c.prec == SULI_PREC_FP64)
    hipLaunchKernelGGL((k_capture<double>), g1(n), dim3(256), 0, c.st, g, d, r, (const double *)c.rho, rs, rh);
  else
    hipLaunchKernelGGL((k_capture<float>), g1(n), dim3(256), 0, c.st, g, d, r, (const float *)c.rho, rs, rh);
  u->ds_n++;
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
  for (int l = 0; l < 9; l++) {
    // He-uniform for the ReLU layers; the linear output layer starts small
    const double fan_in = 27.0 * kLayers[l].ci;
    const double a = (l == 8 ? 0.1 : 1.0) * std::sqrt(6.0 / fan_in);
    std::uniform_real_distribution<double> U(-a, a);
    for (int64_t q = 0; q < wcount(l); q++) h[(size_t)(u.woff[l] + q)] = (float)U(rng);
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

extern "C" int suli_unet_capture(int max_samples) {
  NEEDU();
  Ctx &c = *g_ctx;
  UNetPrecond &u = *fcg_unet(c);
  if (!dims_ok(c.g)) {
    suli_set_error("suli_unet_capture: nx, ny and nz must be multiples of 4");
    return 1;
  }
  if (max_samples < 0) {
    suli_set_error("suli_unet_capture: max_samples must be >= 0");
    return 1;
  }
  if (max_samples == 0) {  // stop capturing, keep the set
    u.ds_max = u.ds_n;
    return 0;
  }
  const Dims d{c.g.nx, c.g.ny, c.g.nz};
  const int64_t n = (int64_t)d.nx * d.ny * d.nz;
  if (max_samples > u.ds_cap) {
    float *a = nullptr, *b = nullptr;
    if (alloc_f(c, &a, (int64_t)max_samples * n) || alloc_f(c, &b, (int64_t)max_samples * n)) return 1;
    u.ds_r = a;
    u.ds_rho = b;
    u.ds_cap = max_samples;
  }
  u.dsd = d;
  u.ds_n = 0;
  u.ds_max = max_samples;
  return 0;
}

extern "C" int suli_unet_dataset_size(void) {
  if (!g_ctx || !g_ctx->fcg) return -1;
  return fcg_unet(*g_ctx)->ds_n;
}

extern "C" int suli_unet_loss_grad(const double *r, const double *rho, double *loss, float *grad, float *z_out) {
  NEEDU();
  Ctx &c = *g_ctx;
  UNetPrecond &u = *fcg_unet(c);
  if (!u.ready) {
    suli_set_error("suli_unet_loss_grad: no weights loaded (suli_unet_init or suli_unet_set_weights)");
    return 1;
  }
  if (!dims_ok(c.g)) {
    suli_set_error("suli_unet_loss_grad: nx, ny and nz must be multiples of 4");
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

extern "C" int suli_unet_train(int iters, double lr, uint64_t seed, double *loss) {
  NEEDU();
  Ctx &c = *g_ctx;
  UNetPrecond &u = *fcg_unet(c);
  if (!u.ready) {
    suli_set_error("suli_unet_train: no weights loaded (suli_unet_init or suli_unet_set_weights)");
    return 1;
  }
  if (u.ds_n <= 0) {
    suli_set_error("suli_unet_train: the training set is empty (suli_unet_capture, then solves)");
    return 1;
  }
  if (iters < 0 || !(lr > 0)) {
    suli_set_error("suli_unet_train: iters must be >= 0 and lr > 0");
    return 1;
  }
  if (alloc_acts(c, u, u.dsd)) return 1;
  const int64_t n = (int64_t)u.dsd.nx * u.dsd.ny * u.dsd.nz;
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  // the last min(iters, 32) losses are read back for the report
  const int nrep = std::min(iters, 32);
  std::vector<double> rep((size_t)nrep * 3, 0.0);
  double *rep_d = nullptr;
  HIPCHK(hipMalloc((void **)&rep_d, std::max<size_t>(1, rep.size()) * sizeof(double)));
  for (int it = 0; it < iters; it++) {
    const int s = (int)(rng() % (uint64_t)u.ds_n);
    loss_and_dz(c, u, u.ds_r + (int64_t)s * n, u.ds_rho + (int64_t)s * n);
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
  (void)hipFree(rep_d);
  HIPCHK(hipGetLastError());
  double m = 0;
  for (int q = 0; q < nrep; q++) m += 1.0 - rep[3 * q] * rep[3 * q] / (rep[3 * q + 2] * rep[3 * q + 1]);
  if (loss) *loss = nrep > 0 ? m / nrep : 0.0;
  return 0;
}
