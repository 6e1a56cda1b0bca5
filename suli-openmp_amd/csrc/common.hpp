// Shared device-side definitions of the SuLi-IC MI355X library.
//
// Field layout in HBM ("frame"): every 3-D field of the time step - the
// staggered velocities with their ghost layers, the pressure, cell arrays and
// face arrays - lives in one common frame indexed with the reference's own
// Fortran indices i in [0,nx+2], j in [0,ny+2], k in [0,nz+2] (the union of
// the extents declared in 0_modules.f90:92-133), i fastest:
//
//     idx(i,j,k) = off + i + sx*(j + sy*k)
//
// sx is nx+3 rounded up to 32 elements and off = 31, so the first interior
// point of every row (i = 1) sits on a 128-byte boundary for FP32 fields and
// every wavefront of 64 consecutive i reads whole cache lines.  Because all
// fields share one index map, a fused kernel computes one address per thread
// and offsets it by the same strides (1, sx, sx*sy) for every field.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SULI_LDS_RED 16  // max wavefronts per block in the reductions below

// offset of an element in a frame field
typedef uint32_t fidx;

// Stage updates of tempo (11_diversos.f90:10-228) done by the momentum kernel.
enum MomUpdate {
  MOM_RK_FIRST = 1,  // RK2/RK3 tt=1: u0 = u; [fu0 = F]; u = u + dt F
  MOM_FROM_U0 = 2,   // RK2 tt=2: u = u0 + dt F
  MOM_EULER = 3,     // Euler and the first AB2 step: u = u + dt F; [fu0 = F]
  MOM_RK3_2 = 4,     // fu1 = F; u = u0 - dt fu0 + (2 dt) F
  MOM_RK3_3 = 5,     // u = u0 + (dt/6) (fu0 + 4 fu1 + F)
  MOM_AB2 = 6        // u = u + dt 0.5 (3 F - fu0); fu0 = F
};

struct Grid {
  int nx, ny, nz;
  int sx, sy;        // pitches in i and j
  uint32_t sz;       // sx*sy (k stride)
  uint32_t off;      // offset of (0,0,0)
  int64_t total;     // elements per frame field (below 2^29: checked at set-up)
  double dx, dy, dz;
  int ccx0, ccxf, ccy0, ccyf, ccz0, cczf;
  // 1: the momentum predictor's inner box in LDS tiles (k_momentum_tile),
  // 0: every stencil value read from memory; the same values either way
  int fast;
  // Frame offsets are 32-bit unsigned: with a uniform base pointer the loads
  // take the scalar-base + 32-bit-offset form, one VGPR per address instead
  // of two.  Neighbour offsets may wrap below zero in intermediate sums; the
  // final offset of every access lies in the frame.
  __host__ __device__ __forceinline__ fidx idx(int i, int j, int k) const {
    return off + (fidx)i + (fidx)sx * ((fidx)j + (fidx)sy * (fidx)k);
  }
};

// Boundary planes of the prescribed-velocity walls (wall code 3; bxy0 also
// serves code 4 at x = 0), in the declaration order of 0_modules.f90:96-120.
// All 24 live in one buffer of the field precision; plane n holds its
// module extents (0:e1-1, 0:e2-1) in Fortran order at off[n].
enum BPlane {
  BXX0, BXX1, BXY0, BXZ0, BXXF, BXXF1, BXYF, BXZF,
  BYX0, BYY0, BYY1, BYZ0, BYXF, BYYF, BYYF1, BYZF,
  BZX0, BZY0, BZZ0, BZZ1, BZXF, BZYF, BZZF, BZZF1, BP_N
};
struct BPlanes {
  int64_t off[BP_N + 1];  // element offsets; off[BP_N] = total
  int e1[BP_N], e2[BP_N]; // extents along the plane's two axes
  __host__ __device__ __forceinline__ int64_t at(int n, int a, int b) const { return off[n] + a + (int64_t)e1[n] * b; }
};
// wall axis (0 x, 1 y, 2 z) and velocity component (0 u, 1 v, 2 w) of plane n
__host__ __device__ __forceinline__ int bplane_wall(int n) { return n / 8; }
__host__ __device__ __forceinline__ int bplane_comp(int n) {
  // per wall: the normal component (4 planes: ghost, first face, last face,
  // outer ghost) and the two tangential components at each end
  constexpr int c[BP_N] = {0, 0, 1, 2, 0, 0, 1, 2, 0, 1, 1, 2, 0, 1, 1, 2, 0, 1, 2, 2, 0, 1, 2, 2};
  return c[n];
}
inline BPlanes bplanes_layout(int nx, int ny, int nz) {
  BPlanes L{};
  const int n[3] = {nx, ny, nz};
  int64_t o = 0;
  for (int q = 0; q < BP_N; q++) {
    const int wall = bplane_wall(q), comp = bplane_comp(q);
    const int a0 = wall == 0 ? 1 : 0, a1 = wall == 2 ? 1 : 2;  // the two in-plane axes
    L.e1[q] = n[a0] + (comp == a0 ? 3 : 2);
    L.e2[q] = n[a1] + (comp == a1 ? 3 : 2);
    L.off[q] = o;
    o += (int64_t)L.e1[q] * L.e2[q];
  }
  L.off[BP_N] = o;
  return L;
}

// A frame field read through 32-bit byte offsets: base + zext(c * sizeof(T)).
// Frame fields hold fewer than 2^29 elements (checked at set-up), so the byte
// offset fits in 32 bits, and the load takes the scalar-base + 32-bit-offset
// form (one VGPR per address instead of a 64-bit address pair).
template <typename T>
struct Fld {
  const T *__restrict__ p;
  __device__ __forceinline__ T operator[](fidx c) const {
    return *(const T *)((const char *)p + (uint32_t)(c * (uint32_t)sizeof(T)));
  }
};

// Fluid properties and model constants (14_level_set.f90:25-31,
// 0_modules.f90:141-143,181,197-198).
struct Phys {
  double rho_f1, rho_f2, mi_f1, mi_f2, sigma, alpha1;
  double csmag, gx, gz, chezy;
  double c23;   // (dx*dy*dz)**(2/3), evaluated on the host (5_viscosidade.f90:94)
  double dx1;   // max(dx,dy,dz)
  double dt1;   // pseudo time step of the reinitialisation, 0.1*dx1
  int m_turb;
};

// Thread layout of the stencil kernels: 64 lanes along i (one wavefront per
// row segment), 4 rows along j, one k plane per block column.
#define SULI_BX 64
#define SULI_BY 4

template <typename T>
__device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <typename T>
__device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
template <typename T>
__device__ __forceinline__ T tabs(T a) { return a < T(0) ? -a : a; }

// ---------------------------------------------------------------------------
// Deterministic block reductions in FP64.
//
// A kernel that produces a global sum writes one partial per block (no
// atomics); a one-work-group finish kernel (k_finish below) then adds the
// partials in block order and stores the result, so the sum is the same on
// every run.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double w = __shfl_down(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

// Work-group order of the grid stencils.  The dispatcher places
// work-group b on XCD b mod 8 and each XCD has its own L2; in grid order the
// block rows of one XCD are every fourth row, so the rows above and below a
// block (the j-neighbours of its stencil) sit in other XCDs' caches.  The
// blocks are renumbered so that XCD x owns a contiguous band of block rows
// through all planes, walked plane by plane: its j- and k-neighbours are its
// own.  (Identity when the block rows do not split into 8 bands.)
__device__ __forceinline__ void xcd_band_block(int &bx, int &by, int &bz) {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  if (gy % 8u) {
    bx = blockIdx.x;
    by = blockIdx.y;
    bz = blockIdx.z;
    return;
  }
  const unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned band = gy / 8u, t = b / 8u;
  bx = (int)(t % gx);
  by = (int)((b % 8u) * band + (t / gx) % band);
  bz = (int)(t / (gx * band));
}

// Sum (or max) over the block of nv values per thread; result valid in thread 0.
template <int NV, bool MAX = false>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double *lds /* NV*SULI_LDS_RED */) {
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int nt = blockDim.x * blockDim.y * blockDim.z;
  const int lane = tid & 63, wid = tid >> 6, nw = (nt + 63) >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double r = MAX ? wave_max(v[q]) : wave_sum(v[q]);
    if (lane == 0) lds[q * SULI_LDS_RED + wid] = r;
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) {
      double r = lds[q * SULI_LDS_RED];
      for (int w = 1; w < nw; w++) {
        double x = lds[q * SULI_LDS_RED + w];
        r = MAX ? (x > r ? x : r) : r + x;
      }
      v[q] = r;
    }
  }
}

// Multigrid face conductances: FP32 value of an FP32 or BF16 conductance, and
// BF16 rounding to nearest even (finite values).
__device__ __forceinline__ float mgc(float v) { return v; }
__device__ __forceinline__ float mgc(uint16_t v) { return __uint_as_float((uint32_t)v << 16); }
__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// Reduction slots: per-block partials (NV values per block, value-major) and
// the NV results.  A producer kernel writes its block partials with
// block_partial (no fence, no atomic); a one-block finish kernel launched
// right after it on the same stream adds the partials in block order, so the
// sum is the same on every run.  (A last-block-done scheme needs a
// device-scope fence per block, which on this multi-die part writes the L2
// back and costs far more than the extra launch.)
struct RedSlot {
  double *partials;
  double *result;
};

template <int NV>
__device__ __forceinline__ void block_partial(const double (&v)[NV], double *partials) {
  if (threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0) {
    const unsigned int nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
    for (int q = 0; q < NV; q++) partials[(size_t)q * nb + b] = v[q];
  }
}

// After an epilogue (or in place of one that is not enabled) thread 0 may
// also publish loop-control state to the host: epi_publish overloads for the
// epilogue types that do; none by default.
template <typename Epi>
__device__ __forceinline__ void epi_publish(const Epi &) {}

// Epilogue of a plain reduction: nothing beyond storing the result.
struct EpiNone {
  __device__ bool enabled() const { return true; }
  __device__ void operator()(const double *) const {}
};

// Sum (or max) of the nb block partials of NV values in a fixed order; thread
// 0 stores result[0..NV) and runs the epilogue (loop control of the solvers).
// The partials were written by blocks on every XCD and come from memory: a
// block of SULI_FINISH_THREADS threads with all of a thread's loads issued
// before its adds keeps the serial load latency to one or two round trips.
#define SULI_FINISH_THREADS 1024
template <int NV, bool MAX, typename Epi>
__global__ __launch_bounds__(SULI_FINISH_THREADS) void k_finish(const double *__restrict__ partials, int nb,
                                                                double *__restrict__ result, Epi epi) {
  if (!epi.enabled()) {
    if (threadIdx.x == 0) epi_publish(epi);
    return;
  }
  __shared__ double lds[NV * SULI_LDS_RED];
  double acc[NV];
  const double init = MAX ? -1.0e300 : 0.0;
  const int bd = blockDim.x;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    const double *pq = partials + (size_t)q * nb;
    double a0 = init, a1 = init;
    for (int t = threadIdx.x; t < nb; t += 2 * bd) {
      const double x0 = pq[t];
      const double x1 = t + bd < nb ? pq[t + bd] : init;
      a0 = MAX ? (x0 > a0 ? x0 : a0) : a0 + x0;
      a1 = MAX ? (x1 > a1 ? x1 : a1) : a1 + x1;
    }
    acc[q] = MAX ? (a1 > a0 ? a1 : a0) : a0 + a1;
  }
  block_reduce<NV, MAX>(acc, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) result[q] = acc[q];
    epi(result);
    epi_publish(epi);
  }
}

template <int NV, bool MAX = false, typename Epi = EpiNone>
inline void launch_finish(hipStream_t s, dim3 producer_grid, RedSlot rs, Epi epi = Epi()) {
  const int nb = (int)(producer_grid.x * producer_grid.y * producer_grid.z);
  hipLaunchKernelGGL((k_finish<NV, MAX, Epi>), dim3(1), dim3(SULI_FINISH_THREADS), 0, s, rs.partials, nb, rs.result,
                     epi);
}

// ---------------------------------------------------------------------------
// WENO5-JS one-sided derivative from five consecutive divided differences
// (14_level_set.f90:580-608), evaluated in the reference's operation order.
template <typename T>
__device__ __forceinline__ T weno5(T v1, T v2, T v3, T v4, T v5) {
  const T aux1 = T(13.) / T(12.), aux2 = T(1.) / T(4.), aux3 = T(1.) / T(6.);
  const T eps = T(0.00000001);
  T a, b;
  a = v1 - 2 * v2 + v3;
  b = v1 - 4 * v2 + 3 * v3;
  const T is1 = aux1 * a * a + aux2 * b * b;
  a = v2 - 2 * v3 + v4;
  b = v2 - v4;
  const T is2 = aux1 * a * a + aux2 * b * b;
  a = v3 - 2 * v4 + v5;
  b = 3 * v3 - 4 * v4 + v5;
  const T is3 = aux1 * a * a + aux2 * b * b;
  const T al1 = T(0.1) / ((eps + is1) * (eps + is1));
  const T al2 = T(0.6) / ((eps + is2) * (eps + is2));
  const T al3 = T(0.3) / ((eps + is3) * (eps + is3));
  const T o1 = al1 / (al1 + al2 + al3);
  const T o2 = al2 / (al1 + al2 + al3);
  const T o3 = al3 / (al1 + al2 + al3);
  return aux3 * (o1 * (2 * v1 - 7 * v2 + 11 * v3) + o2 * (-v2 + 5 * v3 + 2 * v4) +
                 o3 * (2 * v3 + 5 * v4 - v5));
}

// The mixed-precision form: the same smoothness indicators and stencils,
// with the weights 0.1/d1, 0.6/d2, 0.3/d3 (d = (eps + IS)^2) normalised
// through the products 0.1 d2 d3, 0.6 d1 d3, 0.3 d1 d2 and one reciprocal
// instead of four (a reciprocal issues at a quarter of the VALU rate).  With
// eps = 1e-8, d >= 1e-16 and a product of two d stays above the FP32 normal
// range's floor; the divided differences of a level set are O(1), far from
// its ceiling.
template <>
__device__ __forceinline__ float weno5<float>(float v1, float v2, float v3, float v4, float v5) {
  const float aux1 = 13.f / 12.f, aux2 = 1.f / 4.f, aux3 = 1.f / 6.f;
  const float eps = 0.00000001f;
  float a, b;
  a = v1 - 2 * v2 + v3;
  b = v1 - 4 * v2 + 3 * v3;
  const float is1 = aux1 * a * a + aux2 * b * b;
  a = v2 - 2 * v3 + v4;
  b = v2 - v4;
  const float is2 = aux1 * a * a + aux2 * b * b;
  a = v3 - 2 * v4 + v5;
  b = 3 * v3 - 4 * v4 + v5;
  const float is3 = aux1 * a * a + aux2 * b * b;
  const float d1 = (eps + is1) * (eps + is1), d2 = (eps + is2) * (eps + is2), d3 = (eps + is3) * (eps + is3);
  const float w1 = 0.1f * (d2 * d3), w2 = 0.6f * (d1 * d3), w3 = 0.3f * (d1 * d2);
  const float r = 1.f / (w1 + w2 + w3);
  return aux3 * r *
         (w1 * (2 * v1 - 7 * v2 + 11 * v3) + w2 * (-v2 + 5 * v3 + 2 * v4) + w3 * (2 * v3 + 5 * v4 - v5));
}

// Seven values phi(pos-3..pos+3) of a pencil of n points with stride st,
// centred on the point at address c (1 <= pos <= n).  Points outside 1..n are
// the ghost values of 14_level_set.f90:532-570: ihs 0 periodic with period
// n-1, ihs 1 extrapolation (1/5 on the low side, 1/11 on the high side), ihs 2
// the 1/11 formula on both sides.  The ghosts depend only on phi(1..3) and
// phi(n-2..n), which always lie inside the thread's own window.
template <typename T>
__device__ __forceinline__ void window7(const T *__restrict__ f, fidx c, fidx st, int pos,
                                        int n, int ihs, T (&w)[7]) {
  const Fld<T> F{f};  // 32-bit byte offsets from the field's base
  if (ihs == 0) {
#pragma unroll
    for (int m = 0; m < 7; m++) {
      int q = pos - 3 + m;
      int qs = q < 1 ? q + n - 1 : (q > n ? q - n + 1 : q);
      w[m] = F[c + (fidx)(qs - pos) * st];
    }
    return;
  }
#pragma unroll
  for (int m = 0; m < 7; m++) {
    int q = pos - 3 + m;
    int qc = q < 1 ? 1 : (q > n ? n : q);
    w[m] = F[c + (fidx)(qc - pos) * st];
  }
  if (pos <= 3) {
    const T p1 = F[c + (fidx)(1 - pos) * st], p2 = F[c + (fidx)(2 - pos) * st],
            p3 = F[c + (fidx)(3 - pos) * st];
    T g0, gm1, gm2;
    if (ihs == 1) {
      g0 = T(1.) / T(5.) * (T(12.) * p1 - T(9.) * p2 + T(2.) * p3);
      gm1 = T(1.) / T(5.) * (T(12.) * g0 - T(9.) * p1 + T(2.) * p2);
      gm2 = T(1.) / T(5.) * (T(12.) * gm1 - T(9.) * g0 + T(2.) * p1);
    } else {
      g0 = T(1.) / T(11.) * (T(18.) * p1 - T(9.) * p2 + T(2.) * p3);
      gm1 = T(1.) / T(11.) * (T(18.) * g0 - T(9.) * p1 + T(2.) * p2);
      gm2 = T(1.) / T(11.) * (T(18.) * gm1 - T(9.) * g0 + T(2.) * p1);
    }
#pragma unroll
    for (int m = 0; m < 7; m++) {
      int q = pos - 3 + m;
      w[m] = q == 0 ? g0 : (q == -1 ? gm1 : (q == -2 ? gm2 : w[m]));
    }
  }
  if (pos >= n - 2) {
    const T pn = F[c + (fidx)(n - pos) * st], pn1 = F[c + (fidx)(n - 1 - pos) * st],
            pn2 = F[c + (fidx)(n - 2 - pos) * st];
    const T g1 = T(1.) / T(11.) * (T(18.) * pn - T(9.) * pn1 + T(2.) * pn2);
    const T g2 = T(1.) / T(11.) * (T(18.) * g1 - T(9.) * pn + T(2.) * pn1);
    const T g3 = T(1.) / T(11.) * (T(18.) * g2 - T(9.) * g1 + T(2.) * pn);
#pragma unroll
    for (int m = 0; m < 7; m++) {
      int q = pos - 3 + m;
      w[m] = q == n + 1 ? g1 : (q == n + 2 ? g2 : (q == n + 3 ? g3 : w[m]));
    }
  }
}

// Both one-sided WENO derivatives at the window centre:
// dp = phi_x^+ (dphidxp), dn = phi_x^- (dphidxn), 14_level_set.f90:572-608.
template <typename T>
__device__ __forceinline__ void weno_pm(const T (&w)[7], T h, T &dp, T &dn) {
  T d[6];
#pragma unroll
  for (int m = 0; m < 6; m++) d[m] = (w[m + 1] - w[m]) / h;
  dp = weno5<T>(d[5], d[4], d[3], d[2], d[1]);
  dn = weno5<T>(d[0], d[1], d[2], d[3], d[4]);
}

// der_weno at one cell: the six one-sided derivatives of f, with the ihs of
// 14_level_set.f90:353-377 (x, y: 0 if periodic else 2; z: 1) or, when
// all_ihs1 is set, ihs 1 in every direction (as mod_ls1 uses it).
template <typename T>
__device__ __forceinline__ void der_weno_cell(const T *__restrict__ f, const Grid &g, int i, int j,
                                              int k, fidx c, bool all_ihs1, T &ta, T &tb,
                                              T &tc, T &td, T &te, T &tf) {
  T w[7];
  const int ihx = all_ihs1 ? 1 : (g.ccx0 == 0 ? 0 : 2);
  const int ihy = all_ihs1 ? 1 : (g.ccy0 == 0 ? 0 : 2);
  window7<T>(f, c, 1, i, g.nx, ihx, w);
  weno_pm<T>(w, T(g.dx), ta, td);
  window7<T>(f, c, g.sx, j, g.ny, ihy, w);
  weno_pm<T>(w, T(g.dy), tb, te);
  window7<T>(f, c, g.sz, k, g.nz, 1, w);
  weno_pm<T>(w, T(g.dz), tc, tf);
}
