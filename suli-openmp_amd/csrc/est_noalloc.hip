// max |div u| of est (4_plot.f90:500-525) without any allocation, for callers
// that cannot allocate device memory: code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc.  It needs no library context: the
// velocities are the caller's device arrays in the reference extents, the
// result goes to the caller's device memory, and the block partials of the
// reduction live in a fixed file-scope device buffer (`scratch`), one copy
// per GPU, held until the process ends.
//
// Every call on a GPU shares that GPU's buffer: a call writes the partials
// of its blocks and reads back exactly those, so nothing needs clearing, and
// two calls that overlap on one GPU (different streams, no ordering) race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ctx.hpp"

// partials of the reduction: one per work-group, the grid never has more
// (16 KiB per GPU); grids: ny*nz rows of at most 2^30 (row indices stay
// int), velocity arrays of at most 2^36 elements
#define SULI_NOALLOC_SLOTS 2048
#define SULI_NOALLOC_MAX_ROWS ((int64_t)1 << 30)
#define SULI_NOALLOC_MAX_ELEMS ((int64_t)1 << 36)
static __device__ __align__(256) double scratch[SULI_NOALLOC_SLOTS];

// address of each GPU's copy of `scratch`, looked up once per device with
// that device current (the host-side name is only a stand-in for it)
#define SULI_NOALLOC_MAXDEV 64
static double *g_scratch_dev[SULI_NOALLOC_MAXDEV];
static std::mutex g_scratch_mu;

// Looks up the current device's scratch address (once per device); called
// by suli_init for its device and by the entry point on its first call on a
// device.  Returns 0 and the address, or nonzero with the error set.
int suli_noalloc_setup(double **out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= SULI_NOALLOC_MAXDEV) {
    suli_set_error("suli_est_div_noalloc: device ordinal " + std::to_string(dev) + " beyond the " +
                   std::to_string(SULI_NOALLOC_MAXDEV) + " this library tracks");
    return 1;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = (double *)ptr;
  }
  if (out) *out = g_scratch_dev[dev];
  return 0;
}

// One work-group per stretch of rows (j, k): 64 lanes along i, 4 rows per
// pass, grid-strided over the ny*nz rows; each group writes its maximum to
// partials[blockIdx.x].  u(0:nx+2,0:ny+1,0:nz+1), v(0:nx+1,0:ny+2,0:nz+1),
// w(0:nx+1,0:ny+1,0:nz+2), Fortran order.  The divergence is est's
// expression in its order, (du/dx + dv/dy) + dw/dz, in FP64.
__global__ __launch_bounds__(256) void k_div_rows(const double *__restrict__ u, const double *__restrict__ v,
                                                  const double *__restrict__ w, int nx, int ny, int nz,
                                                  double dx, double dy, double dz, double *__restrict__ partials) {
  const int64_t ux = nx + 3, uy = ny + 2;  // u extents
  const int64_t vx = nx + 2, vy = ny + 3;  // v
  const int64_t wx = nx + 2, wy = ny + 2;  // w
  const int64_t rows = (int64_t)ny * nz;
  double m = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * 4 + threadIdx.y; r < rows; r += (int64_t)gridDim.x * 4) {
    const int j = 1 + (int)(r % ny), k = 1 + (int)(r / ny);
    const double *ur = u + ux * (j + uy * k);
    const double *vr0 = v + vx * (j + vy * k), *vr1 = vr0 + vx;
    const double *wr0 = w + wx * (j + wy * k), *wr1 = wr0 + wx * wy;
    for (int i = 1 + threadIdx.x; i <= nx; i += 64) {
      const double d = (ur[i + 1] - ur[i]) / dx + (vr1[i] - vr0[i]) / dy + (wr1[i] - wr0[i]) / dz;
      const double a = d < 0 ? -d : d;
      m = a > m ? a : m;
    }
  }
  double acc[1] = {m};
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1, true>(acc, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) partials[blockIdx.x] = acc[0];
}

// One work-group: the largest of the nb partials to *out.
__global__ __launch_bounds__(256) void k_div_finish(const double *__restrict__ partials, int nb,
                                                    double *__restrict__ out) {
  double m = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) m = partials[b] > m ? partials[b] : m;
  double acc[1] = {m};
  __shared__ double lds[SULI_LDS_RED];
  block_reduce<1, true>(acc, lds);
  if (threadIdx.x == 0) *out = acc[0];
}

extern "C" int suli_est_div_noalloc(const double *u, const double *v, const double *w, int nx, int ny, int nz,
                                    double dx, double dy, double dz, double *div_max, void *stream) {
  if (!u || !v || !w || !div_max) {
    suli_set_error("suli_est_div_noalloc: null array");
    return 1;
  }
  if (nx < 1 || ny < 1 || nz < 1 || (int64_t)ny * nz > SULI_NOALLOC_MAX_ROWS ||
      (int64_t)(nx + 3) * (ny + 3) * (int64_t)(nz + 3) > SULI_NOALLOC_MAX_ELEMS) {
    suli_set_error("suli_est_div_noalloc: grid " + std::to_string(nx) + "x" + std::to_string(ny) + "x" +
                   std::to_string(nz) + " outside the supported range (each extent >= 1, ny*nz <= " +
                   std::to_string((long long)SULI_NOALLOC_MAX_ROWS) + ", velocity arrays of at most " +
                   std::to_string((long long)SULI_NOALLOC_MAX_ELEMS) + " elements)");
    return 1;
  }
  if (!(dx > 0) || !(dy > 0) || !(dz > 0)) {
    suli_set_error("suli_est_div_noalloc: grid spacings must be positive");
    return 1;
  }
  double *part = nullptr;
  if (suli_noalloc_setup(&part)) return 1;
  const int64_t rows = (int64_t)ny * nz;
  const int nb = (int)(((rows + 3) / 4) < SULI_NOALLOC_SLOTS ? (rows + 3) / 4 : SULI_NOALLOC_SLOTS);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_div_rows, dim3(nb), dim3(64, 4), 0, s, u, v, w, nx, ny, nz, dx, dy, dz, part);
  hipLaunchKernelGGL(k_div_finish, dim3(1), dim3(256), 0, s, part, nb, div_max);
  HIPCHK(hipGetLastError());
  return 0;
}
