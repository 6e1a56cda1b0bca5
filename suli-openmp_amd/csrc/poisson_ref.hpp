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
__device__ __forceinline__ double face_coef(const T *__restrict__ rho, fidx cf, fidx st, int f,
                                            int n, double dth2) {
  const int fc = f < 2 ? 2 : (f > n ? n : f);
  const fidx cc = cf + (fidx)(fc - f) * st;
  const double rf = ((double)rho[cc] + (double)rho[cc - st]) * 0.5;
  return dth2 / rf;
}

// q = -div u* at a cell (:80-87)
template <typename T>
__device__ __forceinline__ double div_q(const Grid &g, const T *__restrict__ u, const T *__restrict__ v,
                                        const T *__restrict__ w, fidx c) {
  return ((double)u[c] - (double)u[c + 1]) / g.dx + ((double)v[c] - (double)v[c + g.sx]) / g.dy +
         ((double)w[c] - (double)w[c + g.sz]) / g.dz;
}

// D and q = -div u* (:80-87), sqrt(D); q only when q is set.
template <typename T>
__global__ __launch_bounds__(256) void k_pD(Grid g, double dt, const T *__restrict__ u,
                                            const T *__restrict__ v, const T *__restrict__ w,
                                            const T *__restrict__ rho, double *__restrict__ D,
                                            double *__restrict__ sD, double *__restrict__ q) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const fidx c = g.idx(i, j, k);
  const fidx X = 1, Y = g.sx, Z = g.sz;
  const double mx = dt / (g.dx * g.dx), my = dt / (g.dy * g.dy), mz = dt / (g.dz * g.dz);
  const double d = face_coef<T>(rho, c + X, X, i + 1, g.nx, mx) + face_coef<T>(rho, c, X, i, g.nx, mx) +
                   face_coef<T>(rho, c + Y, Y, j + 1, g.ny, my) + face_coef<T>(rho, c, Y, j, g.ny, my) +
                   face_coef<T>(rho, c + Z, Z, k + 1, g.nz, mz) + face_coef<T>(rho, c, Z, k, g.nz, mz);
  D[c] = d;
  sD[c] = sqrt(d);
  if (q) q[c] = div_q<T>(g, u, v, w, c);
}

// Ghost cells of a P-extent array, (0:nx+1,0:ny+1,0:nz+1), filled with the
// value at the mirrored interior cell (composition of the x, y, z copies).
template <int UNUSED = 0>
__global__ void k_mirror_ghosts(Grid g, double *__restrict__ a, double *__restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y,
            k = blockIdx.z;
  if (i > g.nx + 1 || j > g.ny + 1 || k > g.nz + 1) return;
  if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz) return;
  const int mi = mirror_idx(i, g.nx, g.ccx0 == 0), mj = mirror_idx(j, g.ny, g.ccy0 == 0),
            mk = mirror_idx(k, g.nz, false);
  const fidx c = g.idx(i, j, k), s = g.idx(mi, mj, mk);
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
  const fidx c = g.idx(i, j, k);
  const fidx X = 1, Y = g.sx, Z = g.sz;
  if (j <= ny && k <= nz)
    ax[c] = face_coef<T>(rho, c, X, i, nx, dt / (g.dx * g.dx)) / sqrt(D[c] * D[c - X]);
  if (i <= nx && k <= nz)
    ay[c] = face_coef<T>(rho, c, Y, j, ny, dt / (g.dy * g.dy)) / sqrt(D[c] * D[c - Y]);
  if (i <= nx && j <= ny)
    az[c] = face_coef<T>(rho, c, Z, k, nz, dt / (g.dz * g.dz)) / sqrt(D[c] * D[c - Z]);
}

// ---- faces of the mixed-precision FCG set-up, in one pass ---------------------
// Faces: the scaled coefficients a_f of
// k_pcoef, their FP32 copies, and the finest multigrid level's conductances
// c_f rounded to BF16 (zero on walls); the diagonal of that level follows
// from the rounded faces (k_mg_diag_bf16).
template <typename T>
__global__ __launch_bounds__(256) void k_fcg_faces(Grid g, double dt, const T *__restrict__ rho,
                                                   const double *__restrict__ D, FcgFaceOut o) {
  const int i = 1 + blockIdx.x * SULI_BX + threadIdx.x, j = 1 + blockIdx.y * SULI_BY + threadIdx.y,
            k = 1 + blockIdx.z;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx + 1 || j > ny + 1 || k > nz + 1) return;
  const fidx c = g.idx(i, j, k);
  const fidx X = 1, Y = g.sx, Z = g.sz;
  // D on both sides of the face; a side outside the domain reads the mirrored
  // interior cell (the ghost value k_mirror_ghosts would have stored)
  const auto face = [&](fidx st, int f, int n, bool per, double h, double *a, float *a32, uint16_t *hf) {
    const double fc = face_coef<T>(rho, c, st, f, n, dt / (h * h));
    const double dlo = f == 1 ? D[c + (fidx)(mirror_idx(0, n, per) - 1) * st] : D[c - st];
    const double dhi = f == n + 1 ? D[c + (fidx)(mirror_idx(n + 1, n, per) - (n + 1)) * st] : D[c];
    const double av = fc / sqrt(dhi * dlo);
    a[c] = av;
    a32[c] = (float)av;
    hf[c] = to_bf16((f == 1 || f == n + 1) ? 0.f : (float)fc);
  };
  if (j <= ny && k <= nz) face(X, i, nx, g.ccx0 == 0, g.dx, o.ax, o.ax32, o.hx);
  if (i <= nx && k <= nz) face(Y, j, ny, g.ccy0 == 0, g.dy, o.ay, o.ay32, o.hy);
  if (i <= nx && j <= ny) face(Z, k, nz, false, g.dz, o.az, o.az32, o.hz);
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
