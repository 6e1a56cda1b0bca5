// Wave inflow at x = 0 (wave_t of 0_modules.f90:35): the coefficients of
// waves_coef (3_contorno.f90:1070-1205) on the host, and boundary_waves
// (:952-1068) on the device.
//
// waves_coef: for Stokes I and II (wave_t 1, 2) the wavelength of the linear
// dispersion relation L = g T^2/(2 pi) tanh(2 pi h0/L) by 1000 fixed-point
// passes; for Stokes V (wave_t 5, the fifth-order theory of Fenton 1985) the
// wavelength among the 10000 candidates L0/2 + L0 (ii/10000) 2.5 (L0 the
// deep-water wavelength) at which the two estimates of the steepness
// parameter lambda (10000 fixed-point passes of each, from the wave height
// and from the dispersion relation) agree best, then the fifth-order
// coefficients of the surface (aeta1..5) and of the velocity (avel1..5).
// The candidates are scanned by host threads; the selection keeps the first
// candidate of smallest disagreement, as the sequential scan does.
//
// boundary_waves, at the time t of the stage: the surface h_f and the
// velocities u, w of the wave at the first two x positions (i = 0, 1:
// x = (i - 1/2) dx for the surface and w, (i - 1) dx for u), zero above the
// surface; the planes bxx0, bxx1, bxz0 of the prescribed wall x = 0 take
// u(0), u(1), w(0) (bxy0 is zero: the reference copies there a v that it
// never sets), and the level set of the first cell column becomes
// h_f(1) - z.  The reference holds its level-set copy ls1 in real(4): the
// surface test uses the single-precision value, and the whole level set
// passes through single precision (a no-op for the mixed-precision fields).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "ctx.hpp"

#define NEEDCTX()                                              \
  do {                                                         \
    if (!g_ctx) {                                              \
      suli_set_error("library not initialised (suli_init)");   \
      return 2;                                                \
    }                                                          \
  } while (0)

namespace {

struct WavePar {
  int t;                 // 1 Stokes I, 2 Stokes II, 5 Stokes V
  double a, h0, n, f;    // amplitude, still depth, wave number, angular frequency
  double aeta[5], avel[5];
};

// fifth-order coefficients that depend on the depth ratio only, through
// c = cosh(2 pi h0/L), s = sinh(2 pi h0/L) (3_contorno.f90:1124-1131)
struct B5 {
  double b33, b35, b55, c1, c2;
};
B5 stokes5_b(double c, double s) {
  const double cc = c * c, c4i = cc * cc, c12i = c4i * (c4i * c4i);  // c**12 with an integer exponent
  B5 r;
  r.b33 = 3. * (8. * std::pow(c, 6.) + 1.) / (64. * std::pow(s, 6.));
  r.b35 = (88128. * std::pow(c, 14.) - 208224. * c12i + 70848. * std::pow(c, 10.) + 54000. * std::pow(c, 8.) -
           21816. * std::pow(c, 6.) + 6264. * std::pow(c, 4.) - 54. * c * c - 81.) /
          (12288. * std::pow(s, 12.) * (6. * c * c - 1.));
  r.b55 = (192000. * std::pow(c, 16.) - 262720. * std::pow(c, 14.) + 83680. * c12i + 20160. * std::pow(c, 10.) -
           7280. * std::pow(c, 8.) + 7160. * std::pow(c, 6.) - 1800. * std::pow(c, 4.) - 1050. * c * c + 225.) /
          (12288 * std::pow(s, 10.) * (6. * c * c - 1.) * (8. * std::pow(c, 4.) - 11. * c * c + 3.));
  r.c1 = (8. * std::pow(c, 4.) - 8. * c * c + 9.) / (8. * std::pow(s, 4.));
  r.c2 = (3840. * std::pow(c, 12.) - 4096. * std::pow(c, 10.) + 2592. * std::pow(c, 8.) - 1008. * std::pow(c, 6.) +
          5944. * std::pow(c, 4.) - 1830. * c * c + 147.) /
         (512. * std::pow(s, 10.) * (6. * c * c - 1.));
  return r;
}

// waves_coef: fills w (and the wavelength, celerity) from amplitude a_w,
// period p_w, still depth h0_f and gravity gz
void waves_coef(int wave_t, double a_w, double p_w, double h0_f, double gz, WavePar &w, double &l_w, double &c_w) {
  const double pi = M_PI;
  const double l0_w = gz * p_w * p_w / (2. * pi);
  double lamb = 0.;
  l_w = 0.;
  if (wave_t <= 2) {
    l_w = gz * p_w * p_w / (2. * pi);
    for (int i = 1; i <= 1000; i++) l_w = gz * p_w * p_w / (2. * pi) * std::tanh(2. * pi * h0_f / l_w);
  } else {
    const int nxii = 10000;
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    struct Best { double erro = 999., lamb = 0., l_w = 0.; int ii = 0; };
    std::vector<Best> best(nth);
    auto scan = [&](unsigned q) {
      const int lo = 1 + (int)((int64_t)nxii * q / nth), hi = (int)((int64_t)nxii * (q + 1) / nth);
      Best b;
      for (int ii = lo; ii <= hi; ii++) {
        double lamb1 = 0., lamb2 = 0.;
        const double l_w1 = l0_w * 0.5 + l0_w * ((double)ii / nxii) * 2.5;
        const double s = std::sinh(2. * pi * h0_f / l_w1), c = std::cosh(2. * pi * h0_f / l_w1);
        const B5 k = stokes5_b(c, s);
        for (int i = 1; i <= 10000; i++)
          lamb1 = pi * a_w * 2. / l_w1 - std::pow(lamb1, 3.) * k.b33 - std::pow(lamb1, 5.) * (k.b35 + k.b55);
        const double th = std::tanh(2. * pi * h0_f / l_w1);
        for (int i = 1; i <= 10000; i++)
          lamb2 = std::sqrt((l_w1 / (l0_w * th) - 1.) / (k.c1 + std::pow(lamb2, 2.) * k.c2));
        const double erro = std::fabs(lamb1 - lamb2);
        if (erro < b.erro) b = Best{erro, lamb1, l_w1, ii};
      }
      best[q] = b;
    };
    std::vector<std::thread> th;
    for (unsigned q = 1; q < nth; q++) th.emplace_back(scan, q);
    scan(0);
    for (auto &x : th) x.join();
    double erro0 = 999.;
    for (const Best &b : best)  // chunks in candidate order: the first strict minimum wins
      if (b.erro < erro0) {
        erro0 = b.erro;
        lamb = b.lamb;
        l_w = b.l_w;
      }
  }
  c_w = l_w / p_w;
  w.t = wave_t;
  w.a = a_w;
  w.h0 = h0_f;
  w.f = 2. * pi / p_w;
  w.n = 2. * pi / l_w;
  for (int q = 0; q < 5; q++) w.aeta[q] = w.avel[q] = 0.;
  if (wave_t == 5) {
    const double n_w = w.n, s = std::sinh(2. * pi * h0_f / l_w), c = std::cosh(2. * pi * h0_f / l_w);
    const B5 k = stokes5_b(c, s);
    const double a11 = 1. / s;
    const double a13 = -c * c * (5. * c * c + 1.) / (8. * std::pow(s, 5.));
    const double a15 = -(1184. * std::pow(c, 10.) - 1440. * std::pow(c, 8.) - 1992. * std::pow(c, 6.) +
                         2641. * std::pow(c, 4.) - 249. * std::pow(c, 2.) + 18.) / (1536. * std::pow(s, 11.));
    const double a22 = 3. / (8. * std::pow(s, 4.));
    const double a24 = (192. * std::pow(c, 8.) - 424. * std::pow(c, 6.) - 312. * std::pow(c, 4.) + 480. * c * c - 17.) /
                       (768. * std::pow(s, 10.));
    const double a33 = (13. - 4. * c * c) / (64. * std::pow(s, 7.));
    const double a35 = (512. * std::pow(c, 12.) + 4224. * std::pow(c, 10.) - 6800. * std::pow(c, 8.) -
                        12808. * std::pow(c, 6.) + 16704. * std::pow(c, 4.) - 3154. * c * c + 107.) /
                       (4096. * std::pow(s, 13.) * (6. * c * c - 1.));
    const double a44 = (80. * std::pow(c, 6.) - 816. * std::pow(c, 4.) + 1338. * c * c - 197.) /
                       (1536. * std::pow(s, 10.) * (6. * c * c - 1.));
    const double a55 = -(2880. * std::pow(c, 10.) - 72480. * std::pow(c, 8.) + 324000. * std::pow(c, 6.) -
                         432000. * std::pow(c, 4.) + 163470. * c * c - 16245.) /
                       (61440. * std::pow(s, 11.) * (6. * c * c - 1.) * (8. * std::pow(c, 4.) - 11. * c * c + 3.));
    const double b22 = c * (2. * c * c + 1.) / (4. * std::pow(s, 3.));
    const double b24 = c * (272. * std::pow(c, 8.) - 504. * std::pow(c, 6.) - 192. * std::pow(c, 4.) + 322 * c * c + 21.) /
                       (384. * std::pow(s, 9.));
    const double b44 = c * (768. * std::pow(c, 10.) - 448. * std::pow(c, 8.) - 48. * std::pow(c, 6.) +
                            48. * std::pow(c, 4.) + 106. * c * c - 21.) /
                       (384. * std::pow(s, 9.) * (6. * c * c - 1.));
    const double L = lamb;
    w.aeta[0] = L / n_w;
    w.aeta[1] = (L * L * b22 + std::pow(L, 4.) * b24) / n_w;
    w.aeta[2] = (std::pow(L, 3.) * k.b33 + std::pow(L, 5.) * k.b35) / n_w;
    w.aeta[3] = std::pow(L, 4.) * b44 / n_w;
    w.aeta[4] = std::pow(L, 5.) * k.b55 / n_w;
    const double v = 2. * pi / (p_w * n_w);
    w.avel[0] = 1. * v * (L * a11 + std::pow(L, 3.) * a13 + std::pow(L, 5.) * a15);
    w.avel[1] = 2. * v * (std::pow(L, 2.) * a22 + std::pow(L, 4.) * a24);
    w.avel[2] = 3. * v * (std::pow(L, 3.) * a33 + std::pow(L, 5.) * a35);
    w.avel[3] = 4. * v * (std::pow(L, 4.) * a44);
    w.avel[4] = 5. * v * (std::pow(L, 5.) * a55);
  }
}

// the wave at position index i (0, 1) and height kp = (k - 1/2) dz: surface
// level set (single precision, as the reference's ls1) and velocities u, w
__device__ inline void wave_point(const WavePar &w, double dx, double dz, double t, int i, double kp, float &ls1,
                                  double &u1, double &w1) {
  const double n = w.n, f = w.f, a = w.a, h0 = w.h0;
  const double aux2 = n * dx * (i - 0.5) - f * t, aux3 = n * dx * (i - 1.) - f * t;
  double h;
  if (w.t == 1) {
    h = h0 + a * cos(aux2);
  } else if (w.t == 2) {
    const double s = tanh(n * h0);
    h = h0 + a * cos(aux2) + n * a * a * (3. - s * s) / (4. * s * s * s) * cos(2. * aux2);
  } else {
    h = h0;
#pragma unroll
    for (int m = 1; m <= 5; m++) h += w.aeta[m - 1] * cos(m * aux2);
  }
  ls1 = (float)(h - kp);
  u1 = 0.;
  w1 = 0.;
  if (ls1 >= 0.f) {
    const double zw = kp - 0.5 * dz;
    if (w.t == 1) {
      u1 = a * f * cosh(n * kp) / sinh(n * h0) * cos(aux3);
      w1 = a * f * sinh(n * zw) / sinh(n * h0) * sin(aux2);
    } else if (w.t == 2) {
      const double s4 = pow(sinh(n * h0), 4.);
      u1 = a * f * cosh(n * kp) / sinh(n * h0) * cos(aux3) + 3. / 4. * a * a * f * n * cosh(2. * n * kp) / s4 * cos(2. * aux3);
      // the second-order term of w takes the phase of u (3_contorno.f90:1013)
      w1 = a * f * sinh(n * zw) / sinh(n * h0) * sin(aux2) + 3. / 4. * a * a * f * n * sinh(2. * n * zw) / s4 * sin(2. * aux3);
    } else {
#pragma unroll
      for (int m = 1; m <= 5; m++) {
        u1 += w.avel[m - 1] * cosh(m * (n * kp)) * cos(m * aux3);
        w1 += w.avel[m - 1] * sinh(m * (n * zw)) * sin(m * aux2);
      }
    }
  }
}

// planes and the first level-set column; one thread per (j, k) of the x = 0
// wall, j in [0, ny+2], k in [0, nz+2]
template <typename T>
__global__ void k_waves(Grid g, WavePar w, double t, T *bp, BPlanes L, T *ls) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y * blockDim.y + threadIdx.y;
  const int ny = g.ny, nz = g.nz;
  if (j > ny + 2 || k > nz + 2) return;
  if (k <= nz + 1) bp[L.at(BXY0, j, k)] = T(0);
  if (j > ny + 1) return;
  if (k == nz + 2) {
    bp[L.at(BXZ0, j, k)] = T(0);
    return;
  }
  const double kp = (k - 0.5) * g.dz;
  float l0, l1;
  double u0, w0, u1, w1;
  wave_point(w, g.dx, g.dz, t, 0, kp, l0, u0, w0);
  wave_point(w, g.dx, g.dz, t, 1, kp, l1, u1, w1);
  bp[L.at(BXX0, j, k)] = (T)u0;
  bp[L.at(BXZ0, j, k)] = (T)w0;
  bp[L.at(BXX1, j, k)] = (T)u1;
  if (j >= 1 && j <= ny && k >= 1 && k <= nz) ls[g.idx(1, j, k)] = (T)(double)l1;
}

// the rest of the level set through single precision (FP64 fields)
__global__ void k_ls_single(Grid g, double *ls) {
  const int i = 2 + blockIdx.x * blockDim.x + threadIdx.x, j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.nx || j > g.ny) return;
  const fidx c = g.idx(i, j, k);
  ls[c] = (double)(float)ls[c];
}

}  // namespace

// waves_coef (3_contorno.f90:1070-1205; 2_iniciais.f90:52-54): wave_t 0
// (off), 1 Stokes I, 2 Stokes II or 5 Stokes V; the reference's waves_coef
// fixes a_w = 0.03 m, p_w = 2.5 s, h0_f = 1 m.  coef (may be null)
// receives l_w, n_w, f_w, c_w, aeta1..5, avel1..5.
extern "C" int suli_waves_coef(int wave_t, double a_w, double p_w, double h0_f, double *coef) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  if (wave_t != 0 && wave_t != 1 && wave_t != 2 && wave_t != 5) {
    suli_set_error("suli_waves_coef: wave_t must be 0 (none), 1 (Stokes I), 2 (Stokes II) or 5 (Stokes V)");
    return 1;
  }
  if (wave_t != 0 && !(a_w > 0 && p_w > 0 && h0_f > 0 && c.ph.gz > 0)) {
    suli_set_error("suli_waves_coef: amplitude, period, depth and gravity gz must be positive");
    return 1;
  }
  if (wave_t != 0 && c.g.ccx0 != 3) {
    suli_set_error("suli_waves_coef: the wave enters through the prescribed wall x = 0 (ccx0 = 3)");
    return 1;
  }
  c.wave.t = wave_t;
  if (wave_t == 0) return 0;
  WavePar w{};
  double l_w = 0, c_w = 0;
  waves_coef(wave_t, a_w, p_w, h0_f, c.ph.gz, w, l_w, c_w);
  c.wave.a = a_w;
  c.wave.p = p_w;
  c.wave.h0 = h0_f;
  c.wave.wc[0] = l_w; c.wave.wc[1] = w.n; c.wave.wc[2] = w.f; c.wave.wc[3] = c_w;
  for (int q = 0; q < 5; q++) {
    c.wave.wc[4 + q] = w.aeta[q];
    c.wave.wc[9 + q] = w.avel[q];
  }
  if (coef)
    for (int q = 0; q < 14; q++) coef[q] = c.wave.wc[q];
  return 0;
}

int waves_now(Ctx &c) {
  if (c.wave.t == 0) return 0;
  WavePar w{};
  w.t = c.wave.t;
  w.a = c.wave.a;
  w.h0 = c.wave.h0;
  w.n = c.wave.wc[1];
  w.f = c.wave.wc[2];
  for (int q = 0; q < 5; q++) {
    w.aeta[q] = c.wave.wc[4 + q];
    w.avel[q] = c.wave.wc[9 + q];
  }
  const Grid &g = c.g;
  const BPlanes L = bplanes_layout(g.nx, g.ny, g.nz);
  if (c.prec == SULI_PREC_FP64)
    hipLaunchKernelGGL(k_ls_single, dim3((g.nx - 1 + 63) / 64, (g.ny + 3) / 4, g.nz), dim3(64, 4, 1), 0, c.st, g,
                       (double *)c.ls);
  DISPATCH(c, hipLaunchKernelGGL(k_waves<TT>, dim3((g.ny + 3 + 63) / 64, (g.nz + 3 + 3) / 4, 1), dim3(64, 4, 1), 0,
                                 c.st, g, w, c.t_now, (TT *)c.bp, L, (TT *)c.ls));
  HIPCHK(hipGetLastError());
  return 0;
}

// boundary_waves (3_contorno.f90:952-1068) at the library's time
extern "C" int suli_boundary_waves(void) {
  NEEDCTX();
  if (g_ctx->wave.t == 0) {
    suli_set_error("suli_boundary_waves: no wave (suli_waves_coef)");
    return 1;
  }
  return waves_now(*g_ctx);
}
