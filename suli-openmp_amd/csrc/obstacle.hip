// Bottom-obstacle geometry of obstaculo (3_contorno.f90:494-760), on the
// host: the obstacle height auxx on the half-step grid (-1:2(nx+1)+2,
// -1:2(ny+1)+2) for the reference's obst_t cases, then the columns ku, kv, kw
// the contorno kernels use.  No device work; needs no context.
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/suli.h"

void suli_set_error(const std::string &msg);

namespace {
// Fortran NINT: nearest integer, halves away from zero
int fnint(double x) { return (int)(x >= 0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5)); }
}  // namespace

extern "C" int suli_obstacle_columns(int obst_t, int nx, int ny, int nz, double dx, double dy, double dz,
                                     int32_t *ku, int32_t *kv, int32_t *kw) {
  if (!(obst_t >= 1 && obst_t <= 7) && obst_t != 10) {
    // 8 and 9 leave part of auxx unset in the reference (:685-712)
    suli_set_error("suli_obstacle_columns: obst_t must be 1..7 or 10");
    return 1;
  }
  if (nx < 1 || ny < 1 || nz < 1 || !(dx > 0) || !(dy > 0) || !(dz > 0) || !ku || !kv || !kw) {
    suli_set_error("suli_obstacle_columns: bad grid or null output");
    return 1;
  }
  const double pi = std::acos(-1.);
  const int nxx = (nx + 1) * 2, nyy = (ny + 1) * 2;
  const int sx = nxx + 4;  // auxx(-1:nxx+2, -1:nyy+2)
  std::vector<double> auxx((size_t)sx * (nyy + 4), 0.0);
  auto A = [&](int i, int j) -> double & { return auxx[(size_t)(i + 1) + (size_t)sx * (j + 1)]; };
  for (int j = -1; j <= nyy + 2; j++) {
    double aux1 = 0, aux2 = 0;  // carried along x by obst_t = 2 (:525-557)
    for (int i = -1; i <= nxx + 2; i++) {
      double &a = A(i, j);
      if (obst_t == 1) {  // sinusoidal dunes, :518-523 (elev, amp, comp, fase of 0_modules.f90:80-82)
        const double x = (i - 1.) * (dx * 0.5);
        const int elev = (int)(50. * dz);
        const double amp = 0.50, comp = .5, fase = -pi / 2.;
        a = elev + amp * (1. + std::sin(fase + 2. * pi * x / comp));
      } else if (obst_t == 2) {  // dune of Yue et al. (2003), :525-565
        const double x = (i - 1.) * (dx * 0.5);
        if (x <= 0.01) a = 0.02;
        else if (x <= 0.05) a = -std::tan(26. * pi / 180.) * x + 0.02;
        else if (x <= 0.075) a = 0.0005;
        else if (x <= 0.1375) { a = std::tan(1.8 * pi / 180.) * x + 0.0005; aux1 = a; }
        else if (x <= 0.325) { a = std::tan(5. * pi / 180.) * x + aux1; aux2 = a; }
        else if (x <= 0.39) a = std::tan(1.8 * pi / 180.) * x + aux2;
        else a = 0.02;
      } else if (obst_t == 3) {  // Gaussian bump, :567-585
        const double d = 0.2, x0 = 3., y0 = ny / 2. * dy, ah = 2.0, sigx = 0.5, sigy = 0.5;
        const double y = (j - 1.) * (dy * 0.5), x = (i - 1.) * (dx * 0.5);
        const double b1 = (y - y0) * (y - y0) / (2. * sigy * sigy);
        double b2 = (x - x0) * (x - x0) / (2. * sigx * sigx);
        b2 = -b2 - b1;
        a = std::exp(b2) * ah + d;
      } else if (obst_t == 4) {  // Beji and Battjes (1994), :587-605
        const double x = (i - 1) * (dx * 0.5);
        if (x < 6.) a = 0.;
        else if (x < 12.) a = std::tan(1. / 20.) * (x - 6.);
        else if (x < 14.) a = 0.30;
        else if (x < 17.) a = -std::tan(1. / 10.) * (x - 14.) + 0.30;
        else a = 0.;
      } else if (obst_t == 5) {  // Delft 1980 channel 0, :607-617
        const double x = (i - 2) * (dx * 0.5);
        a = x <= 1.0 ? 0.2 : 0.;
      } else if (obst_t == 6) {  // Delft 1980 channel 1/2, :619-637
        const double x = (i - 1.) * dx * 0.5;
        if (x < 1.) a = 0.2;
        else if (x < 1.4) a = -1. / 2. * (x - 1.) + 0.2;
        else if (x < 2.4) a = 0.;
        else if (x < 2.8) a = 1. / 2. * (x - 2.4);
        else a = 0.2;
      } else if (obst_t == 7) {  // holes and humps, :639-652
        const double raio = 0.12, x = (i - 1.) * dx * 0.5;
        if (x < 1. - raio) a = 0.;
        else if (x < 1.) a = std::sqrt(-(x - 1.) * (x - 1.) + raio * raio);
        else a = raio;
      } else {  // 10: Koshizuka (1995), Kleefsman (2005) box, :714-737
        const double x = (i - 1.) * dx * 0.5, y = (j - 1.) * dy * 0.5;
        a = -dz;
        if (x >= 0.75 - 0.16 * 0.5 && x <= 0.75 + 0.16 * 0.5) a = 0.16;
        if (y > 0.5 - 0.4 * 0.5 && y < 0.5 + 0.4 * 0.5) a = -dz;
      }
    }
  }
  // columns, :741-760
  for (int j = 0; j <= ny + 1; j++)
    for (int i = 0; i <= nx + 2; i++) {
      int k = fnint(A(i * 2 - 1, j * 2) / dz + 0.5);
      ku[i + (int64_t)(nx + 3) * j] = k > nz ? nz : k;
    }
  for (int j = 0; j <= ny + 2; j++)
    for (int i = 0; i <= nx + 1; i++) {
      int k = fnint(A(i * 2, j * 2 - 1) / dz + 0.5);
      kv[i + (int64_t)(nx + 2) * j] = k > nz ? nz : k;
    }
  for (int j = 0; j <= ny + 1; j++)
    for (int i = 0; i <= nx + 1; i++) {
      int k = (int)std::floor(A(i * 2, j * 2) / dz + 1.);
      kw[i + (int64_t)(nx + 2) * j] = k > nz + 1 ? nz + 1 : k;
    }
  return 0;
}
