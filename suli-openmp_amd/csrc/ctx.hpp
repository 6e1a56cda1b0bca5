// Host-side state of the library: one GPU, one stream, every field resident.
#pragma once
#include <string>
#include <vector>

#include "../../include/suli.h"
#include "launch.hpp"
#include "poisson_ref.hpp"

struct PoissonFCG;  // poisson_fcg.hip

// reduction result slots (indices into Ctx::res)
enum { RS_VOL = 0, RS_BAND = 2, RS_DIV = 4, RS_POIS = 8, RS_NSLOT = 64 };

struct Ctx {
  suli_params P{};
  Grid g{};
  Phys ph{};
  int prec = SULI_PREC_FP64;
  size_t esz = 8;
  hipStream_t st = nullptr;
  // fields of type T (double or float by precision)
  void *u = nullptr, *v = nullptr, *w = nullptr, *un = nullptr, *vn = nullptr, *wn = nullptr;
  void *u0 = nullptr, *v0 = nullptr, *w0 = nullptr;
  void *ls = nullptr, *lsA = nullptr, *lsB = nullptr, *S = nullptr;
  void *hs = nullptr, *hsx = nullptr, *hsy = nullptr, *hsz = nullptr, *rho = nullptr, *lsnu = nullptr;
  void *modls = nullptr, *kurv = nullptr, *nut = nullptr, *nrx = nullptr, *nry = nullptr, *nrz = nullptr;
  void *ddx = nullptr, *ddy = nullptr, *ddz = nullptr, *Fu = nullptr, *Fv = nullptr, *Fw = nullptr;
  void *tfu = nullptr, *tfv = nullptr, *tfw = nullptr;  // optional momentum sources
  double *prd1 = nullptr;
  // projection (FP64)
  double *D = nullptr, *sD = nullptr, *q = nullptr, *ax = nullptr, *ay = nullptr, *az = nullptr;
  double *e = nullptr, *r = nullptr, *pp = nullptr, *mp = nullptr;
  CGCtl *ctl = nullptr, *ctl_h = nullptr;
  // reductions
  double *partials = nullptr;
  unsigned int *counters = nullptr;
  double *res = nullptr, *res_h = nullptr;
  bool convdiff_pending = false;
  double vol_ini = 0, vol_ins = 0;
  PoissonFCG *fcg = nullptr;
  std::vector<void *> allocs;

  RedSlot slot(int off, int ctr) const { return RedSlot{partials, counters + ctr, res + off}; }
};

extern Ctx *g_ctx;
void suli_set_error(const std::string &m);

#define HIPCHK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      suli_set_error(std::string(#x) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + \
                     ":" + std::to_string(__LINE__) + ")");                          \
      return 1;                                                                      \
    }                                                                                \
  } while (0)

// dispatch on the field precision
#define DISPATCH(ctx, CALL)              \
  do {                                   \
    if ((ctx).prec == SULI_PREC_FP64) {  \
      using TT = double;                 \
      CALL;                              \
    } else {                             \
      using TT = float;                  \
      CALL;                              \
    }                                    \
  } while (0)

// Poisson solvers (poisson_fcg.hip)
int fcg_create(Ctx &c);
void fcg_destroy(Ctx &c);
int fcg_solve(Ctx &c, double dt, suli_poisson_stats *st);
