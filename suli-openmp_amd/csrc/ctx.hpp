// Host-side state of the library: one GPU, one stream, every field resident.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/suli.h"
#include "launch.hpp"
#include "poisson_ref.hpp"

struct PoissonFCG;  // poisson_fcg.hip

// Per-kernel GPU timing with HIP events recorded on the library's stream
// around each launch (enabled by suli_prof_enable).  Event pairs are
// recycled: when many are pending, the stream is synchronised and their
// elapsed times are folded into per-kernel totals.
struct Prof {
  bool on = false;
  std::string only;  // when set, only the scope of this name is timed
  int every = 1;     // and only every every-th instance of it
  long long seen = 0;
  bool sample() { return every <= 1 || (seen++ % every) == 0; }
  std::vector<std::string> names;
  std::vector<double> total_ms;
  std::vector<long long> launches;
  struct Pending { int id; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
};

// reduction result slots (indices into Ctx::res)
enum { RS_VOL = 0, RS_BAND = 2, RS_DIV = 4, RS_POIS = 8, RS_NSLOT = 64 };

struct Ctx {
  suli_params P{};
  Grid g{};
  Phys ph{};
  // time-dependent horizontal forcing gx(t) = gx + amp*sin(2 pi t/period + phase)
  double force_amp = 0.0, force_period = 1.0, force_phase = 0.0, t_now = 0.0;
  Phys phys_now() const {
    Phys q = ph;
    if (force_amp != 0.0) q.gx = ph.gx + force_amp * std::sin(2.0 * M_PI * t_now / force_period + force_phase);
    return q;
  }
  // multigrid levels below the finest with at most this many cells smooth in
  // one fused launch per side of the V-cycle (0: every half sweep a launch)
  int64_t mg_fuse_max = 65536;
  int mg_coarse_sweeps = 24;  // red-black sweeps on the coarsest level
  // input/output scaling, prolongations and z.q carried by the neighbouring half
  // sweeps (1), and the finest level's colour-split iterate and right-hand side
  // in BF16 (2); 0: separate launches
  int mg_fuse_fine = 2;
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
  void *tfu = nullptr, *tfv = nullptr, *tfw = nullptr;  // momentum sources tf_u, tf_v, tf_w
  // kept forces of the RK3 and AB2 schemes (fu0, fu1 of 11_diversos.f90)
  void *fu0 = nullptr, *fv0 = nullptr, *fw0 = nullptr, *fu1 = nullptr, *fv1 = nullptr, *fw1 = nullptr;
  int t_tempo = SULI_TIME_RK2;  // time scheme
  int adv = 11;                 // advection: 10*adv_type + der (6_convdiff.f90:70-82)
  // bottom obstacle: columns ku, kv, kw on the device (plane index i + sx*j)
  // and as given (reference extents, for the restart file)
  bool obst = false;
  int *ku_d = nullptr, *kv_d = nullptr, *kw_d = nullptr;
  std::vector<int32_t> ku_h, kv_h, kw_h;
  // face viscosities xnut, ynut, znut of the directional model (m_turb = 2)
  void *xnf = nullptr, *ynf = nullptr, *znf = nullptr;
  // boundary planes of the prescribed-velocity walls (wall code 3; layout
  // bplanes_layout in common.hpp), field precision, zero until set
  void *bp = nullptr;
  // manufactured solution of 13_mms.f90 (t = mms_t: 0 off, 1 steady, 2 unsteady)
  struct {
    int t = 0;
    double a = 0.0, h0 = 2.0, coef = 0.0;
    bool src_done = false;     // the steady sources are computed once (it = 1)
    double *erro_p = nullptr;  // squared pressure errors (nx*ny*nz), kept between calls
    double *part = nullptr;    // block partials of the error sums
  } mms;
  // wave inflow at x = 0 (waves.hip): t = wave_t (0 none, 1, 2, 5), amplitude,
  // period, still depth and the coefficients l_w, n_w, f_w, c_w, aeta1..5, avel1..5
  struct {
    int t = 0;
    double a = 0.03, p = 2.5, h0 = 1.0;
    double wc[14] = {};
  } wave;
  bool ab2_started = false;     // fu0 holds the forces of the last AB2 step
  bool tf_on = false;                                    // set by the first upload of a source
  double *prd1 = nullptr;
  // projection (FP64)
  double *D = nullptr, *sD = nullptr, *q = nullptr, *ax = nullptr, *ay = nullptr, *az = nullptr;
  double *e = nullptr, *r = nullptr, *pp = nullptr, *mp = nullptr;
  CGCtl *ctl = nullptr, *ctl_h = nullptr;
  // reductions
  double *partials = nullptr;
  double *res = nullptr, *res_h = nullptr;
  bool convdiff_pending = false;
  double vol_ini = 0, vol_ins = 0;
  PoissonFCG *fcg = nullptr;
  // bumped whenever the density (or anything else the projection's operator
  // is built from) may have changed: heaviside, uploads, restarts, solver
  // settings.  The FCG projection rebuilds its operator only when it moved.
  uint64_t rho_epoch = 1;
  // initial guess of the FCG projection: the kind of stage set by the last
  // tempo call (1: first stage of a multi-stage step, step dt/2; 2: last
  // stage, step dt; 0: other).  Per kind k of the first and last stages, the
  // offsets of its last solutions from the pressure each solve started from,
  // newest first, in FP32: off[0] first stage, g(n) - h(n-1); off[1] last
  // stage, h(n) - g(n) (g, h the first- and last-stage pressures of step n),
  // offn[k] of them valid, all of the base step hdt.  last: the kind (1, 2)
  // of the solve whose pressure prd1 holds, 0 after any other change of it;
  // an offset joins its history only when the solve started from the other
  // kind's pressure.  (guess = 1: extrapolate the offsets.)
  int guess = 1, stage = 0, last = 0;
  int offn[2] = {0, 0};
  double hdt = 0.0;
  float *off[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  // field output: the record on the device, its pinned host copy, coordinate tables
  float *out_dev = nullptr, *out_host = nullptr, *out_tab = nullptr;
  int64_t out_cap = 0, out_host_cap = 0;
  std::vector<void *> allocs;
  Prof prof;
  // GPU time of a projection solve: events recorded on the stream at the
  // solve's start and after its last kernel (suli_poisson_stats.ms)
  hipEvent_t solve_ev[2] = {nullptr, nullptr};

  RedSlot slot(int off, int ctr) const { (void)ctr; return RedSlot{partials, res + off}; }
};

extern Ctx *g_ctx;
void suli_set_error(const std::string &m);
// address of the current device's scratch buffer of suli_est_div_noalloc
// (est_noalloc.hip), looked up once per device
int suli_noalloc_setup(double **out);
// manufactured-solution mode (mms.hip): the boundary planes and the momentum
// sources at the context's time t_now (no-ops when the mode is off; the
// steady planes only at t = 0)
int mms_planes_now(Ctx &c);
int mms_sources_now(Ctx &c);
extern "C" int suli_mms_set(int mms_t, double a, double h0, double coef);
// boundary_waves at t_now when a wave is set (waves.hip)
int waves_now(Ctx &c);

#define HIPCHK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      suli_set_error(std::string(#x) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + \
                     ":" + std::to_string(__LINE__) + ")");                          \
      return 1;                                                                      \
    }                                                                                \
  } while (0)

// the pressure changed other than by a projection: no history to extrapolate
inline void guess_reset(Ctx &c) {
  c.offn[0] = c.offn[1] = 0;
  c.last = 0;
}

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

// kernel timing scope: records events around the launches in its lifetime
int prof_id(Ctx &c, const char *name);
hipEvent_t prof_event(Ctx &c);
void prof_flush(Ctx &c);
struct ProfScope {
  Ctx &c;
  int id = -1;
  hipEvent_t a = nullptr;
  ProfScope(Ctx &cc, const char *name) : c(cc) {
    if (c.prof.on && (c.prof.only.empty() || c.prof.only == name) && c.prof.sample()) {
      id = prof_id(c, name);
      a = prof_event(c);
      (void)hipEventRecord(a, c.st);
    }
  }
  ~ProfScope() {
    if (id >= 0) {
      hipEvent_t b = prof_event(c);
      (void)hipEventRecord(b, c.st);
      c.prof.pending.push_back({id, a, b});
      if (c.prof.pending.size() > 4096) prof_flush(c);
    }
  }
};

// Times one kernel launch exactly: the events go into the launch itself
// (hipExtLaunchKernelGGL records them at the dispatch's start and end), so
// the dispatch gap before the kernel is not counted as ProfScope would.
struct ProfKernel {
  Ctx &c;
  int id = -1;
  hipEvent_t a = nullptr, b = nullptr;
  ProfKernel(Ctx &cc, const char *name) : c(cc) {
    if (c.prof.on && (c.prof.only.empty() || c.prof.only == name) && c.prof.sample()) {
      id = prof_id(c, name);
      a = prof_event(c);
      b = prof_event(c);
    }
  }
  ~ProfKernel() {
    if (id >= 0) {
      c.prof.pending.push_back({id, a, b});
      if (c.prof.pending.size() > 4096) prof_flush(c);
    }
  }
};
#define SULI_TIMED_LAUNCH(pk, kernel, grid, block, shmem, stream, ...) \
  hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, (pk).a, (pk).b, 0, __VA_ARGS__)

// Poisson solvers (poisson_ref.hip, poisson_fcg.hip)
int poisson_ref_solve(Ctx &c, double dt, suli_poisson_stats *st);
// milliseconds between the two solve events (the second recorded and waited for here)
int solve_elapsed_ms(Ctx &c, double *ms);
int fcg_create(Ctx &c);
void fcg_destroy(Ctx &c);
int fcg_solve(Ctx &c, double dt, suli_poisson_stats *st);
// z = M r (device FP32 frames) with the operator of the last fcg_solve
int fcg_precond_apply(Ctx &c, const float *r, float *z);
int fcg_readbacks(Ctx &c, int64_t *published, int64_t *copied);
