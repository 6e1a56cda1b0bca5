// C ABI of the SuLi-IC MI355X library: state management, host <-> device
// exchange in the reference's array extents, and the time-step entry points
// that replace the reference subroutines called by 1_principal.f90:64-108.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"

Ctx *g_ctx = nullptr;
static std::string g_err;

void suli_set_error(const std::string &m) { g_err = m; }

#define NEEDCTX()                                              \
  do {                                                         \
    if (!g_ctx) {                                              \
      suli_set_error("library not initialised (suli_init)");   \
      return 2;                                                \
    }                                                          \
  } while (0)

extern "C" const char *suli_version(void) { return "suli-mi355x 0.1 (gfx950)"; }

extern "C" int suli_last_error(char *buf, int len) {
  if (!buf || len <= 0) return 1;
  std::snprintf(buf, (size_t)len, "%s", g_err.c_str());
  return 0;
}

// --------------------------------------------------------------------------
// allocation

static int dalloc(Ctx &c, void **p, size_t bytes) {
  HIPCHK(hipMalloc(p, bytes));
  HIPCHK(hipMemsetAsync(*p, 0, bytes, c.st));
  c.allocs.push_back(*p);
  return 0;
}

static void free_all(Ctx &c) {
  prof_flush(c);
  for (hipEvent_t e : c.prof.pool) (void)hipEventDestroy(e);
  c.prof.pool.clear();
  for (hipEvent_t &e : c.solve_ev)
    if (e) (void)hipEventDestroy(e), e = nullptr;
  for (void *p : c.allocs) (void)hipFree(p);
  c.allocs.clear();
  if (c.ctl_h) (void)hipHostFree(c.ctl_h);
  if (c.res_h) (void)hipHostFree(c.res_h);
  if (c.out_host) (void)hipHostFree(c.out_host);
  if (c.out_dev) (void)hipFree(c.out_dev);
  if (c.st) (void)hipStreamDestroy(c.st);
}

int solve_elapsed_ms(Ctx &c, double *ms) {
  HIPCHK(hipEventRecord(c.solve_ev[1], c.st));
  HIPCHK(hipEventSynchronize(c.solve_ev[1]));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, c.solve_ev[0], c.solve_ev[1]));
  *ms = t;
  return 0;
}

extern "C" int suli_finalize(void) {
  if (!g_ctx) return 0;
  (void)hipStreamSynchronize(g_ctx->st);
  fcg_destroy(*g_ctx);
  free_all(*g_ctx);
  delete g_ctx;
  g_ctx = nullptr;
  return 0;
}

extern "C" int suli_init(const suli_params *p) {
  if (!p) {
    suli_set_error("suli_init: null params");
    return 1;
  }
  if (g_ctx) suli_finalize();
  if (p->nx < 4 || p->ny < 4 || p->nz < 4) {
    suli_set_error("suli_init: every grid dimension must be at least 4 (WENO5 stencils)");
    return 1;
  }
  auto okc = [](int v, bool x) { return (v >= 0 && v <= 3) || (x && v == 4); };
  if (!okc(p->ccx0, true) || !okc(p->ccxf, true) || !okc(p->ccy0, false) || !okc(p->ccyf, false) ||
      p->ccz0 < 1 || p->ccz0 > 3 || p->cczf < 1 || p->cczf > 3) {
    suli_set_error("suli_init: supported wall codes are 0 periodic, 1 free slip, 2 no slip, 3 prescribed, "
                   "4 (x walls only: inflow at x0, outflow at xf); z walls 1, 2 or 3");
    return 1;
  }
  if ((p->ccx0 == 0) != (p->ccxf == 0) || (p->ccy0 == 0) != (p->ccyf == 0)) {
    suli_set_error("suli_init: a periodic wall needs its opposite wall periodic too");
    return 1;
  }
  if (p->precision != SULI_PREC_FP64 && p->precision != SULI_PREC_MIXED) {
    suli_set_error("suli_init: unknown precision");
    return 1;
  }
  if (p->m_turb < 0 || p->m_turb > 2) {
    suli_set_error("suli_init: m_turb must be 0 (none), 1 (Smagorinsky) or 2 (directional Smagorinsky)");
    return 1;
  }
  if (p->poisson != SULI_POISSON_REF && p->poisson != SULI_POISSON_FCG) {
    suli_set_error("suli_init: unknown poisson solver");
    return 1;
  }
  if (p->precond != SULI_PRECOND_JACOBI && p->precond != SULI_PRECOND_GMG && p->precond != SULI_PRECOND_LEARNED) {
    suli_set_error("suli_init: unknown preconditioner");
    return 1;
  }
  {
    // frame offsets are 32-bit (fidx) and the stencils read through 32-bit
    // byte offsets (Fld, common.hpp): a frame field must stay below 2^29
    // elements (2^32 bytes in FP64)
    const int64_t sx = ((p->nx + 3 + 31) / 32) * 32;
    const int64_t tot = 31 + sx * (p->ny + 3) * (int64_t)(p->nz + 3) + 64;
    if (tot >= (int64_t(1) << 29)) {
      suli_set_error("suli_init: grid too large: a field frame of " + std::to_string(tot) +
                     " elements exceeds the 2^29 of the 32-bit byte offsets of the stencils");
      return 1;
    }
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    suli_set_error("suli_init: no HIP device visible");
    return 3;
  }
  Ctx *c = new Ctx();
  c->P = *p;
  if (c->P.max_it <= 0) c->P.max_it = 10000;
  if (c->P.tol_abs <= 0) c->P.tol_abs = 0.0001 / ((double)p->nx * p->ny * p->nz);
  if (c->P.tol_rel <= 0) c->P.tol_rel = 1e-8;
  g_ctx = c;
  HIPCHK(hipSetDevice(p->device));
  if (suli_noalloc_setup(nullptr)) return 1;  // this device's scratch of suli_est_div_noalloc
  HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
  c->prec = p->precision;
  c->esz = c->prec == SULI_PREC_FP64 ? 8 : 4;
  Grid &g = c->g;
  g.nx = p->nx; g.ny = p->ny; g.nz = p->nz;
  g.sx = ((p->nx + 3 + 31) / 32) * 32;
  g.sy = p->ny + 3;
  g.sz = (fidx)g.sx * g.sy;
  g.off = 31;
  g.total = (int64_t)g.off + (int64_t)g.sz * (p->nz + 3) + 64;
  g.dx = p->dx; g.dy = p->dy; g.dz = p->dz;
  g.ccx0 = p->ccx0; g.ccxf = p->ccxf; g.ccy0 = p->ccy0; g.ccyf = p->ccyf; g.ccz0 = p->ccz0; g.cczf = p->cczf;
  g.fast = 1;
  Phys &ph = c->ph;
  ph.rho_f1 = p->rho_f1; ph.rho_f2 = p->rho_f2; ph.mi_f1 = p->mi_f1; ph.mi_f2 = p->mi_f2;
  ph.sigma = p->sigma; ph.alpha1 = p->alpha1; ph.csmag = p->csmag; ph.gx = p->gx; ph.gz = p->gz;
  ph.chezy = p->chezy;
  ph.c23 = std::pow(p->dx * p->dy * p->dz, 2. / 3.);
  ph.dx1 = std::fmax(std::fmax(p->dx, p->dy), p->dz);
  ph.dt1 = 0.1 * ph.dx1;
  ph.m_turb = p->m_turb;
  const size_t fb = (size_t)g.total * c->esz, db = (size_t)g.total * 8;
  void **tf[] = {&c->u, &c->v, &c->w, &c->un, &c->vn, &c->wn, &c->u0, &c->v0, &c->w0, &c->ls, &c->lsA,
                 &c->lsB, &c->S, &c->hs, &c->hsx, &c->hsy, &c->hsz, &c->rho, &c->lsnu, &c->modls,
                 &c->kurv, &c->nut, &c->nrx, &c->nry, &c->nrz, &c->ddx, &c->ddy, &c->ddz, &c->Fu,
                 &c->Fv, &c->Fw, &c->tfu, &c->tfv, &c->tfw, &c->fu0, &c->fv0, &c->fw0, &c->fu1,
                 &c->fv1, &c->fw1, &c->xnf, &c->ynf, &c->znf};
  for (void **q : tf)
    if (dalloc(*c, q, fb)) return 1;
  double **td[] = {&c->prd1, &c->D, &c->sD, &c->q, &c->ax, &c->ay, &c->az, &c->e, &c->r, &c->pp, &c->mp};
  for (double **q : td)
    if (dalloc(*c, (void **)q, db)) return 1;
  if (dalloc(*c, (void **)&c->ctl, sizeof(CGCtl))) return 1;
  if (dalloc(*c, &c->bp, (size_t)bplanes_layout(g.nx, g.ny, g.nz).off[BP_N] * c->esz)) return 1;
  const size_t maxblocks = (size_t)((g.nx + 3 + 63) / 64) * ((g.ny + 3 + 3) / 4) * (g.nz + 3) + 1024;
  if (dalloc(*c, (void **)&c->partials, maxblocks * 4 * sizeof(double))) return 1;
  if (dalloc(*c, (void **)&c->res, RS_NSLOT * sizeof(double))) return 1;
  HIPCHK(hipHostMalloc((void **)&c->ctl_h, sizeof(CGCtl), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void **)&c->res_h, RS_NSLOT * sizeof(double), hipHostMallocDefault));
  HIPCHK(hipEventCreate(&c->solve_ev[0]));
  HIPCHK(hipEventCreate(&c->solve_ev[1]));
  // density and viscosity start as the light fluid until a level set is set
  if (fcg_create(*c)) return 1;
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

// --------------------------------------------------------------------------
// kernel timing

int prof_id(Ctx &c, const char *name) {
  for (size_t q = 0; q < c.prof.names.size(); q++)
    if (c.prof.names[q] == name) return (int)q;
  c.prof.names.push_back(name);
  c.prof.total_ms.push_back(0.0);
  c.prof.launches.push_back(0);
  return (int)c.prof.names.size() - 1;
}

hipEvent_t prof_event(Ctx &c) {
  if (c.prof.pool.empty()) {
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  hipEvent_t e = c.prof.pool.back();
  c.prof.pool.pop_back();
  return e;
}

void prof_flush(Ctx &c) {
  (void)hipStreamSynchronize(c.st);
  for (auto &p : c.prof.pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      c.prof.total_ms[p.id] += ms;
      c.prof.launches[p.id] += 1;
    }
    c.prof.pool.push_back(p.a);
    c.prof.pool.push_back(p.b);
  }
  c.prof.pending.clear();
}

extern "C" int suli_prof_enable(int on) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  prof_flush(c);
  c.prof.on = on != 0;
  c.prof.only.clear();
  c.prof.every = 1;
  c.prof.seen = 0;
  c.prof.names.clear();
  c.prof.total_ms.clear();
  c.prof.launches.clear();
  return 0;
}

extern "C" int suli_prof_select(const char *name, int every) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  prof_flush(c);
  c.prof.only = name ? name : "";
  c.prof.every = every > 1 ? every : 1;
  c.prof.seen = 0;
  return 0;
}

extern "C" int suli_prof_count(void) {
  if (!g_ctx) return -1;
  prof_flush(*g_ctx);
  return (int)g_ctx->prof.names.size();
}

extern "C" int suli_prof_get(int idx, char *name, int len, int64_t *launches, double *total_ms) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  prof_flush(c);
  if (idx < 0 || idx >= (int)c.prof.names.size()) {
    suli_set_error("suli_prof_get: index out of range");
    return 1;
  }
  if (name && len > 0) std::snprintf(name, (size_t)len, "%s", c.prof.names[idx].c_str());
  if (launches) *launches = c.prof.launches[idx];
  if (total_ms) *total_ms = c.prof.total_ms[idx];
  return 0;
}

extern "C" int suli_sync(void) {
  NEEDCTX();
  HIPCHK(hipStreamSynchronize(g_ctx->st));
  return 0;
}

// --------------------------------------------------------------------------
// host <-> device in the reference extents

struct Ext {
  int i0, i1, j0, j1, k0, k1;
};
enum Kind { K_U, K_V, K_W, K_P, K_C, K_FX, K_FY, K_FZ };

static Ext ext_of(const Grid &g, Kind k) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  switch (k) {
    case K_U: return {0, nx + 2, 0, ny + 1, 0, nz + 1};
    case K_V: return {0, nx + 1, 0, ny + 2, 0, nz + 1};
    case K_W: return {0, nx + 1, 0, ny + 1, 0, nz + 2};
    case K_P: return {0, nx + 1, 0, ny + 1, 0, nz + 1};
    case K_C: return {1, nx, 1, ny, 1, nz};
    case K_FX: return {1, nx + 1, 1, ny, 1, nz};
    case K_FY: return {1, nx, 1, ny + 1, 1, nz};
    default: return {1, nx, 1, ny, 1, nz + 1};
  }
}
static int64_t ext_len(const Ext &e) {
  return (int64_t)(e.i1 - e.i0 + 1) * (e.j1 - e.j0 + 1) * (e.k1 - e.k0 + 1);
}

struct FieldRef {
  Kind kind;
  void *dev;
  bool f64;      // stored as double regardless of precision
  bool derived;  // download-only
};

static bool field_ref(Ctx &c, int f, FieldRef &r) {
  switch (f) {
    case SULI_U: r = {K_U, c.u, false, false}; return true;
    case SULI_V: r = {K_V, c.v, false, false}; return true;
    case SULI_W: r = {K_W, c.w, false, false}; return true;
    case SULI_PRD1: r = {K_P, c.prd1, true, false}; return true;
    case SULI_RHO: r = {K_C, c.rho, false, false}; return true;
    case SULI_LS_NU: r = {K_C, c.lsnu, false, false}; return true;
    case SULI_LS: r = {K_C, c.ls, false, false}; return true;
    case SULI_MOD_LS: r = {K_C, c.modls, false, false}; return true;
    case SULI_KURV: r = {K_C, c.kurv, false, false}; return true;
    case SULI_HS: r = {K_C, c.hs, false, false}; return true;
    case SULI_HSX: r = {K_C, c.hsx, false, false}; return true;
    case SULI_HSY: r = {K_C, c.hsy, false, false}; return true;
    case SULI_HSZ: r = {K_C, c.hsz, false, false}; return true;
    case SULI_DDLSDX: r = {K_C, c.ddx, false, false}; return true;
    case SULI_DDLSDY: r = {K_C, c.ddy, false, false}; return true;
    case SULI_DDLSDZ: r = {K_C, c.ddz, false, false}; return true;
    case SULI_NUT: r = {K_C, c.nut, false, false}; return true;
    case SULI_XNUT:  // stored by the directional model, else rebuilt from nut and ls_nu
      if (c.ph.m_turb == 2) r = {K_FX, c.xnf, false, false}; else r = {K_FX, nullptr, false, true};
      return true;
    case SULI_YNUT:
      if (c.ph.m_turb == 2) r = {K_FY, c.ynf, false, false}; else r = {K_FY, nullptr, false, true};
      return true;
    case SULI_ZNUT:
      if (c.ph.m_turb == 2) r = {K_FZ, c.znf, false, false}; else r = {K_FZ, nullptr, false, true};
      return true;
    case SULI_FU: r = {K_FX, c.Fu, false, false}; return true;
    case SULI_FV: r = {K_FY, c.Fv, false, false}; return true;
    case SULI_FW: r = {K_FZ, c.Fw, false, false}; return true;
    case SULI_U0: r = {K_FX, c.u0, false, false}; return true;
    case SULI_V0: r = {K_FY, c.v0, false, false}; return true;
    case SULI_W0: r = {K_FZ, c.w0, false, false}; return true;
    case SULI_TFU: r = {K_FX, c.tfu, false, false}; return true;
    case SULI_TFV: r = {K_FY, c.tfv, false, false}; return true;
    case SULI_TFW: r = {K_FZ, c.tfw, false, false}; return true;
    case SULI_FU0: r = {K_FX, c.fu0, false, false}; return true;
    case SULI_FV0: r = {K_FY, c.fv0, false, false}; return true;
    case SULI_FW0: r = {K_FZ, c.fw0, false, false}; return true;
    case SULI_FU1: r = {K_FX, c.fu1, false, false}; return true;
    case SULI_FV1: r = {K_FY, c.fv1, false, false}; return true;
    case SULI_FW1: r = {K_FZ, c.fw1, false, false}; return true;
    default: return false;
  }
}

// boundary plane n (field id SULI_BXX0 + n) between host doubles in its
// module extents and the device buffer of the field precision
static int plane_io(Ctx &c, int n, double *host, bool up) {
  const BPlanes L = bplanes_layout(c.g.nx, c.g.ny, c.g.nz);
  const int64_t m = (int64_t)L.e1[n] * L.e2[n];
  char *dev = (char *)c.bp + L.off[n] * (int64_t)c.esz;
  std::vector<float> tmp;
  void *h = host;
  if (c.esz == 4) {
    tmp.resize((size_t)m);
    if (up) for (int64_t q = 0; q < m; q++) tmp[q] = (float)host[q];
    h = tmp.data();
  }
  HIPCHK(hipMemcpyAsync(up ? (void *)dev : h, up ? h : (void *)dev, (size_t)m * c.esz,
                        up ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  if (c.esz == 4 && !up) for (int64_t q = 0; q < m; q++) host[q] = tmp[q];
  return 0;
}
static bool is_plane(int field) { return field >= SULI_BXX0 && field < SULI_BXX0 + BP_N; }

extern "C" int64_t suli_field_len(int field) {
  if (!g_ctx) return -1;
  if (is_plane(field)) {
    const BPlanes L = bplanes_layout(g_ctx->g.nx, g_ctx->g.ny, g_ctx->g.nz);
    return (int64_t)L.e1[field - SULI_BXX0] * L.e2[field - SULI_BXX0];
  }
  FieldRef r;
  if (!field_ref(*g_ctx, field, r)) return -1;
  return ext_len(ext_of(g_ctx->g, r.kind));
}

// whole frame field to host doubles
static int frame_to_host(Ctx &c, const void *dev, bool f64, std::vector<double> &out) {
  const Grid &g = c.g;
  out.resize((size_t)g.total);
  if (f64 || c.esz == 8) {
    HIPCHK(hipMemcpyAsync(out.data(), dev, (size_t)g.total * 8, hipMemcpyDeviceToHost, c.st));
  } else {
    std::vector<float> tmp((size_t)g.total);
    HIPCHK(hipMemcpyAsync(tmp.data(), dev, (size_t)g.total * 4, hipMemcpyDeviceToHost, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
    for (size_t q = 0; q < tmp.size(); q++) out[q] = tmp[q];
  }
  HIPCHK(hipStreamSynchronize(c.st));
  return 0;
}
static int host_to_frame(Ctx &c, void *dev, bool f64, const std::vector<double> &in) {
  const Grid &g = c.g;
  if (f64 || c.esz == 8) {
    HIPCHK(hipMemcpyAsync(dev, in.data(), (size_t)g.total * 8, hipMemcpyHostToDevice, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
  } else {
    std::vector<float> tmp((size_t)g.total);
    for (size_t q = 0; q < tmp.size(); q++) tmp[q] = (float)in[q];
    HIPCHK(hipMemcpyAsync(dev, tmp.data(), (size_t)g.total * 4, hipMemcpyHostToDevice, c.st));
    HIPCHK(hipStreamSynchronize(c.st));
  }
  return 0;
}

extern "C" int suli_upload(int field, const double *host, int64_t n) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  c.rho_epoch++;  // the projection rebuilds its operator
  if (is_plane(field)) {
    if (n != suli_field_len(field)) {
      suli_set_error("suli_upload: plane " + std::to_string(field) + " expects " +
                     std::to_string(suli_field_len(field)) + " values, got " + std::to_string(n));
      return 1;
    }
    return plane_io(c, field - SULI_BXX0, const_cast<double *>(host), true);
  }
  FieldRef r;
  if (!field_ref(c, field, r) || r.derived) {
    suli_set_error("suli_upload: field " + std::to_string(field) + " cannot be uploaded");
    return 1;
  }
  const Ext e = ext_of(c.g, r.kind);
  if (n != ext_len(e)) {
    suli_set_error("suli_upload: field " + std::to_string(field) + " expects " + std::to_string(ext_len(e)) +
                   " values, got " + std::to_string(n));
    return 1;
  }
  std::vector<double> fr;
  if (frame_to_host(c, r.dev, r.f64, fr)) return 1;
  int64_t q = 0;
  for (int k = e.k0; k <= e.k1; k++)
    for (int j = e.j0; j <= e.j1; j++)
      for (int i = e.i0; i <= e.i1; i++) fr[(size_t)c.g.idx(i, j, k)] = host[q++];
  if (host_to_frame(c, r.dev, r.f64, fr)) return 1;
  if (field == SULI_U || field == SULI_V || field == SULI_W) c.convdiff_pending = false;
  if (field == SULI_TFU || field == SULI_TFV || field == SULI_TFW) c.tf_on = true;
  if (field == SULI_PRD1) guess_reset(c);  // a new pressure: no history to extrapolate
  return 0;
}

extern "C" int suli_download(int field, double *host, int64_t n) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  if (is_plane(field)) {
    if (n != suli_field_len(field)) {
      suli_set_error("suli_download: plane " + std::to_string(field) + " expects " +
                     std::to_string(suli_field_len(field)) + " values, got " + std::to_string(n));
      return 1;
    }
    return plane_io(c, field - SULI_BXX0, host, false);
  }
  FieldRef r;
  if (!field_ref(c, field, r)) {
    suli_set_error("suli_download: unknown field " + std::to_string(field));
    return 1;
  }
  const Ext e = ext_of(c.g, r.kind);
  if (n != ext_len(e)) {
    suli_set_error("suli_download: field " + std::to_string(field) + " expects " +
                   std::to_string(ext_len(e)) + " values, got " + std::to_string(n));
    return 1;
  }
  const Grid &g = c.g;
  if (r.derived) {
    // face viscosities xnut/ynut/znut = interp*_cf(ls_nu) + interp*_cf(nut)
    // (5_viscosidade.f90:99-101,178-211), rebuilt from the cell arrays
    std::vector<double> nu, nt;
    if (frame_to_host(c, c.lsnu, false, nu) || frame_to_host(c, c.nut, false, nt)) return 1;
    int64_t q = 0;
    for (int k = e.k0; k <= e.k1; k++)
      for (int j = e.j0; j <= e.j1; j++)
        for (int i = e.i0; i <= e.i1; i++) {
          int a = i, b = j, cc = k, da = 0, db = 0, dc = 0, n1;
          if (field == SULI_XNUT) { n1 = g.nx; a = a < 2 ? 2 : (a > n1 ? n1 : a); da = 1; }
          else if (field == SULI_YNUT) { n1 = g.ny; b = b < 2 ? 2 : (b > n1 ? n1 : b); db = 1; }
          else { n1 = g.nz; cc = cc < 2 ? 2 : (cc > n1 ? n1 : cc); dc = 1; }
          const size_t p0 = (size_t)g.idx(a, b, cc), p1 = (size_t)g.idx(a - da, b - db, cc - dc);
          host[q++] = (nu[p0] + nu[p1]) * 0.5 + (nt[p0] + nt[p1]) * 0.5;
        }
    return 0;
  }
  std::vector<double> fr;
  if (frame_to_host(c, r.dev, r.f64, fr)) return 1;
  int64_t q = 0;
  for (int k = e.k0; k <= e.k1; k++)
    for (int j = e.j0; j <= e.j1; j++)
      for (int i = e.i0; i <= e.i1; i++) host[q++] = fr[(size_t)g.idx(i, j, k)];
  return 0;
}

// --------------------------------------------------------------------------
// level set

static int read_res(Ctx &c, int off, int n) {
  HIPCHK(hipMemcpyAsync(c.res_h + off, c.res + off, sizeof(double) * n, hipMemcpyDeviceToHost, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  return 0;
}

static int heavi_curv(Ctx &c, bool curv, double *vol) {
  c.rho_epoch++;  // the density follows the heaviside
  { ProfScope ps_(c, "heaviside_normals");
  DISPATCH(c, Launch<TT>::heavi(c.g, c.ph, c.st, c.ls, c.hs, c.hsx, c.hsy, c.hsz, c.rho, c.lsnu, c.modls,
                                c.nrx, c.nry, c.nrz, c.slot(RS_VOL, 0))); }
  if (curv) { ProfScope ps_(c, "curvature");
    DISPATCH(c, Launch<TT>::curv(c.g, c.st, c.nrx, c.nry, c.nrz, c.ddx, c.ddy, c.ddz, c.kurv)); }
  HIPCHK(hipGetLastError());
  if (vol) {
    if (read_res(c, RS_VOL, 1)) return 1;
    *vol = c.res_h[RS_VOL];
  }
  return 0;
}

extern "C" int suli_level_set_ini(double *vol_ini) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  double v = 0;
  if (heavi_curv(c, true, &v)) return 1;
  c.vol_ini = v;
  if (vol_ini) *vol_ini = v;
  return 0;
}

extern "C" int suli_set_vol_ini(double vol_ini) {
  NEEDCTX();
  if (!(vol_ini > 0)) {
    suli_set_error("suli_set_vol_ini: the reference volume must be positive");
    return 1;
  }
  g_ctx->vol_ini = vol_ini;
  return 0;
}

extern "C" int suli_get_volumes(double *vol_ini, double *vol_ins) {
  NEEDCTX();
  if (vol_ini) *vol_ini = g_ctx->vol_ini;
  if (vol_ins) *vol_ins = g_ctx->vol_ins;
  return 0;
}

extern "C" int suli_init_dambreak(double width, double height, double *vol_ini) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  { ProfScope ps_(c, "init_dambreak"); DISPATCH(c, Launch<TT>::dambreak(c.g, c.st, width, height, c.ls)); }
  HIPCHK(hipGetLastError());
  return suli_level_set_ini(vol_ini);
}

extern "C" int suli_level_set(double dt, suli_ls_stats *st) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  if (!(c.vol_ini > 0)) {
    suli_set_error("suli_level_set: no reference volume; call suli_level_set_ini, suli_init_dambreak "
                   "or suli_set_vol_ini first");
    return 1;
  }
  const Grid &g = c.g;
  // SSP-RK3 advection with the full dt (14_level_set.f90:231-239)
  { ProfScope ps_(c, "ls_advect_rk3"); DISPATCH(c, Launch<TT>::ls_adv(g, c.st, c.ls, c.ls, c.u, c.v, c.w, c.lsA, 1., 0., 1., dt)); }
  { ProfScope ps_(c, "ls_advect_rk3"); DISPATCH(c, Launch<TT>::ls_adv(g, c.st, c.lsA, c.ls, c.u, c.v, c.w, c.lsB, 3. / 4., 1. / 4., 1. / 4., dt)); }
  { ProfScope ps_(c, "ls_advect_rk3"); DISPATCH(c, Launch<TT>::ls_adv(g, c.st, c.lsB, c.ls, c.u, c.v, c.w, c.lsA, 1. / 3., 2. / 3., 2. / 3., dt)); }
  std::swap(c.ls, c.lsA);
  // reinitialisation (:386-509)
  const double dt1 = c.ph.dt1, dx1 = c.ph.dx1;
  // (fast form: the sign function fused with the first stage, the same values)
  if (!g.fast) { ProfScope ps_(c, "reinit_sign"); DISPATCH(c, Launch<TT>::reinit_sign(g, c.ph, c.st, c.ls, c.S)); }
  int il = 0;
  double error = 999.;
  while (error > dt1 * dx1 * dx1) {
    il = il + 1;
    if (il == 1 && g.fast) {
      ProfScope ps_(c, "reinit_rk3");
      DISPATCH(c, Launch<TT>::reinit_first(g, c.ph, c.st, c.ls, c.S, c.lsA, 1. * dt1));
    } else {
      ProfScope ps_(c, "reinit_rk3");
      DISPATCH(c, Launch<TT>::reinit_stage(g, c.ph, c.st, c.ls, c.ls, c.S, c.lsA, 1., 0., 1. * dt1, 0,
                                           c.slot(RS_BAND, 1)));
    }
    { ProfScope ps_(c, "reinit_rk3"); DISPATCH(c, Launch<TT>::reinit_stage(g, c.ph, c.st, c.lsA, c.ls, c.S, c.lsB, 3. / 4., 1. / 4.,
                                         1. / 4. * dt1, 0, c.slot(RS_BAND, 1))); }
    { ProfScope ps_(c, "reinit_rk3"); DISPATCH(c, Launch<TT>::reinit_stage(g, c.ph, c.st, c.lsB, c.ls, c.S, c.lsA, 1. / 3., 2. / 3.,
                                         2. / 3. * dt1, il < 3 ? 1 : 0, c.slot(RS_BAND, 1))); }
    std::swap(c.ls, c.lsA);
    HIPCHK(hipGetLastError());
    if (il == 3) {
      error = 0.;
    } else {
      if (read_res(c, RS_BAND, 2)) return 1;
      const double nr = c.res_h[RS_BAND + 1];
      error = c.res_h[RS_BAND] / nr;  // nr = 0 gives NaN and stops, as in the reference
    }
  }
  // mod_ls1 + heaviside, volume, volume correction (:243-282)
  double vol = 0;
  if (heavi_curv(c, false, &vol)) return 1;
  int nc = 0;
  // (the unsteady manufactured solution corrects nothing, :257)
  while (std::fabs(vol - c.vol_ini) / vol > 0.1 && c.mms.t != 2) {
    { ProfScope ps_(c, "volume_correction"); DISPATCH(c, Launch<TT>::masscorr(g, c.st, c.ls, c.modls, dt1 * (vol - c.vol_ini), c.vol_ini)); }
    if (heavi_curv(c, false, &vol)) return 1;
    nc++;
    if (nc > 1000) {
      suli_set_error("suli_level_set: volume correction did not converge");
      return 4;
    }
  }
  { ProfScope ps_(c, "curvature"); DISPATCH(c, Launch<TT>::curv(g, c.st, c.nrx, c.nry, c.nrz, c.ddx, c.ddy, c.ddz, c.kurv)); }
  HIPCHK(hipGetLastError());
  c.vol_ins = vol;
  if (st) {
    st->vol_ini = c.vol_ini;
    st->vol_ins = vol;
    st->reinit_iters = il;
    st->mass_corr_iters = nc;
  }
  return 0;
}

// --------------------------------------------------------------------------
// predictor

extern "C" int suli_visco(void) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  if (c.ph.m_turb == 1) {
    { ProfScope ps_(c, "visco"); DISPATCH(c, Launch<TT>::visco(c.g, c.ph, c.st, c.u, c.v, c.w, c.rho, c.nut)); }
    HIPCHK(hipGetLastError());
  } else if (c.ph.m_turb == 2) {
    { ProfScope ps_(c, "visco"); DISPATCH(c, Launch<TT>::visco_dir(c.g, c.ph, c.st, c.u, c.v, c.w, c.lsnu, c.xnf, c.ynf, c.znf)); }
    HIPCHK(hipGetLastError());
  }
  return 0;
}

extern "C" int suli_convdiff(void) {
  NEEDCTX();
  g_ctx->convdiff_pending = true;
  return 0;
}

static MomPtrs mom_ptrs(Ctx &c) {
  MomPtrs m;
  m.u = c.u; m.v = c.v; m.w = c.w; m.u0 = c.u0; m.v0 = c.v0; m.w0 = c.w0;
  m.rho = c.rho; m.lsnu = c.lsnu; m.nut = c.nut; m.hsx = c.hsx; m.hsy = c.hsy; m.hsz = c.hsz;
  m.kurv = c.kurv;
  // momentum sources only once some were uploaded (no traffic otherwise)
  m.tfu = c.tf_on ? c.tfu : nullptr;
  m.tfv = c.tf_on ? c.tfv : nullptr;
  m.tfw = c.tf_on ? c.tfw : nullptr;
  m.un = c.un; m.vn = c.vn; m.wn = c.wn; m.u0o = c.u0; m.v0o = c.v0; m.w0o = c.w0;
  m.Fu = nullptr; m.Fv = nullptr; m.Fw = nullptr;
  m.fu0 = m.fv0 = m.fw0 = m.fu1 = m.fv1 = m.fw1 = nullptr;
  const bool dir = c.ph.m_turb == 2;
  m.xnf = dir ? c.xnf : nullptr; m.ynf = dir ? c.ynf : nullptr; m.znf = dir ? c.znf : nullptr;
  m.adv = c.adv;
  m.ku = c.obst ? c.ku_d : nullptr; m.kv = c.obst ? c.kv_d : nullptr; m.kw = c.obst ? c.kw_d : nullptr;
  return m;
}

extern "C" int suli_convdiff_forces(void) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  MomPtrs m = mom_ptrs(c);
  m.u0 = c.u; m.v0 = c.v; m.w0 = c.w;  // stage-2 form: writes only the scratch swap buffers
  m.Fu = c.Fu; m.Fv = c.Fv; m.Fw = c.Fw;
  { ProfScope ps_(c, "convdiff_tempo"); DISPATCH(c, Launch<TT>::momentum(c.g, c.phys_now(), c.st, m, 0.0, MOM_FROM_U0)); }
  HIPCHK(hipGetLastError());
  return 0;
}

static int time_stages(int t_tempo) {  // ntt, 2_iniciais.f90:85-102
  return t_tempo == SULI_TIME_RK3 ? 3 : t_tempo == SULI_TIME_RK2 ? 2 : 1;
}

extern "C" int suli_tempo(int tt, double dt) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  if (!c.convdiff_pending) {
    suli_set_error("suli_tempo: needs a preceding suli_convdiff (the forces are fused into the update)");
    return 1;
  }
  const int ntt = time_stages(c.t_tempo);
  if (tt < 1 || tt > ntt) {
    suli_set_error("suli_tempo: stage " + std::to_string(tt) + " outside 1.." + std::to_string(ntt) +
                   " of the selected time scheme");
    return 1;
  }
  MomPtrs m = mom_ptrs(c);
  int mode = MOM_EULER;
  switch (c.t_tempo) {  // 11_diversos.f90:19-227
    case SULI_TIME_RK2: mode = tt == 1 ? MOM_RK_FIRST : MOM_FROM_U0; break;
    case SULI_TIME_RK3:
      mode = tt == 1 ? MOM_RK_FIRST : tt == 2 ? MOM_RK3_2 : MOM_RK3_3;
      m.fu0 = c.fu0; m.fv0 = c.fv0; m.fw0 = c.fw0;
      m.fu1 = c.fu1; m.fv1 = c.fv1; m.fw1 = c.fw1;
      break;
    case SULI_TIME_AB2:
      mode = c.ab2_started ? MOM_AB2 : MOM_EULER;
      m.fu0 = c.fu0; m.fv0 = c.fv0; m.fw0 = c.fw0;
      break;
    default: break;
  }
  { ProfScope ps_(c, "convdiff_tempo"); DISPATCH(c, Launch<TT>::momentum(c.g, c.phys_now(), c.st, m, dt, mode)); }
  HIPCHK(hipGetLastError());
  std::swap(c.u, c.un);
  std::swap(c.v, c.vn);
  std::swap(c.w, c.wn);
  c.convdiff_pending = false;
  if (c.t_tempo == SULI_TIME_AB2) c.ab2_started = true;
  // kind of stage for the projection's initial guess: first stage of a
  // multi-stage step (dt/2), last stage of the step (dt), other
  c.stage = tt == ntt ? 2 : tt == 1 ? 1 : 0;
  return 0;
}

extern "C" int suli_set_time_scheme(int t_tempo) {
  NEEDCTX();
  g_ctx->rho_epoch++;  // the projection rebuilds its operator
  if (t_tempo < SULI_TIME_EULER || t_tempo > SULI_TIME_AB2) {
    suli_set_error("suli_set_time_scheme: t_tempo must be 0 (Euler), 1 (RK2), 2 (RK3) or 3 (AB2)");
    return 1;
  }
  Ctx &c = *g_ctx;
  c.t_tempo = t_tempo;
  c.ab2_started = false;
  guess_reset(*g_ctx);  // the pressure history belongs to the previous scheme's stages
  return 0;
}

extern "C" int suli_set_advection(int adv_type, int der) {
  NEEDCTX();
  const bool ok = (adv_type == 1 && der >= 1 && der <= 3) || ((adv_type == 2 || adv_type == 3) && der == 1);
  if (!ok) {  // the reference stops on the other pairs (7_advectivo.f90:426-429,649-652)
    suli_set_error("suli_set_advection: adv_type 1 takes der 1, 2 or 3; adv_type 2 and 3 take der 1");
    return 1;
  }
  g_ctx->adv = 10 * adv_type + der;
  return 0;
}

extern "C" int suli_set_obstacle(const int32_t *ku, const int32_t *kv, const int32_t *kw) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  c.rho_epoch++;  // the projection rebuilds its operator
  const Grid &g = c.g;
  if (!ku && !kv && !kw) {
    c.obst = false;
    c.ku_h.clear(); c.kv_h.clear(); c.kw_h.clear();
    return 0;
  }
  if (!ku || !kv || !kw) {
    suli_set_error("suli_set_obstacle: give all of ku, kv, kw (or none to remove the obstacle)");
    return 1;
  }
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  struct M { const int32_t *a; int n0, n1, top; const char *name; } ms[3] = {
      {ku, nx + 3, ny + 2, nz, "ku"}, {kv, nx + 2, ny + 3, nz, "kv"}, {kw, nx + 2, ny + 2, nz + 1, "kw"}};
  const size_t plane = (size_t)g.sx * g.sy;
  std::vector<int> h[3];
  for (int q = 0; q < 3; q++) {
    h[q].assign(plane, 0);
    for (int j = 0; j < ms[q].n1; j++)
      for (int i = 0; i < ms[q].n0; i++) {
        const int32_t x = ms[q].a[i + (int64_t)ms[q].n0 * j];
        if (x < -1 || x > ms[q].top) {
          suli_set_error(std::string("suli_set_obstacle: ") + ms[q].name + "(" + std::to_string(i) + "," +
                         std::to_string(j) + ") = " + std::to_string(x) + " outside -1.." + std::to_string(ms[q].top));
          return 1;
        }
        h[q][(size_t)i + (size_t)g.sx * j] = x;
      }
  }
  int **d[3] = {&c.ku_d, &c.kv_d, &c.kw_d};
  for (int q = 0; q < 3; q++) {
    if (!*d[q]) {
      if (hipMalloc((void **)d[q], plane * sizeof(int)) != hipSuccess) {
        suli_set_error("suli_set_obstacle: out of device memory");
        return 1;
      }
      c.allocs.push_back(*d[q]);
    }
    HIPCHK(hipMemcpyAsync(*d[q], h[q].data(), plane * sizeof(int), hipMemcpyHostToDevice, c.st));
  }
  HIPCHK(hipStreamSynchronize(c.st));
  c.ku_h.assign(ku, ku + (size_t)(nx + 3) * (ny + 2));
  c.kv_h.assign(kv, kv + (size_t)(nx + 2) * (ny + 3));
  c.kw_h.assign(kw, kw + (size_t)(nx + 2) * (ny + 2));
  c.obst = true;
  return 0;
}

extern "C" int suli_time_stages(void) {
  if (!g_ctx) return -1;
  return time_stages(g_ctx->t_tempo);
}

extern "C" int suli_contorno(int nlock) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  if (nlock == 3) {  // only a bottom obstacle acts here (3_contorno.f90:445-490)
    if (!c.obst) return 0;
    { ProfScope ps_(c, "contorno"); DISPATCH(c, Launch<TT>::obst_ls(c.g, c.st, c.ls, c.kw_d)); }
    // heaviside (14_level_set.f90:611); mod_ls is not recomputed there, so
    // the fused mod_ls output goes to the free reinitialisation buffer
    c.rho_epoch++;
    { ProfScope ps_(c, "heaviside_normals");
      DISPATCH(c, Launch<TT>::heavi(c.g, c.ph, c.st, c.ls, c.hs, c.hsx, c.hsy, c.hsz, c.rho, c.lsnu, c.S, c.nrx,
                                    c.nry, c.nrz, c.slot(RS_VOL, 0))); }
    HIPCHK(hipGetLastError());
    return 0;
  }
  if (int rc = mms_planes_now(c)) return rc;  // contorno(1|2) calls mms_bc first (3_contorno.f90:30)
  { ProfScope ps_(c, "contorno");
    DISPATCH(c, Launch<TT>::contorno(c.g, c.st, nlock, c.u, c.v, c.w, c.prd1, c.bp, c.obst ? c.ku_d : nullptr,
                                     c.obst ? c.kv_d : nullptr, c.obst ? c.kw_d : nullptr)); }
  HIPCHK(hipGetLastError());
  return 0;
}

// --------------------------------------------------------------------------
// projection

extern "C" int suli_graddin(double dt, suli_poisson_stats *st) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  if (c.P.poisson == SULI_POISSON_REF) {
    c.rho_epoch++;  // its set-up rewrites the operator arrays the FCG solve keeps
    c.last = 0;     // nor does its pressure continue the FCG guess's history
    return poisson_ref_solve(c, dt, st);
  }
  return fcg_solve(c, dt, st);
}

extern "C" int suli_set_guess(int mode) {
  NEEDCTX();
  if (mode != 0 && mode != 1) {
    suli_set_error("suli_set_guess: mode must be 0 (the current pressure) or 1 (extrapolated for stage 1)");
    return 1;
  }
  g_ctx->guess = mode;
  guess_reset(*g_ctx);
  return 0;
}

extern "C" int suli_set_solver(int poisson, int precond, double tol_rel, int max_it) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  c.rho_epoch++;  // the projection rebuilds its operator
  if (poisson != SULI_POISSON_REF && poisson != SULI_POISSON_FCG) {
    suli_set_error("suli_set_solver: unknown poisson solver");
    return 1;
  }
  if (precond != SULI_PRECOND_JACOBI && precond != SULI_PRECOND_GMG && precond != SULI_PRECOND_LEARNED) {
    suli_set_error("suli_set_solver: unknown preconditioner");
    return 1;
  }
  c.P.poisson = poisson;
  c.P.precond = precond;
  if (tol_rel > 0) c.P.tol_rel = tol_rel;
  if (max_it > 0) c.P.max_it = max_it;
  return 0;
}

extern "C" int suli_set_mg(int64_t fuse_max_cells, int coarse_sweeps, int fuse_fine) {
  NEEDCTX();
  if (fuse_fine > 2) {
    suli_set_error("suli_set_mg: fuse_fine must be 0, 1 or 2 (or negative to keep it)");
    return 1;
  }
  g_ctx->rho_epoch++;  // the projection rebuilds its operator
  if (fuse_max_cells >= 0) g_ctx->mg_fuse_max = fuse_max_cells;
  if (coarse_sweeps >= 0) g_ctx->mg_coarse_sweeps = coarse_sweeps;
  if (fuse_fine >= 0) g_ctx->mg_fuse_fine = fuse_fine;
  return 0;
}

extern "C" int suli_set_fast_stencils(int on) {
  NEEDCTX();
  if (on >= 0) g_ctx->g.fast = on ? 1 : 0;
  return 0;
}

extern "C" int suli_fcg_readbacks(int64_t *published, int64_t *copied) {
  NEEDCTX();
  if (!published || !copied) {
    suli_set_error("suli_fcg_readbacks: null output pointer");
    return 1;
  }
  return fcg_readbacks(*g_ctx, published, copied);
}

extern "C" int suli_precond_apply(const double *r, double *z, int64_t n) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  const Grid &g = c.g;
  if (n != (int64_t)g.nx * g.ny * g.nz) {
    suli_set_error("suli_precond_apply: expects nx*ny*nz = " + std::to_string((int64_t)g.nx * g.ny * g.nz) +
                   " values, got " + std::to_string(n));
    return 1;
  }
  if (!c.fcg) {
    suli_set_error("suli_precond_apply: no FCG solver");
    return 1;
  }
  std::vector<float> fr((size_t)g.total, 0.f);
  int64_t q = 0;
  for (int k = 1; k <= g.nz; k++)
    for (int j = 1; j <= g.ny; j++)
      for (int i = 1; i <= g.nx; i++) fr[(size_t)g.idx(i, j, k)] = (float)r[q++];
  float *dr = nullptr, *dz = nullptr;
  const size_t bytes = (size_t)g.total * sizeof(float);
  HIPCHK(hipMalloc((void **)&dr, bytes));
  if (hipMalloc((void **)&dz, bytes) != hipSuccess) {
    (void)hipFree(dr);
    suli_set_error("suli_precond_apply: out of device memory");
    return 1;
  }
  int rc = 0;
  if (hipMemcpyAsync(dr, fr.data(), bytes, hipMemcpyHostToDevice, c.st) != hipSuccess ||
      hipMemsetAsync(dz, 0, bytes, c.st) != hipSuccess) {
    suli_set_error("suli_precond_apply: copy to the device failed");
    rc = 1;
  }
  if (!rc) rc = fcg_precond_apply(c, dr, dz);
  if (!rc && (hipMemcpyAsync(fr.data(), dz, bytes, hipMemcpyDeviceToHost, c.st) != hipSuccess ||
              hipStreamSynchronize(c.st) != hipSuccess)) {
    suli_set_error("suli_precond_apply: copy from the device failed");
    rc = 1;
  }
  (void)hipFree(dr);
  (void)hipFree(dz);
  if (rc) return 1;
  q = 0;
  for (int k = 1; k <= g.nz; k++)
    for (int j = 1; j <= g.ny; j++)
      for (int i = 1; i <= g.nx; i++) z[q++] = fr[(size_t)g.idx(i, j, k)];
  return 0;
}

extern "C" int suli_posdin(double dt) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  { ProfScope ps_(c, "posdin"); DISPATCH(c, Launch<TT>::posdin(c.g, c.st, dt, c.prd1, c.rho, c.u, c.v, c.w)); }
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int suli_est_div(double *maxdiv) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  { ProfScope ps_(c, "est_div"); DISPATCH(c, Launch<TT>::divmax(c.g, c.st, c.u, c.v, c.w, c.slot(RS_DIV, 2))); }
  HIPCHK(hipGetLastError());
  if (read_res(c, RS_DIV, 1)) return 1;
  if (maxdiv) *maxdiv = c.res_h[RS_DIV];
  return 0;
}

// --------------------------------------------------------------------------
// composed step (1_principal.f90:74-95)

extern "C" int suli_stage(int tt, double dt_base, suli_poisson_stats *st) {
  NEEDCTX();
  // a_dt(tt), 2_iniciais.f90:85-102: dt/2 for the first of several stages
  const double dt = tt == 1 && time_stages(g_ctx->t_tempo) > 1 ? 0.5 * dt_base : 1.0 * dt_base;
  int rc;
  if ((rc = suli_visco())) return rc;
  if ((rc = suli_convdiff())) return rc;
  if ((rc = suli_tempo(tt, dt))) return rc;
  if ((rc = waves_now(*g_ctx))) return rc;  // boundary_waves when wave_t > 0 (1_principal.f90:84)
  if ((rc = suli_contorno(2))) return rc;
  if (g_ctx->mms.t > 0) {  // the manufactured-solution run has no projection (1_principal.f90:89-93)
    if (st) *st = suli_poisson_stats{};
    return 0;
  }
  if ((rc = suli_graddin(dt, st))) return rc;
  if ((rc = suli_posdin(dt))) return rc;
  if ((rc = suli_contorno(1))) return rc;
  return 0;
}

extern "C" int suli_set_time(double t) {
  NEEDCTX();
  g_ctx->t_now = t;
  return 0;
}

extern "C" int suli_set_forcing(double amplitude, double period, double phase) {
  NEEDCTX();
  if (!(period > 0.0)) {
    suli_set_error("suli_set_forcing: period must be positive");
    return 1;
  }
  Ctx &c = *g_ctx;
  c.force_amp = amplitude;
  c.force_period = period;
  c.force_phase = phase;
  return 0;
}

extern "C" int suli_step(double dt, suli_ls_stats *ls, suli_poisson_stats *ps) {
  NEEDCTX();
  g_ctx->t_now += dt;  // t = it*dt of the step (1_principal.f90:65) when started from t = 0
  int rc;
  // manufactured-solution sources: the steady ones at the first step, the
  // unsteady ones every step (1_principal.f90:70-71)
  if (g_ctx->mms.t == 2 || (g_ctx->mms.t == 1 && !g_ctx->mms.src_done))
    if ((rc = mms_sources_now(*g_ctx))) return rc;
  if ((rc = suli_level_set(dt, ls))) return rc;
  if ((rc = suli_contorno(3))) return rc;
  const int ntt = time_stages(g_ctx->t_tempo);
  for (int tt = 1; tt <= ntt; tt++) {
    suli_poisson_stats a{};
    if ((rc = suli_stage(tt, dt, &a))) return rc;
    if (ps) ps[tt - 1] = a;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Checkpoint / restart (restart_salva, restart_ini: 11_diversos.f90:232-329)

// Fields of a binary checkpoint: everything the time step carries from one
// step to the next or reads before recomputing it.
static const int kCkptFields[] = {SULI_U,      SULI_V,      SULI_W,      SULI_PRD1,   SULI_RHO,  SULI_LS_NU,
                                  SULI_LS,     SULI_MOD_LS, SULI_KURV,   SULI_HS,     SULI_HSX,  SULI_HSY,
                                  SULI_HSZ,    SULI_DDLSDX, SULI_DDLSDY, SULI_DDLSDZ, SULI_NUT,  SULI_U0,
                                  SULI_V0,     SULI_W0,     SULI_TFU,    SULI_TFV,    SULI_TFW,  SULI_FU0,
                                  SULI_FV0,    SULI_FW0};
static const char kCkptMagic[8] = {'S', 'U', 'L', 'I', 'C', 'K', 'P', '7'};

struct CkptHeader {
  char magic[8];
  int32_t nx, ny, nz, prec;
  int64_t frame_total;
  int32_t nfields, tf_on;
  double vol_ini, vol_ins, t_now, force_amp, force_period, force_phase;
  int32_t guess, stage, last, t_tempo;
  double hdt;
  int32_t ab2_started, adv;
  int32_t obst, pad_;  // 1: the bottom-obstacle columns ku, kv, kw follow the pressure history
  // manufactured-solution mode; the boundary planes (field precision) follow
  // the obstacle columns, then, in the mode, its pressure errors (nx*ny*nz FP64)
  int32_t mms_t, mms_src_done;
  double mms_a, mms_h0, mms_coef;
  // wave inflow: wave_t, amplitude, period, depth and its coefficients
  int32_t wave_t, pad2_;
  double wave_a, wave_p, wave_h0, wave_wc[14];
  int32_t offn[2];  // the guess's FP32 offsets, first then last stage (after the fields)
};

// Binary checkpoint: the device frames of the state fields as they are (the
// precision of the run, ghosts included) and the scalars of the run, so that
// a run resumed from it continues bit for bit.  It loads into a context of
// the same grid and precision only.
extern "C" int suli_checkpoint_save(const char *path) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  if (!path) {
    suli_set_error("suli_checkpoint_save: null path");
    return 1;
  }
  CkptHeader h{};
  std::memcpy(h.magic, kCkptMagic, 8);
  h.nx = c.g.nx; h.ny = c.g.ny; h.nz = c.g.nz; h.prec = c.prec;
  h.frame_total = c.g.total;
  h.nfields = (int32_t)(sizeof(kCkptFields) / sizeof(int));
  h.tf_on = c.tf_on ? 1 : 0;
  h.vol_ini = c.vol_ini; h.vol_ins = c.vol_ins; h.t_now = c.t_now;
  h.force_amp = c.force_amp; h.force_period = c.force_period; h.force_phase = c.force_phase;
  h.guess = c.guess; h.stage = c.stage; h.last = c.last; h.t_tempo = c.t_tempo;
  h.offn[0] = c.offn[0]; h.offn[1] = c.offn[1];
  h.hdt = c.hdt;
  h.ab2_started = c.ab2_started ? 1 : 0; h.adv = c.adv;
  h.obst = c.obst ? 1 : 0;
  h.mms_t = c.mms.t; h.mms_src_done = c.mms.src_done ? 1 : 0;
  h.mms_a = c.mms.a; h.mms_h0 = c.mms.h0; h.mms_coef = c.mms.coef;
  h.wave_t = c.wave.t; h.wave_a = c.wave.a; h.wave_p = c.wave.p; h.wave_h0 = c.wave.h0;
  for (int q = 0; q < 14; q++) h.wave_wc[q] = c.wave.wc[q];
  HIPCHK(hipStreamSynchronize(c.st));
  FILE *f = std::fopen(path, "wb");
  if (!f) {
    suli_set_error(std::string("suli_checkpoint_save: cannot open ") + path);
    return 1;
  }
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
  std::vector<char> buf;
  for (int q = 0; q < h.nfields && ok; q++) {
    FieldRef r;
    field_ref(c, kCkptFields[q], r);
    const size_t bytes = (size_t)c.g.total * (r.f64 ? 8 : c.esz);
    buf.resize(bytes);
    if (hipMemcpy(buf.data(), r.dev, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
      std::fclose(f);
      suli_set_error("suli_checkpoint_save: device copy failed");
      return 1;
    }
    const int32_t id = kCkptFields[q];
    ok = std::fwrite(&id, 4, 1, f) == 1 && std::fwrite(buf.data(), 1, bytes, f) == bytes;
  }
  // the offsets of the first- and last-stage pressures (the guess's history)
  for (int q = 0; q < h.offn[0] + h.offn[1] && ok; q++) {
    const size_t bytes = (size_t)c.g.total * 4;
    buf.resize(bytes);
    const float *src = q < h.offn[0] ? c.off[0][q] : c.off[1][q - h.offn[0]];
    if (hipMemcpy(buf.data(), src, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
      std::fclose(f);
      suli_set_error("suli_checkpoint_save: device copy failed");
      return 1;
    }
    ok = std::fwrite(buf.data(), 1, bytes, f) == bytes;
  }
  if (h.obst && ok)  // the obstacle columns in the reference extents, as suli_set_obstacle took them
    for (const std::vector<int32_t> *k : {&c.ku_h, &c.kv_h, &c.kw_h})
      ok = ok && std::fwrite(k->data(), 4, k->size(), f) == k->size();
  {
    const size_t pb = (size_t)bplanes_layout(c.g.nx, c.g.ny, c.g.nz).off[BP_N] * c.esz;
    buf.resize(pb);
    if (ok && hipMemcpy(buf.data(), c.bp, pb, hipMemcpyDeviceToHost) != hipSuccess) ok = false;
    ok = ok && std::fwrite(buf.data(), 1, pb, f) == pb;
  }
  if (h.mms_t > 0 && ok) {
    const size_t eb = (size_t)c.g.nx * c.g.ny * c.g.nz * 8;
    buf.resize(eb);
    if (hipMemcpy(buf.data(), c.mms.erro_p, eb, hipMemcpyDeviceToHost) != hipSuccess) ok = false;
    ok = ok && std::fwrite(buf.data(), 1, eb, f) == eb;
  }
  if (std::fclose(f) != 0 || !ok) {
    suli_set_error(std::string("suli_checkpoint_save: write failed on ") + path);
    return 1;
  }
  return 0;
}

extern "C" int suli_checkpoint_load(const char *path) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  c.rho_epoch++;  // the projection rebuilds its operator
  FILE *f = path ? std::fopen(path, "rb") : nullptr;
  if (!f) {
    suli_set_error(std::string("suli_checkpoint_load: cannot open ") + (path ? path : "(null)"));
    return 1;
  }
  CkptHeader h{};
  std::string err;
  if (std::fread(&h, sizeof h, 1, f) != 1 || std::memcmp(h.magic, kCkptMagic, 8) != 0) err = "not a checkpoint";
  else if (h.t_tempo < SULI_TIME_EULER || h.t_tempo > SULI_TIME_AB2) err = "unknown time scheme";
  else if (h.nx != c.g.nx || h.ny != c.g.ny || h.nz != c.g.nz || h.prec != c.prec || h.frame_total != c.g.total)
    err = "grid " + std::to_string(h.nx) + "x" + std::to_string(h.ny) + "x" + std::to_string(h.nz) + " precision " +
          std::to_string(h.prec) + " differs from this context's";
  std::vector<char> buf;
  for (int q = 0; q < h.nfields && err.empty(); q++) {
    int32_t id = -1;
    FieldRef r;
    if (std::fread(&id, 4, 1, f) != 1 || !field_ref(c, id, r) || r.derived) {
      err = "bad field record " + std::to_string(q);
      break;
    }
    const size_t bytes = (size_t)c.g.total * (r.f64 ? 8 : c.esz);
    buf.resize(bytes);
    if (std::fread(buf.data(), 1, bytes, f) != bytes) {
      err = "truncated at field " + std::to_string(id);
      break;
    }
    if (hipMemcpy(r.dev, buf.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) err = "device copy failed";
  }
  if (err.empty() && (h.offn[0] < 0 || h.offn[0] > 3 || h.offn[1] < 0 || h.offn[1] > 3 || h.last < 0 ||
                      h.last > 2))
    err = "bad pressure history";
  for (int q = 0; q < h.offn[0] + h.offn[1] && err.empty(); q++) {
    const size_t bytes = (size_t)c.g.total * 4;
    buf.resize(bytes);
    float *&slot = q < h.offn[0] ? c.off[0][q] : c.off[1][q - h.offn[0]];
    if (!slot) {
      if (hipMalloc((void **)&slot, bytes) != hipSuccess) err = "device allocation failed";
      else c.allocs.push_back(slot);
    }
    if (err.empty() && std::fread(buf.data(), 1, bytes, f) != bytes) err = "truncated at the pressure history";
    if (err.empty() && hipMemcpy(slot, buf.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
      err = "device copy failed";
  }
  const int nx = c.g.nx, ny = c.g.ny;
  std::vector<int32_t> kcol[3];
  const size_t kn[3] = {(size_t)(nx + 3) * (ny + 2), (size_t)(nx + 2) * (ny + 3), (size_t)(nx + 2) * (ny + 2)};
  for (int q = 0; q < 3 && h.obst && err.empty(); q++) {
    kcol[q].resize(kn[q]);
    if (std::fread(kcol[q].data(), 4, kn[q], f) != kn[q]) err = "truncated at the obstacle columns";
  }
  std::vector<char> planes, errp;
  if (err.empty()) {
    planes.resize((size_t)bplanes_layout(c.g.nx, c.g.ny, c.g.nz).off[BP_N] * c.esz);
    if (std::fread(planes.data(), 1, planes.size(), f) != planes.size()) err = "truncated at the boundary planes";
  }
  if (err.empty() && (h.mms_t < 0 || h.mms_t > 2)) err = "unknown manufactured-solution mode";
  if (err.empty() && h.mms_t > 0) {
    errp.resize((size_t)c.g.nx * c.g.ny * c.g.nz * 8);
    if (std::fread(errp.data(), 1, errp.size(), f) != errp.size()) err = "truncated at the pressure errors";
  }
  std::fclose(f);
  if (!err.empty()) {
    suli_set_error(std::string("suli_checkpoint_load: ") + path + ": " + err);
    return 1;
  }
  // the checkpoint's obstacle replaces the context's (none: the obstacle is removed)
  if (h.obst ? suli_set_obstacle(kcol[0].data(), kcol[1].data(), kcol[2].data())
             : suli_set_obstacle(nullptr, nullptr, nullptr))
    return 1;
  HIPCHK(hipMemcpy(c.bp, planes.data(), planes.size(), hipMemcpyHostToDevice));
  if (int rc = suli_mms_set(h.mms_t, h.mms_a, h.mms_h0, h.mms_coef)) return rc;
  if (h.mms_t > 0) HIPCHK(hipMemcpy(c.mms.erro_p, errp.data(), errp.size(), hipMemcpyHostToDevice));
  c.mms.src_done = h.mms_src_done != 0;
  c.wave.t = h.wave_t; c.wave.a = h.wave_a; c.wave.p = h.wave_p; c.wave.h0 = h.wave_h0;
  for (int q = 0; q < 14; q++) c.wave.wc[q] = h.wave_wc[q];
  c.tf_on = h.tf_on != 0;
  c.vol_ini = h.vol_ini; c.vol_ins = h.vol_ins; c.t_now = h.t_now;
  c.force_amp = h.force_amp; c.force_period = h.force_period; c.force_phase = h.force_phase;
  c.guess = h.guess; c.stage = h.stage; c.last = h.last; c.offn[0] = h.offn[0]; c.offn[1] = h.offn[1];
  c.hdt = h.hdt;
  c.t_tempo = h.t_tempo; c.ab2_started = h.ab2_started != 0; c.adv = h.adv;
  c.convdiff_pending = false;
  return 0;
}

// Reference restart file (dados/arquivorestart): list-directed text of the
// arrays and scalars of restart_salva (11_diversos.f90:319-321), in that
// order, with the reference extents (0_modules.f90).  The obstacle columns
// ku, kv, kw are written from the context's obstacle (zeros without one)
// and read back into it through suli_set_obstacle.  The boundary planes
// bxx0..bzzf1 of the prescribed walls go both ways, and so do the momentum
// sources tf_u, tf_v, tf_w (read only when nonzero, restart_ini zeroes them
// when mms_t = 0, 11_diversos.f90:268-273).  Items this library does not
// model are written as the reference holds them in its configuration here
// and skipped on reading: d_max, d_min, ls_m (0), the friction planes ub,
// vb, wb (rebuilt by contorno), prd0 and prd (copies of prd1) and the face
// viscosities (rebuilt by visco).
namespace {
struct RItem {
  int field;          // >= 0: a library field; -1: zeros; -2: scalar; -3 prd1 copy; -4 face viscosity; -5 source
  int n0, n1, n2;     // extents (values written); scalar: 1
  const char *name;
};
std::vector<RItem> restart_items(const Grid &g) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  std::vector<RItem> v = {
      {SULI_U, nx + 3, ny + 2, nz + 2, "u"},     {SULI_V, nx + 2, ny + 3, nz + 2, "v"},
      {SULI_W, nx + 2, ny + 2, nz + 3, "w"},     {-1, nx + 3, ny + 2, 1, "ku"},
      {-1, nx + 2, ny + 3, 1, "kv"},             {-1, nx + 2, ny + 2, 1, "kw"},
      {-3, nx + 2, ny + 2, nz + 2, "prd0"},      {SULI_PRD1, nx + 2, ny + 2, nz + 2, "prd1"},
      {-3, nx + 2, ny + 2, nz + 2, "prd"},       {SULI_RHO, nx, ny, nz, "rho"},
      {SULI_LS_NU, nx, ny, nz, "ls_nu"},         {SULI_LS, nx, ny, nz, "ls"},
      {SULI_MOD_LS, nx, ny, nz, "mod_ls"}};
  const int planes[24][2] = {{ny + 2, nz + 2}, {ny + 2, nz + 2}, {ny + 3, nz + 2}, {ny + 2, nz + 3},   // bxx0 bxx1 bxy0 bxz0
                             {ny + 2, nz + 2}, {ny + 2, nz + 2}, {ny + 3, nz + 2}, {ny + 2, nz + 3},   // bxxf bxxf1 bxyf bxzf
                             {nx + 3, nz + 2}, {nx + 2, nz + 2}, {nx + 2, nz + 2}, {nx + 2, nz + 3},   // byx0 byy0 byy1 byz0
                             {nx + 3, nz + 2}, {nx + 2, nz + 2}, {nx + 2, nz + 2}, {nx + 2, nz + 3},   // byxf byyf byyf1 byzf
                             {nx + 3, ny + 2}, {nx + 2, ny + 3}, {nx + 2, ny + 2}, {nx + 2, ny + 2},   // bzx0 bzy0 bzz0 bzz1
                             {nx + 3, ny + 2}, {nx + 2, ny + 3}, {nx + 2, ny + 2}, {nx + 2, ny + 2}};  // bzxf bzyf bzzf bzzf1
  // the boundary planes of the prescribed walls (wall code 3, the manufactured solution)
  for (int q = 0; q < 24; q++) v.push_back({SULI_BXX0 + q, planes[q][0], planes[q][1], 1, "bc plane"});
  v.push_back({-1, nx + 3, ny + 2, nz + 2, "ub"});
  v.push_back({-1, nx + 2, ny + 3, nz + 2, "vb"});
  v.push_back({-1, nx + 2, ny + 2, nz + 3, "wb"});
  v.push_back({-1, 1, 1, 1, "d_max"});
  v.push_back({-1, 1, 1, 1, "d_min"});
  v.push_back({-1, 1, 1, 1, "ls_m"});
  v.push_back({SULI_HSX, nx, ny, nz, "hsx"});
  v.push_back({SULI_HSY, nx, ny, nz, "hsy"});
  v.push_back({SULI_HSZ, nx, ny, nz, "hsz"});
  v.push_back({SULI_HS, nx, ny, nz, "hs"});
  v.push_back({SULI_KURV, nx, ny, nz, "kurv"});
  v.push_back({-4, nx + 1, ny, nz, "xnut"});
  v.push_back({-4, nx, ny + 1, nz, "ynut"});
  v.push_back({-4, nx, ny, nz + 1, "znut"});
  // momentum sources (-5: read only when nonzero, so that a file without
  // sources leaves the step's convdiff without the source term)
  v.push_back({-5, nx + 1, ny, nz, "tf_u"});
  v.push_back({-5, nx, ny + 1, nz, "tf_v"});
  v.push_back({-5, nx, ny, nz + 1, "tf_w"});
  v.push_back({-2, 1, 1, 1, "vol_ini"});
  v.push_back({-2, 1, 1, 1, "vol_ins"});
  return v;
}
}  // namespace

extern "C" int suli_restart_write(const char *path) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  FILE *f = path ? std::fopen(path, "w") : nullptr;
  if (!f) {
    suli_set_error(std::string("suli_restart_write: cannot open ") + (path ? path : "(null)"));
    return 1;
  }
  std::vector<double> host;
  bool ok = true;
  int col = 0;
  auto put = [&](double x) {
    ok = ok && std::fprintf(f, col == 3 ? " %.17g\n" : " %.17g", x) > 0;
    col = (col + 1) % 4;
  };
  for (const RItem &it : restart_items(c.g)) {
    const int64_t n = (int64_t)it.n0 * it.n1 * it.n2;
    int field = it.field;
    if (field == -3) field = SULI_PRD1;  // prd0, prd: copies of prd1
    if (field == -4) field = it.name[0] == 'x' ? SULI_XNUT : (it.name[0] == 'y' ? SULI_YNUT : SULI_ZNUT);
    if (field == -5) field = it.name[3] == 'u' ? SULI_TFU : (it.name[3] == 'v' ? SULI_TFV : SULI_TFW);
    if (field >= 0) {
      host.resize((size_t)n);
      if (suli_download(field, host.data(), n)) {
        std::fclose(f);
        return 1;
      }
      for (double x : host) put(x);
    } else if (field == -2) {
      put(std::strcmp(it.name, "vol_ini") == 0 ? c.vol_ini : c.vol_ins);
    } else if (c.obst && it.name[0] == 'k' && it.name[2] == 0) {  // ku, kv, kw
      const std::vector<int32_t> &m = it.name[1] == 'u' ? c.ku_h : (it.name[1] == 'v' ? c.kv_h : c.kw_h);
      for (int32_t x : m) put((double)x);
    } else {
      for (int64_t q = 0; q < n; q++) put(0.0);
    }
  }
  if (col) ok = ok && std::fputc('\n', f) != EOF;
  if (std::fclose(f) != 0 || !ok) {
    suli_set_error(std::string("suli_restart_write: write failed on ") + path);
    return 1;
  }
  return 0;
}

// list-directed values: separators are blanks, commas and line ends; "r*c"
// stands for r copies of c
namespace {
struct LdReader {
  FILE *f;
  double rep_val = 0;
  int64_t rep_left = 0;
  bool next(double *x) {
    if (rep_left > 0) {
      rep_left--;
      *x = rep_val;
      return true;
    }
    char tok[128];
    int n = 0, ch;
    while ((ch = std::fgetc(f)) != EOF && (ch == ' ' || ch == '\n' || ch == '\r' || ch == '\t' || ch == ','));
    if (ch == EOF) return false;
    while (ch != EOF && !(ch == ' ' || ch == '\n' || ch == '\r' || ch == '\t' || ch == ',') && n < 127) {
      tok[n++] = (char)ch;
      ch = std::fgetc(f);
    }
    tok[n] = 0;
    for (int q = 0; q < n; q++)
      if (tok[q] == 'd' || tok[q] == 'D') tok[q] = 'e';  // Fortran D exponents
    char *star = std::strchr(tok, '*');
    char *end = nullptr;
    if (star) {
      *star = 0;
      const long long r = std::strtoll(tok, &end, 10);
      const double v = std::strtod(star + 1, &end);
      if (r < 1 || end == star + 1) return false;
      rep_val = v;
      rep_left = r - 1;
      *x = v;
      return true;
    }
    *x = std::strtod(tok, &end);
    return end != tok && *end == 0;
  }
};
}  // namespace

extern "C" int suli_restart_read(const char *path) {
  NEEDCTX();
  Ctx &c = *g_ctx;
  c.rho_epoch++;  // the projection rebuilds its operator
  FILE *f = path ? std::fopen(path, "r") : nullptr;
  if (!f) {
    suli_set_error(std::string("suli_restart_read: cannot open ") + (path ? path : "(null)"));
    return 1;
  }
  LdReader rd{f};
  std::vector<double> host;
  double vol[2] = {0, 0};
  int nvol = 0;
  std::vector<int32_t> km[3];
  bool any_k = false;
  for (const RItem &it : restart_items(c.g)) {
    const int64_t n = (int64_t)it.n0 * it.n1 * it.n2;
    host.resize((size_t)n);
    for (int64_t q = 0; q < n; q++)
      if (!rd.next(&host[(size_t)q])) {
        std::fclose(f);
        suli_set_error(std::string("suli_restart_read: ") + path + ": file ends or holds a non-number in " +
                       it.name + " (value " + std::to_string(q + 1) + " of " + std::to_string(n) + ")");
        return 1;
      }
    if (it.field >= 0 && suli_upload(it.field, host.data(), n)) {
      std::fclose(f);
      return 1;
    }
    if (it.field == -5) {  // tf_u, tf_v, tf_w
      bool nz_ = false;
      for (int64_t q = 0; q < n && !nz_; q++) nz_ = host[(size_t)q] != 0.0;
      const int fld = it.name[3] == 'u' ? SULI_TFU : (it.name[3] == 'v' ? SULI_TFV : SULI_TFW);
      if (nz_ && suli_upload(fld, host.data(), n)) {
        std::fclose(f);
        return 1;
      }
    }
    if (it.field == -2) vol[nvol++] = host[0];
    if (it.field == -1 && it.name[0] == 'k' && it.name[2] == 0) {  // ku, kv, kw
      std::vector<int32_t> &m = it.name[1] == 'u' ? km[0] : (it.name[1] == 'v' ? km[1] : km[2]);
      m.resize((size_t)n);
      for (int64_t q = 0; q < n; q++) {
        m[(size_t)q] = (int32_t)std::lround(host[(size_t)q]);
        any_k = any_k || m[(size_t)q] != 0;
      }
    }
  }
  std::fclose(f);
  // the file's obstacle columns replace the context's (restart_ini reads
  // ku, kv, kw, 11_diversos.f90:243); zero columns leave a context without
  // an obstacle as it is
  if ((c.obst || any_k) && suli_set_obstacle(km[0].data(), km[1].data(), km[2].data())) return 1;
  c.vol_ini = vol[0];
  c.vol_ins = vol[1];
  c.convdiff_pending = false;
  guess_reset(c);
  c.ab2_started = false;  // the reference restarts its step counter: it = 1 takes the Euler update
  return 0;
}
