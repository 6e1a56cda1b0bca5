// Interface between the flexible-CG projection and the learned preconditioner.
#pragma once
#include "ctx.hpp"

struct UNetPrecond;  // unet.hip

int unet_create(Ctx &c, UNetPrecond **u);
void unet_destroy(Ctx &c, UNetPrecond *u);
bool unet_ready(const UNetPrecond *u);
// per-solve check and activation buffers; nonzero (error set) when the
// network has no weights or the grid does not fit it
int unet_prepare(Ctx &c, UNetPrecond *u);
// z = M(r) with the network on the scaled FP32 residual (frame layout);
// every kernel returns at once when *active == 0
void unet_apply(Ctx &c, UNetPrecond *u, const float *r, float *z, const int *active);
// append (r, rho) to the training set while capture is armed: iter 0 at the
// start of a refinement pass, iter > 0 inside it (kept every `every`-th);
// nothing is stored while *active == 0
void unet_capture(Ctx &c, UNetPrecond *u, const float *r, const int *active, int iter);
// Multigrid level the network solves: 0 = the network is the whole
// preconditioner (unet_apply); e >= 1 = the V-cycle smooths levels 0..e-1
// and the network replaces the coarse solve of level e (unet_coarse_solve).
int unet_entry(const UNetPrecond *u);
// directions kept by FCG with the learned preconditioner: 4 (FCG(4), each new
// direction A-orthogonalised against the last four) or 1 (FCG(1), as with
// multigrid; entry levels >= 1 only)
int unet_fcg_m(const UNetPrecond *u);
bool unet_capture_armed(const UNetPrecond *u);
// block-mean density of the entry level for this solve (entry >= 1)
int unet_level_density(Ctx &c, UNetPrecond *u);
// x = alpha D^{-1/2} Net(D^{-1/2} b) on a multigrid level (frame of grid ge,
// face conductances cx, cy, cz zero on the walls, diagonal dg), alpha the
// energy-optimal step (y.b)/(y.A y)
void unet_coarse_solve(Ctx &c, UNetPrecond *u, const Grid &ge, const float *cx, const float *cy, const float *cz,
                       const float *dg, const float *b, float *x, const int *active);
// append the level's problem (D^{-1/2} b, block density, faces) to the
// training set while capture is armed at this entry level (iter as in
// unet_capture)
void unet_capture_level(Ctx &c, UNetPrecond *u, const Grid &ge, const float *cx, const float *cy, const float *cz,
                        const float *dg, const float *b, const int *active, int iter);
// the solver's network (poisson_fcg.hip)
UNetPrecond *fcg_unet(Ctx &c);
