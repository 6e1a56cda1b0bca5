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
// the solver's network (poisson_fcg.hip)
UNetPrecond *fcg_unet(Ctx &c);
