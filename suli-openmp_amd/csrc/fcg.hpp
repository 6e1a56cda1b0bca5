// Interface between the flexible-CG projection and the learned preconditioner.
#pragma once
#include "ctx.hpp"

struct UNetPrecond;  // unet.hip

int unet_create(Ctx &c, UNetPrecond **u);
void unet_destroy(Ctx &c, UNetPrecond *u);
bool unet_ready(const UNetPrecond *u);
// per-solve preparation (log-density input for this rho field)
void unet_prepare(Ctx &c, UNetPrecond *u);
// z = M(r) with the network; kernels return at once when *active == 0
void unet_apply(Ctx &c, UNetPrecond *u, const float *r, float *z, const int *active);
