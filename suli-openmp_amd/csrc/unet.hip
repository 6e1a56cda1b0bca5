// Learned preconditioner (placeholder until the network kernels are built).
#include "fcg.hpp"

struct UNetPrecond {
  int ready = 0;
};

int unet_create(Ctx &c, UNetPrecond **u) {
  (void)c;
  *u = new UNetPrecond();
  return 0;
}
void unet_destroy(Ctx &c, UNetPrecond *u) {
  (void)c;
  delete u;
}
bool unet_ready(const UNetPrecond *u) { return u && u->ready; }
void unet_prepare(Ctx &c, UNetPrecond *u) { (void)c; (void)u; }
void unet_apply(Ctx &c, UNetPrecond *u, const float *r, float *z, const int *active) {
  (void)c; (void)u; (void)r; (void)z; (void)active;
}
