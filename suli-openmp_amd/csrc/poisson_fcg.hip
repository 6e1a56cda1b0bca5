// Mixed-precision flexible CG for the projection (placeholder until built).
#include "ctx.hpp"

int fcg_create(Ctx &c) { (void)c; return 0; }
void fcg_destroy(Ctx &c) { (void)c; }
int fcg_solve(Ctx &c, double dt, suli_poisson_stats *st) {
  (void)c; (void)dt; (void)st;
  suli_set_error("SULI_POISSON_FCG is not available in this build");
  return 5;
}
