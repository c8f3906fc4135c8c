// Reverse sweep: gradient of the objective with respect to the knot vector.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/meamz.h"
#include "mz_adjoint.h"

namespace mz {

void adjoint_release(AdjointState& s) {
  double** bufs[] = {&s.rho, &s.dF, &s.G, &s.ctab, &s.head, &s.grad};
  for (double** b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  s.cap = 0;
}

}  // namespace mz

int mz_grad_impl(mz_ctx* c, int ncand, const double* knots, double* F, double* G) {
  (void)c; (void)ncand; (void)knots; (void)F; (void)G;
  return MZ_ERR_ARG;
}

extern "C" int mz_objective_grad(mz_ctx* c, const double* knots, double* F, double* dF) {
  return mz_grad_impl(c, 1, knots, F, dF);
}
