// Reverse-sweep (adjoint) state of libmeamz.
#pragma once

#include <hip/hip_runtime.h>

#include "mz_kernels.h"

namespace mz {

// Device buffers of the gradient path, sized for `cap` candidates.
struct AdjointState {
  int cap = 0;
  double* rho = nullptr;     // [cap][natoms] unclamped densities
  double* dF = nullptr;      // [cap][natoms] U'(rho)
  double* G = nullptr;       // [cap][npairs*3] three-body leg vectors
  double* ctab = nullptr;    // [cap][ncoef*8] adjoint of the interval records
  double* head = nullptr;    // [cap][nspl*8] adjoint of the end-point data
  double* grad = nullptr;    // [cap][nknot] gradient w.r.t. the knot vector
};

void adjoint_release(AdjointState& s);

}  // namespace mz
