// Reverse-sweep (adjoint) entry points of libmeamz shared by the host units.
#pragma once

struct mz_ctx;

// objective + gradient of ncand knot vectors (F[ncand], G[ncand*nknot])
int mz_grad_impl(mz_ctx* c, int ncand, const double* knots, double* F, double* G);
// releases the gradient buffers and captured gradient graphs of a context
void mz_adjoint_drop(mz_ctx* c);
