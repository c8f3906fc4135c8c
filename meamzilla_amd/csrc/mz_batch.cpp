// Batched evaluation of many knot vectors and the tile autotuner.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/meamz.h"
#include "mz_adjoint.h"

int mz_run_forward(mz_ctx* c, int ncand, double* d_errvec, double* d_ce, double* d_cs, double* d_rho, double* d_dF,
                   double* d_G, double* d_lam, double* d_cadj, double* d_T1);
int mz_stage_knots(mz_ctx* c, int ncand, const double* knots);
int mz_batch_impl(mz_ctx* c, int ncand, const double* knots, double* F, double* G);
int mz_residuals_batch_impl(mz_ctx* c, int ncand, const double* knots, double* F, double* errvec);
int mz_ensure_buffers(mz_ctx* c, int ncand);

extern "C" {

int mz_objective_batch(mz_ctx* c, int ncand, const double* knots, double* F, double* G) {
  return mz_batch_impl(c, ncand, knots, F, G);
}

int mz_residuals_batch(mz_ctx* c, int ncand, const double* knots, double* F, double* errvec) {
  return mz_residuals_batch_impl(c, ncand, knots, F, errvec);
}

}  // extern "C"

namespace {

// Search over the tile: every (candidates per workgroup, threads and pairs
// per chunk) pair, then atoms per chunk with the best of those (about 20
// timings instead of the 54 combinations).  Each timing is one warm-up call and the fastest of `reps`
// calls of mz_objective_batch with the ncand knot vectors `knots` (zeros when
// null), objective only or, with grad, objective and gradient (the forward
// kernel then also stores the leg sums the adjoint reads).  The arithmetic
// does not depend on the knot values but the time does, a little: stores of
// all-zero data are cheaper, which on the ternary workload ranks the tiles
// differently, so callers that have representative vectors pass them.
int autotune(mz_ctx* c, int ncand, const double* knots, bool grad, int reps, mz_tile* best) {
  if (!c || ncand < 1) return MZ_ERR_ARG;
  if (reps < 1) reps = 3;
  mz_info_t info;
  mz_info(c, &info);
  mz_tile bt;
  mz_get_tile(c, &bt);
  std::vector<double> kv((size_t)ncand * info.nknot, 0.0);
  if (knots) std::memcpy(kv.data(), knots, kv.size() * sizeof(double));
  std::vector<double> Fv(ncand);
  std::vector<double> Gv(grad ? (size_t)ncand * info.nknot : 0);
  double* G = grad ? Gv.data() : nullptr;
  auto time_tile = [&](const mz_tile& t, double& dt) {
    if (mz_set_tile(c, &t) != MZ_OK) return false;
    if (mz_objective_batch(c, ncand, kv.data(), Fv.data(), G) != MZ_OK) return false;
    dt = 1e300;
    for (int r = 0; r < reps; ++r) {
      auto t0 = std::chrono::steady_clock::now();
      if (mz_objective_batch(c, ncand, kv.data(), Fv.data(), G) != MZ_OK) return false;
      const double d = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (d < dt) dt = d;
    }
    return true;
  };
  double best_t = 1e300;
  if (!time_tile(bt, best_t)) best_t = 1e300;
  auto consider = [&](mz_tile t) {
    double dt;
    if (time_tile(t, dt) && dt < best_t) {
      best_t = dt;
      bt = t;
    }
  };
  static const int ctiles[] = {1, 2, 4};
  static const int shapes[][2] = {{256, 512}, {256, 256}, {128, 256}, {128, 512}, {512, 512}, {64, 256}};
  static const int atoms[] = {16, 32, 64};
  // candidates per workgroup and shape together: their best pairs differ
  // (registers per lane grow with both), then atoms per chunk
  const mz_tile start = bt;
  for (int ct : ctiles)
    for (auto& sh : shapes)
      if (ct <= ncand && (ct != start.cand_tile || sh[0] != start.block || sh[1] != start.max_pairs)) {
        mz_tile t = start;
        t.cand_tile = ct;
        t.block = sh[0];
        t.max_pairs = sh[1];
        consider(t);
      }
  const mz_tile after_shape = bt;
  for (int na : atoms)
    if (na != after_shape.max_atoms) {
      mz_tile t = bt;
      t.max_atoms = na;
      consider(t);
    }
  mz_set_tile(c, &bt);
  if (best) *best = bt;
  return MZ_OK;
}

}  // namespace

extern "C" {

int mz_autotune(mz_ctx* c, int ncand, int reps, mz_tile* best) {
  return autotune(c, ncand, nullptr, false, reps, best);
}

int mz_autotune_knots(mz_ctx* c, int ncand, const double* knots, int grad, int reps, mz_tile* best) {
  return autotune(c, ncand, knots, grad != 0, reps, best);
}

}  // extern "C"
