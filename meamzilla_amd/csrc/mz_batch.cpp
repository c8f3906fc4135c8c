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

// Coordinate search over the tile: candidates per workgroup, then the
// (threads, pairs per chunk) shape, then atoms per chunk, each dimension timed
// with the others held at the best so far (about 12 timings instead of the 54
// combinations).  Each timing is one warm-up and `reps` calls of
// mz_objective_batch with ncand vectors; the work per call does not depend on
// the knot values, so any finite vector times it.
int mz_autotune(mz_ctx* c, int ncand, int reps, mz_tile* best) {
  if (!c || ncand < 1) return MZ_ERR_ARG;
  if (reps < 1) reps = 3;
  mz_info_t info;
  mz_info(c, &info);
  mz_tile bt;
  mz_get_tile(c, &bt);
  std::vector<double> kv((size_t)ncand * info.nknot, 0.0);
  std::vector<double> Fv(ncand);
  auto time_tile = [&](const mz_tile& t, double& dt) {
    if (mz_set_tile(c, &t) != MZ_OK) return false;
    if (mz_objective_batch(c, ncand, kv.data(), Fv.data(), nullptr) != MZ_OK) return false;
    auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; ++r) mz_objective_batch(c, ncand, kv.data(), Fv.data(), nullptr);
    dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / reps;
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
  const mz_tile start = bt;
  for (int ct : ctiles)
    if (ct != start.cand_tile && ct <= ncand) {
      mz_tile t = bt;
      t.cand_tile = ct;
      consider(t);
    }
  const mz_tile after_ct = bt;
  for (auto& sh : shapes)
    if (sh[0] != after_ct.block || sh[1] != after_ct.max_pairs) {
      mz_tile t = bt;
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

}  // extern "C"
