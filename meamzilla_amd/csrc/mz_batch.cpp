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

int mz_autotune(mz_ctx* c, int ncand, int reps, mz_tile* best) {
  if (!c || ncand < 1) return MZ_ERR_ARG;
  if (reps < 1) reps = 3;
  mz_info_t info;
  mz_info(c, &info);
  mz_tile cur;
  mz_get_tile(c, &cur);
  static const int shapes[][2] = {{256, 512}, {256, 256}, {128, 256}, {128, 512}, {512, 512}, {64, 256}};
  static const int atoms[] = {16, 32, 64};
  static const int ctiles[] = {1, 2, 4};
  std::vector<double> kv((size_t)ncand * info.nknot, 0.0);
  // the work per launch does not depend on the knot values: any finite vector times it
  std::vector<double> Fv(ncand);
  double best_t = 1e300;
  mz_tile bt = cur;
  for (auto& sh : shapes)
    for (int na : atoms)
      for (int ct : ctiles) {
        if (ct > 1 && ncand < ct) continue;
        mz_tile t{na, sh[1], sh[0], ct};
        if (mz_set_tile(c, &t) != MZ_OK) continue;
        if (mz_objective_batch(c, ncand, kv.data(), Fv.data(), nullptr) != MZ_OK) continue;
        auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < reps; ++r) mz_objective_batch(c, ncand, kv.data(), Fv.data(), nullptr);
        double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / reps;
        if (dt < best_t) {
          best_t = dt;
          bt = t;
        }
      }
  mz_set_tile(c, &bt);
  if (best) *best = bt;
  return MZ_OK;
}

}  // extern "C"
