// Allocation-free objective evaluation: mz_objective with every per-evaluation
// device buffer (knot vector, spline tables, pair force vectors, per-atom and
// per-chunk partials, residual outputs) carved out of one static device array
// instead of context-owned allocations.  For callers that cannot allocate
// device memory at evaluation time (code under stream capture, serving loops).
//
// The array lives in this file; host code never passes its name to a kernel or
// a copy: the device address is looked up once per device with
// hipGetSymbolAddress at context creation (mz_create), with that device
// current, and kept per device.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

#include "mz_ctx.h"

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

// Per-evaluation scratch of mz_objective_noalloc.  Each device holds its own
// copy for the life of the process; calls on one device share it.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

#define MZ_HIP(call)                                                                 \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) return mz_fail(MZ_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

int mz_run_forward_on(mz_ctx* c, const mz::EvalBufs& b, int ncand, double* d_errvec, double* d_ce, double* d_cs,
                      double* d_rho, double* d_dF, double* d_G, double* d_lam, double* d_cadj, double* d_T1);

// Device address of the scratch array on `dev`, looked up on first use with
// `dev` current (the caller's current device is restored).
int mz_noalloc_setup(int dev) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (dev < 0 || dev >= kMaxDevices) return mz_fail(MZ_ERR_HIP, "device ordinal out of range for the scratch table");
  if (g_ptr[dev]) return MZ_OK;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) return mz_fail(MZ_ERR_HIP, "hipGetDevice failed");
  if (cur != dev && hipSetDevice(dev) != hipSuccess) return mz_fail(MZ_ERR_HIP, "hipSetDevice failed");
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != dev) (void)hipSetDevice(cur);
  if (e != hipSuccess || !p)
    return mz_fail(MZ_ERR_HIP, std::string("hipGetSymbolAddress of the scratch array: ") + hipGetErrorString(e));
  g_ptr[dev] = p;
  return MZ_OK;
}

extern "C" {

int mz_objective_noalloc(mz_ctx* c, const double* knots, double* F, double* errvec, double* cell_energy,
                         double* cell_stress, double* forces) {
  if (!c || !knots) return mz_fail(MZ_ERR_ARG, "null argument");
  for (int q = 0; q < c->nknot; ++q)
    if (!std::isfinite(knots[q])) return mz_fail(MZ_ERR_ARG, "knot vector holds a non-finite value");
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (c->device >= 0 && c->device < kMaxDevices) base = g_ptr[c->device];
  }
  if (!base) {
    // mz_create looks the address up; this covers a failure there
    const int rc = mz_noalloc_setup(c->device);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    base = g_ptr[c->device];
  }
  // carve the buffers, each on a 256-byte boundary
  const size_t nat = (size_t)c->natoms, nch = c->chunks.size();
  mz::EvalBufs b;
  b.nblocks = mz::resid_blocks(c->natoms, c->ncells);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  const size_t o_knots = take(sizeof(double) * c->nknot);
  const size_t o_tables = take((size_t)c->table_bytes);
  const size_t o_forces = take(sizeof(double) * 3 * nat);
  const size_t o_tvec = take(sizeof(double) * 3 * (size_t)c->npairs);
  const size_t o_fown = take(sizeof(double) * 3 * nat);
  const size_t o_acc = take(sizeof(double) * 8 * nch);
  const size_t o_part = take(sizeof(double) * (size_t)b.nblocks);
  const size_t o_F = take(sizeof(double));
  const size_t o_err = errvec ? take(sizeof(double) * (size_t)c->nerr) : 0;
  const size_t o_ce = cell_energy ? take(sizeof(double) * (size_t)c->ncells) : 0;
  const size_t o_cs = cell_stress ? take(sizeof(double) * 6 * (size_t)c->ncells) : 0;
  if (off > kScratchBytes)
    return mz_fail(MZ_ERR_CAPACITY, "database needs " + std::to_string(off) +
                                        " bytes of evaluation scratch; the allocation-free entry point holds " +
                                        std::to_string(kScratchBytes));
  unsigned char* s = static_cast<unsigned char*>(base);
  b.knots = reinterpret_cast<double*>(s + o_knots);
  b.tables = reinterpret_cast<char*>(s + o_tables);
  b.forces = reinterpret_cast<double*>(s + o_forces);
  b.tvec = reinterpret_cast<double*>(s + o_tvec);
  b.fown = reinterpret_cast<double*>(s + o_fown);
  b.chunk_acc = reinterpret_cast<double*>(s + o_acc);
  b.partial = reinterpret_cast<double*>(s + o_part);
  b.F = reinterpret_cast<double*>(s + o_F);
  double* d_err = errvec ? reinterpret_cast<double*>(s + o_err) : nullptr;
  double* d_ce = cell_energy ? reinterpret_cast<double*>(s + o_ce) : nullptr;
  double* d_cs = cell_stress ? reinterpret_cast<double*>(s + o_cs) : nullptr;

  hipStream_t st = c->stream;
  MZ_HIP(hipMemcpyAsync(b.knots, knots, sizeof(double) * c->nknot, hipMemcpyHostToDevice, st));
  const int rc = mz_run_forward_on(c, b, 1, d_err, d_ce, d_cs, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  if (rc) return rc;
  double Fh = 0.0;
  MZ_HIP(hipMemcpyAsync(&Fh, b.F, sizeof(double), hipMemcpyDeviceToHost, st));
  if (errvec) MZ_HIP(hipMemcpyAsync(errvec, d_err, sizeof(double) * c->nerr, hipMemcpyDeviceToHost, st));
  if (cell_energy) MZ_HIP(hipMemcpyAsync(cell_energy, d_ce, sizeof(double) * c->ncells, hipMemcpyDeviceToHost, st));
  if (cell_stress)
    MZ_HIP(hipMemcpyAsync(cell_stress, d_cs, sizeof(double) * 6 * c->ncells, hipMemcpyDeviceToHost, st));
  if (forces) MZ_HIP(hipMemcpyAsync(forces, b.forces, sizeof(double) * 3 * nat, hipMemcpyDeviceToHost, st));
  MZ_HIP(hipStreamSynchronize(st));
  if (F) *F = Fh;
  return MZ_OK;
}

}  // extern "C"
