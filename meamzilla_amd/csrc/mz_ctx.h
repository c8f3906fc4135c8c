// Context of libmeamz shared by the host translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/meamz.h"
#include "mz_adjoint.h"
#include "mz_common.h"
#include "mz_kernels.h"

namespace mz {

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  int alloc(size_t count) {
    if (count <= n && p) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (count == 0) count = 1;
    if (hipMalloc((void**)&p, count * sizeof(T)) != hipSuccess) return -1;
    n = count;
    return 0;
  }
  int upload(const T* h, size_t count) {
    if (alloc(count)) return -1;
    if (count && hipMemcpy(p, h, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return -1;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Host copy of one spline's knot grid (Spline3Eq::set_spline src/spline3eq.cpp:354-370).
struct Grid {
  int n;
  double x0, xn, step;
  std::vector<double> x;
  void init(int n_, double l, double r) {
    n = n_;
    x0 = l;
    xn = r;
    step = (r - l) / (n - 1);
    x.assign(n, 0.0);
    x[0] = l;
    x[n - 1] = r;
    for (int i = 1; i < n - 1; ++i) x[i] = l + step * i;
  }
  double xcut() const { return x[n - 1] * (1. + 1.e-8); }
};

// Per-evaluation device buffers of the forward launch sequence (refresh,
// atoms, residual, final), each sized for the candidates of one launch.  The
// context owns one set (allocated on demand); the allocation-free entry point
// carves another out of a static device array.
struct EvalBufs {
  double* knots;      // [ncand][nknot]
  char* tables;       // [ncand][table_bytes]
  double* forces;     // [ncand][natoms*3]
  double* tvec;       // [ncand][npairs*3]
  double* fown;       // [ncand][natoms*3]
  double* chunk_acc;  // [ncand][nchunks][8]
  double* partial;    // [ncand][nblocks]
  double* F;          // [ncand]
  int nblocks;        // residual workgroups per candidate
};

// Device state of the gradient path, sized for `cap` candidates.
struct AdjointExtra {
  bool ready = false;
  int cap = 0;
  int nchunks = 0, ntab = 0, gco0 = 0, ngco = 0, dyn_bytes = 0, grep_log2 = 0, nrad = 0, rrep_log2 = -1, blk = 256,
      rrep_stride = 0;
  const void* fn = nullptr;  // the k_adjoint instantiation of the potential
  int resident = 0;          // its workgroups resident on the whole GPU
  DevBuf<Chunk> d_chunks;                 // chunks of <= blk pairs and blk / 4 atoms
  DevBuf<int> d_owner;                    // owning atom of every pair
  DevBuf<double> d_J;                     // refresh Jacobians
  DevBuf<int> d_joff;
  DevBuf<double> d_rho, d_dF, d_G, d_lam, d_cadj, d_grad;
  DevBuf<double> d_slots;                 // [cand][ntab][workgroup] partial coefficient adjoints
  DevBuf<double> d_tabsum;                // [cand][ntab] their sums
  DevBuf<double> d_T1;  // jmeam: [cap][npairs][2] sum_k q p_k of every pair
  double* h_grad = nullptr;               // pinned [cap][nknot]
  struct Graph {
    hipGraphExec_t ex = nullptr;
    int ncand = 0;
  };
  std::vector<Graph> graphs;
};

}  // namespace mz

struct mz_ctx {
  mz::PotDesc pot{};
  int style = 0, nt = 0, nspl = 0;
  bool meam = false, point = false, pair = false;
  double list_ms = 0.0;  // wall time of the pair-list build in mz_create
  int list_device = 0;   // 1: lists built on the device
  int nfpen = 0;  // f-max penalty terms: 1 (meam), one per f spline (gmeam)
  std::vector<int> basis, nknots, bc, koff, coef_off;
  std::vector<double> x0, xn;
  std::vector<mz::Grid> grids;
  int nknot = 0, ncoef = 0, table_bytes = 0;
  // database
  int ncells = 0, natoms = 0, nerr = 0;
  std::vector<int> natoms_c, abase, cell_of, types;
  std::vector<double> vol;
  // lists
  std::vector<int> atom_poff;
  std::vector<uint2> pairs;
  std::vector<long long> ntrip_atom;
  long long npairs = 0, ntriplets = 0;
  std::vector<mz::Chunk> chunks;
  std::vector<int> in_off, in_list;           // CSR: pairs whose neighbour is each atom
  std::vector<int> in_pos;                    // slot of every pair in that CSR (host-built lists)
  std::vector<int> cell_chunk_off, cell_chunks; // CSR: chunks of every cell
  mz_tile tile{64, 512, 256, 2};
  // device: static
  hipStream_t stream = nullptr;
  int device = 0;
  mz::DevBuf<mz::Chunk> d_chunks;
  mz::DevBuf<int> d_atom_poff, d_types, d_cell_of, d_natoms_c, d_abase;
  mz::DevBuf<uint2> d_pairs;
  mz::DevBuf<int> d_in_off, d_in_list, d_cell_chunk_off, d_cell_chunks;
  mz::DevBuf<int> d_in_pos;  // [npairs] slot of every pair in the incoming-pair index
  mz::DevBuf<double> d_pos, d_box, d_weight, d_eweight, d_sweight, d_vol, d_energy0, d_stress0, d_force0;
  mz::DevBuf<int> d_nknots, d_basis, d_koff, d_bc, d_coef_off;
  mz::DevBuf<double> d_x0, d_xn;
  double energy_weight = 0, stress_weight = 0;
  // device: per evaluation (sized for ncap candidates)
  int ncap = 0;
  mz::DevBuf<double> d_knots, d_forces, d_tvec, d_fown, d_chunk_acc, d_partial, d_F, d_errvec, d_cell_e, d_cell_s, d_rho, d_dF, d_G;
  mz::DevBuf<char> d_tables;
  int resid_blocks = 0;
  // gradient path
  mz::AdjointExtra adjx;
  // pinned staging
  double* h_stage = nullptr;
  size_t h_stage_n = 0;
  // captured launch sequences (forward objective of ncand vectors)
  struct Graph {
    hipGraphExec_t ex = nullptr;
    int ncand = 0;
  };
  std::vector<Graph> graphs;
  double* h_F = nullptr;  // pinned result buffer [ncap]
  // timing
  bool timing = false;
  hipEvent_t ev[8] = {};
  mz_timing last{};
};

// Error reporting shared by the host translation units (message kept per thread).
int mz_fail(int code, const std::string& msg);
