// Kernel argument blocks and launchers of libmeamz (host-visible).
#pragma once

#include <hip/hip_runtime.h>

#include "mz_common.h"

namespace mz {

// Static spline layout in device memory (read by k_refresh).
struct RefreshMeta {
  int nspl;
  const int* nknots;
  const int* basis;
  const int* koff;      // knot-vector offset of each spline
  const double* x0;
  const double* xn;
  const int* bc;
  const int* coef_off;  // first interval record of each spline
};

struct AtomArgs {
  PotDesc pot;
  const Chunk* chunks;
  const int* atom_poff;   // [natoms+1] pair offsets
  const uint2* pairs;     // [npairs] (neighbour, packed image)
  const double* pos;      // [natoms*3]
  const double* box;      // [ncells*9]
  const int* types;       // [natoms]
  const double* weight;   // [ncells]
  const char* tables;     // [ncand][table_bytes]
  int table_bytes;
  int ncand;              // candidates of this launch
  int natoms;
  int ncells;
  long long npairs;
  int nchunks;
  const int* in_pos;      // [npairs] slot of every pair in the incoming-pair index of its neighbour
  double* tvec;           // [ncand][npairs][3] total force vector of every pair, at its incoming slot
  double* fown;           // [ncand][natoms*3] sum of T over each atom's own pairs
  double* chunk_acc;      // [ncand][nchunks][8]: E, S(6), constraint of each chunk
  double* rho_out;        // optional [ncand][natoms] density before clamping
  double* dF_out;         // optional [ncand][natoms] U'(rho)
  double* G_out;          // optional [ncand][npairs][4]: dF-free three-body leg vector G_p and
                          // S1_p = sum_k g f_k (the f'-adjoint factor of the reverse sweep)
  double* T1_out;         // optional, jmeam: [ncand][npairs][2] T1_p = sum_k q f_k (p-term p'-adjoint factor)
};

struct ResidArgs {
  int natoms, ncells, nerr;
  const int* cell_of;
  const int* natoms_c;
  const int* abase;
  const double* weight;
  const double* eweight;
  const double* sweight;
  const double* vol;
  const double* energy0;
  const double* stress0;
  const double* force0;
  double energy_weight, stress_weight;
  long long npairs;
  const int* in_off;            // [natoms+1] CSR of the pairs whose neighbour is each atom
  const double* tvec;           // [ncand][npairs][3] force vector of every pair, in incoming-index order
  const double* fown;           // [ncand][natoms*3] sum of T over each atom's own pairs
  double* forces;               // [ncand][natoms*3] F_a = own sum - incoming sum
  const double* chunk_acc;      // [ncand][nchunks][8]
  const int* cell_chunk_off;    // [ncells+1] CSR of the chunks of every cell
  const int* cell_chunks;
  int nchunks;
  double* errvec;         // optional [ncand][nerr]
  double* cell_energy;    // optional [ncand][ncells]
  double* cell_stress;    // optional [ncand][ncells*6]
  double* partial;        // [ncand][nblocks]
  double* lam;            // optional [ncand][natoms*3]: dF/d(force) = 2 w (F - F0)
  double* cell_adj;       // optional [ncand][ncells*8]: dF/dE_raw, dF/dS_raw (6), dF/dP (constraint)
  int nodens;             // pair styles: no constraint rows in the residual vector
};

struct FinalArgs {
  int npartial;
  const double* partial;
  int meam, quintic;
  int ncells, natoms, nerr, nt, nspl, nknot;
  int f0, nf, u0;
  int fper;               // one f-max penalty per f spline (gmeam) instead of one over all
  int nodens;             // pair styles: no penalty term
  int npgrp;              // max-|y| penalty groups: f, then (jmeam) p and p2
  int pgrp0[3];           // first spline of each group (nf splines each)
  const int* koff;
  const int* nknots;
  const double* knots;
  const char* tables;
  int table_bytes;
  double* errvec;
  double* F;              // [ncand]
};

void launch_refresh(hipStream_t st, const double* knots, int nknot, const RefreshMeta& meta, char* tables,
                    int table_bytes, int ncand);
void launch_zero(hipStream_t st, double* p, size_t n);
// CT candidates per workgroup (1, 2 or 4); returns -1 for an unsupported shape.
int launch_atoms(hipStream_t st, const AtomArgs& a, int nchunks, int ncand, int block, int ppl, int ct);
// Dynamic (spline tables) and static LDS bytes of one atom workgroup.
bool atoms_tables_global(const PotDesc& pot, int ct);  // k_atoms reads spline tables from global memory
size_t atoms_dyn_lds(const PotDesc& pot, int ct);
size_t atoms_static_lds(int block, int ppl, int ct, const PotDesc& pot);
int resid_blocks(int natoms, int ncells);
void launch_resid(hipStream_t st, const ResidArgs& a, int nblocks, int ncand);
void launch_final(hipStream_t st, const FinalArgs& a, int ncand);

}  // namespace mz
