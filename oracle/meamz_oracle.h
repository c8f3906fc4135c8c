/* CPU restatement of the MEAM/EAM spline objective (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: this library is the checker the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never links or calls it.
 *
 * Every function restates the reference algorithm; the reference file:line it
 * follows is cited at the definition in meamz_oracle.c.
 */
#ifndef MEAMZ_ORACLE_H
#define MEAMZ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_EAM_SPLINE = 0,
  ORC_MEAM_SPLINE = 1,
  ORC_EAM = 2,
  ORC_MEAM = 3,
  ORC_GMEAM_SPLINE = 4,
  ORC_GMEAM = 5,
  ORC_PAIR_SPLINE = 6,
  ORC_PAIR = 7,
  ORC_JMEAM_SPLINE = 8,
  ORC_JMEAM = 9,
  ORC_JMEAM2_SPLINE = 10,
  ORC_JMEAM2 = 11
};

/* One potential's spline layout.  Splines are listed in file order
 * (phi Pair_ij, rho Pair_j, U Pair_i [, f Pair_ij, g Triplet_i; gmeam styles:
 * g Triplet_i_jk, ntypes * ntypes (ntypes + 1) / 2 splines]; jmeam styles add
 * p Pair_ij, q Triplet_i, jmeam2 also p2 Pair_ij, q2 Triplet_i; pair styles:
 * phi Pair_ij only).
 * bc holds two ints per spline: -1 natural, 0 free slope, 1 fixed slope. */
typedef struct {
  int style;
  int ntypes;
  int nspl;
  const int* basis;     /* 3 or 5 */
  const int* nknots;
  const double* x0;
  const double* xn;
  const int* bc;
  int embed_extrap;
} orc_layout;

/* Database of periodic cells (all arrays concatenated over cells). */
typedef struct {
  int ncells;
  const int* natoms;     /* [ncells] */
  const double* box;     /* [ncells][3][3], rows are lattice vectors */
  const int* types;      /* [natoms_total] */
  const double* pos;     /* [natoms_total][3] */
  const double* force0;  /* [natoms_total][3] */
  const double* energy0; /* [ncells] per-atom target energy */
  const double* stress0; /* [ncells][6] xx yy zz xy yz zx */
  const double* weight, *eweight, *sweight; /* [ncells] */
  double energy_weight, stress_weight;
} orc_db;

typedef struct orc_ctx orc_ctx;

/* Knot vector: per spline [y_0..y_{n-1}, grad_L, grad_R, grad2_L, grad2_R]. */

/* Builds pair lists (and counts triplets).  Returns NULL and writes a message
 * into err (if non-NULL, >= 256 bytes) on failure. */
orc_ctx* orc_create(const orc_layout* lay, const orc_db* db, char* err);
void orc_destroy(orc_ctx* ctx);
void orc_counts(const orc_ctx* ctx, long long* npairs, long long* ntriplets);
int orc_errvec_size(const orc_ctx* ctx);

/* One objective evaluation in the reference's loop order.  Any output pointer
 * may be NULL.  cell_energy is per atom, cell_stress divided by volume. */
double orc_eval(orc_ctx* ctx, const double* knots, double* errvec, double* cell_energy,
                double* cell_stress, double* forces);

/* Same objective with atoms split over nthreads workers balanced by triplet
 * count and one sum-reduction at the end (the reference's MPI decomposition,
 * src/atom_vec.cpp:41-97 and src/potential.cpp:446, emulated with threads). */
double orc_eval_mt(orc_ctx* ctx, const double* knots, int nthreads);
/* orc_eval_mt with the outputs of orc_eval (each may be NULL). */
double orc_eval_mt_out(orc_ctx* ctx, const double* knots, int nthreads, double* errvec, double* cell_energy,
                       double* cell_stress, double* forces);

/* Per-atom densities without clamping (PotMEAMSpline::fast_compute_densities). */
void orc_densities(orc_ctx* ctx, const double* knots, double* dens);

/* Single-spline helpers (for unit tests). */
void orc_spline_refresh(int basis, int n, double x0, double xn, const int* bc, const double* rec,
                        double* ypp, double* yp4);
int orc_knotshift(int n, double x0, double xn, double r, double* shift);
double orc_splint_knot(int basis, int n, double x0, double xn, const int* bc, const double* rec,
                       int knot, double shift, double* grad);
double orc_splint_point(int basis, int n, double x0, double xn, const int* bc, const double* rec,
                        double r, double* grad, double* grad2);

#ifdef __cplusplus
}
#endif
#endif
