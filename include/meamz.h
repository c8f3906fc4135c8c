/* meamz.h -- C-ABI of the MI355X MEAM/EAM objective library (libmeamz.so).
 *
 * The library evaluates the spline-potential fitting objective of the
 * reference program on one gfx950 GPU: every configuration's pair list is
 * resident in HBM, triplets are enumerated on the fly, and one call returns
 * F = sum_i e_i^2 over the residual vector (energies, stresses, forces,
 * density penalties and the style penalty), optionally with the residual
 * vector, per-cell energies/stresses and per-atom forces.
 *
 * Conventions: plain C types, caller-owned host buffers, int status
 * (MZ_OK = 0, negative on error; mz_last_error() gives the message), one host
 * thread per context.  All doubles are IEEE binary64.
 *
 * Knot vector.  A potential is a list of splines in file order (phi Pair_ij,
 * rho Pair_j, U Pair_i and, for MEAM, f Pair_ij and g Triplet_i; jmeam styles
 * then p Pair_ij, q Triplet_i (and, for jmeam2, p2 Pair_ij, q2 Triplet_i); pair
 * styles hold phi Pair_ij alone and their residual vector has no density or
 * penalty rows).  Each spline
 * contributes n+4 doubles: its n knot values y_0..y_{n-1}, then grad_L, grad_R,
 * grad2_L, grad2_R (boundary slopes and, for quintic splines, boundary second
 * derivatives).  Every evaluation entry point takes such a vector.
 */
#ifndef MEAMZ_H
#define MEAMZ_H

#ifdef __cplusplus
extern "C" {
#endif

#define MZ_OK 0
#define MZ_ERR_ARG (-1)       /* bad argument or shape */
#define MZ_ERR_HIP (-2)       /* HIP runtime failure (no device, launch failure) */
#define MZ_ERR_GEOMETRY (-3)  /* pair shorter than a spline's first knot, cell too small */
#define MZ_ERR_LAYOUT (-4)    /* spline layout does not match the style */
#define MZ_ERR_CAPACITY (-5)  /* a list exceeds a compiled limit (knots, neighbours, images) */
#define MZ_ERR_NOMEM (-6)     /* device or host allocation failed */

/* Potential styles (the reference's PotentialStyle keys). */
#define MZ_STYLE_EAM_SPLINE 0  /* "eam/spline"  */
#define MZ_STYLE_MEAM_SPLINE 1 /* "meam/spline" */
#define MZ_STYLE_EAM 2         /* "eam"  (point-evaluated) */
#define MZ_STYLE_MEAM 3        /* "meam" (point-evaluated) */
#define MZ_STYLE_GMEAM_SPLINE 4 /* "gmeam/spline": g indexed by (type i, types j k) */
#define MZ_STYLE_GMEAM 5       /* "gmeam" (point-evaluated) */
#define MZ_STYLE_PAIR_SPLINE 6 /* "pair/spline": phi only (src/pots/pot_pair_spline/pot_pair_spline.cpp:74-141) */
#define MZ_STYLE_PAIR 7        /* "pair" (point-evaluated, src/pots/pot_pair/pot_pair.cpp:73-144) */
#define MZ_STYLE_JMEAM_SPLINE 8  /* "jmeam/spline": meam plus p_ij p_ik q(cos) energy terms */
#define MZ_STYLE_JMEAM 9         /* "jmeam" (point-evaluated, src/pots/pot_jmeam/pot_jmeam.cpp:101-300) */
#define MZ_STYLE_JMEAM2_SPLINE 10 /* "jmeam2/spline": two p p q terms (p, q and p2, q2) */
#define MZ_STYLE_JMEAM2 11       /* "jmeam2" (point-evaluated, src/pots/pot_jmeam2/pot_jmeam2.cpp:107-330) */

/* Spline layout of one potential.  bc holds two ints per spline:
 * -1 natural, 0 free slope, 1 fixed slope (the "#B" constraint line). */
typedef struct {
  int style;
  int ntypes;
  int nspl;
  const int* basis;    /* [nspl] 3 = spline3eq, 5 = spline5eq */
  const int* nknots;   /* [nspl] */
  const double* x0;    /* [nspl] first knot */
  const double* xn;    /* [nspl] last knot */
  const int* bc;       /* [2*nspl] */
  int embed_extrap;    /* the "embed_extrap" parameter */
} mz_layout;

/* Database of periodic cells; arrays are concatenated over cells in file order. */
typedef struct {
  int ncells;
  const int* natoms;      /* [ncells] */
  const double* box;      /* [ncells*9] lattice vectors #X, #Y, #Z as rows */
  const int* types;       /* [natoms_total] */
  const double* pos;      /* [natoms_total*3] */
  const double* force0;   /* [natoms_total*3] target forces */
  const double* energy0;  /* [ncells] target energy per atom (#E) */
  const double* stress0;  /* [ncells*6] target stress xx yy zz xy yz zx (#S) */
  const double* weight;   /* [ncells] #W */
  const double* eweight;  /* [ncells] #EW */
  const double* sweight;  /* [ncells] #SW */
  double energy_weight;   /* "energy_weight" parameter */
  double stress_weight;   /* "stress_weight" parameter */
} mz_db;

typedef struct {
  int ncells;
  int natoms;
  int nspl;
  int nknot;              /* length of a knot vector */
  int nerr;               /* length of the residual vector */
  long long npairs;       /* full (double-counted) pair list */
  long long ntriplets;    /* sum over atoms of kept j<k neighbour pairs */
  int nchunks;            /* work units of the atom kernel */
  int device;             /* HIP device ordinal */
  double list_ms;         /* wall time of the pair-list build inside mz_create */
  int list_device;        /* 1: pair lists built on the device (mz_create), 0: on the host */
} mz_info_t;

typedef struct mz_ctx mz_ctx;

/* Builds the pair lists on the current HIP device (the reference's
 * Config::init / find_pairs / find_triplets, src/config.cpp:46-285, and
 * AtomVec::init, src/atom_vec.cpp:41-140, re-decomposed into atom chunks: a
 * count pass and a fill pass, one thread per atom, the list left resident),
 * counts triplets and uploads the geometry and targets.
 * Replaces: Config::init src/config.cpp:46-119, Config::find_pairs :143-234,
 * Config::find_triplets :240-285, Cell::read_cell/find_scale src/cell.cpp:39-170. */
int mz_create(const mz_layout* layout, const mz_db* db, mz_ctx** out);

/* mz_create with the pair lists built on the host by the same rules, one host
 * thread per cell (the reference's own route, src/config.cpp:143-234).  The
 * lists, and so every result, are identical to mz_create's; it serves as the
 * cross-check of the device builder.
 * No reference counterpart beyond mz_create's. */
int mz_create_host_lists(const mz_layout* layout, const mz_db* db, mz_ctx** out);
void mz_destroy(mz_ctx* ctx);
int mz_info(const mz_ctx* ctx, mz_info_t* out);
const char* mz_last_error(void);

/* One objective evaluation.  Outputs may be NULL.  errvec has info.nerr
 * entries in the reference layout; cell_energy is per atom, cell_stress is
 * divided by the cell volume, forces is [natoms*3].
 * Replaces: virtual double Potential::compute(const Comm&, ErrorVec*)
 * src/potential.h:67, i.e. PotMEAMSpline::compute / fast_compute
 * src/pots/pot_meam_spline/pot_meam_spline.cpp:56-283, PotEAMSpline
 * src/pots/pot_eam_spline/pot_eam_spline.cpp:43-224, PotMEAM::compute
 * src/pots/pot_meam/pot_meam.cpp:95-275, PotEAM::compute
 * src/pots/pot_eam/pot_eam.cpp:129-251, PotPairSpline / PotPair
 * src/pots/pot_pair_spline/pot_pair_spline.cpp:74-141,
 * src/pots/pot_pair/pot_pair.cpp:73-144, with accumulate_error
 * src/potential.cpp:415-491. */
int mz_objective(mz_ctx* ctx, const double* knots, double* F, double* errvec, double* cell_energy,
                 double* cell_stress, double* forces);

/* mz_objective without any device or host allocation at call time: the
 * per-evaluation buffers (knot vector, spline tables, pair force vectors,
 * per-atom and per-chunk partials, residual outputs) live in a static 64 MiB
 * device array held by every GPU the library runs on until the process ends.
 * Same arguments, kernels and results as mz_objective (no captured graph).
 * A database whose buffers exceed the array (about 24 bytes per pair plus
 * 56 bytes per atom, plus the requested outputs) is rejected with
 * MZ_ERR_CAPACITY.  Every call on one GPU uses the same array: calls that
 * overlap on one GPU race (serialize them, one context stream at a time).
 * No counterpart in the reference: for callers that cannot allocate device
 * memory at evaluation time (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc). */
int mz_objective_noalloc(mz_ctx* ctx, const double* knots, double* F, double* errvec, double* cell_energy,
                         double* cell_stress, double* forces);

/* Per-atom densities before clamping (MZ_ERR_LAYOUT for the pair styles,
 * which have none).
 * Replaces: compute_densities src/pots/pot_meam_spline/pot_meam_spline.cpp:79-96,289-349. */
int mz_densities(mz_ctx* ctx, const double* knots, double* dens);

/* Objective and its gradient with respect to every entry of the knot vector
 * (dF has info.nknot entries; entries of fixed knots and natural boundary
 * slots are still filled).  One reverse sweep over the same lists.
 * No reference counterpart: the reference's Powell forms forward differences,
 * src/opt_powell.cpp:348-389. */
int mz_objective_grad(mz_ctx* ctx, const double* knots, double* F, double* dF);

/* Objective (and, if G is non-NULL, gradient) of ncand knot vectors stored
 * row-major in knots[ncand*nknot], all scored over the whole database in one
 * captured launch sequence.  F[ncand], G[ncand*nknot].
 * No reference counterpart: replaces the per-member compute loops of the GA
 * src/opt_genalg.cpp:143-150,237-265 and the trial loops of simulated
 * annealing src/opt_simann.cpp:120-135 and the line searches. */
int mz_objective_batch(mz_ctx* ctx, int ncand, const double* knots, double* F, double* G);

/* Objective and residual vector of ncand knot vectors (rows of
 * knots[ncand*nknot]) in one batched launch sequence: F[ncand],
 * errvec[ncand*nerr] in the layout of mz_objective.
 * No reference counterpart: replaces the ncoeff+1 sequential residual
 * evaluations of Powell's Jacobian, gamma_init src/opt_powell.cpp:348-389. */
int mz_residuals_batch(mz_ctx* ctx, int ncand, const double* knots, double* F, double* errvec);

/* Work-decomposition tile: atoms per chunk limit, pairs per chunk limit and
 * candidates per block.  mz_autotune times tiles on the current database with
 * ncand all-zero knot vectors per call (every pair of candidates per
 * workgroup and threads and pairs per chunk, then atoms per chunk; the
 * fastest of reps calls per tile) and keeps the fastest.  mz_autotune_knots
 * times with the caller's ncand knot vectors knots[ncand*nknot] and, when grad
 * is nonzero, with objective-and-gradient calls: the time depends a little on
 * the values (stores of all-zero data are cheaper), and gradient calls also
 * store the leg sums the adjoint reads, so representative vectors in the mode
 * the caller will use rank the tiles as the real calls do.  mz_get_tile
 * reports the tile in use.  No reference counterpart (work decomposition). */
typedef struct {
  int max_atoms;        /* atoms per chunk (work unit) */
  int max_pairs;        /* pairs per chunk */
  int block;            /* threads per workgroup */
  int cand_tile;        /* candidates per workgroup in batched mode: 1, 2 or 4 */
} mz_tile;
int mz_set_tile(mz_ctx* ctx, const mz_tile* tile);
int mz_get_tile(const mz_ctx* ctx, mz_tile* tile);
int mz_autotune(mz_ctx* ctx, int ncand, int reps, mz_tile* best);
int mz_autotune_knots(mz_ctx* ctx, int ncand, const double* knots, int grad, int reps, mz_tile* best);

/* Device timing of the dominant kernels over the last call made with timing
 * enabled: HIP events recorded on the context stream around each kernel
 * (with timing on, calls launch on the stream instead of replaying their
 * captured graphs), summed over the launches of the call (a batch larger
 * than the free device memory holds runs as equal sub-batches). */
typedef struct {
  double atom_ms;       /* atom (pair + triplet + embedding + force) kernel */
  double resid_ms;      /* force gather, residual and reduction kernels */
  double refresh_ms;    /* spline refresh kernel */
  double adjoint_ms;    /* reverse sweep (gradient) kernel, if run */
  double total_ms;      /* whole evaluation on the stream */
  int launches;         /* launch sequences the call ran (sub-batches) */
} mz_timing;
int mz_enable_timing(mz_ctx* ctx, int on);
int mz_last_timing(const mz_ctx* ctx, mz_timing* out);

/* Peaks this GPU reaches, for roofline reports beside the datasheet figures:
 * an FP64 FMA loop (eight independent chains per lane, TFLOP/s) and a stream
 * copy of 1 GiB (read + write bytes, GB/s; 4 to 32 workgroups per CU), best
 * of three timed rounds each.
 * No reference counterpart. */
int mz_probe_peaks(double* fp64_tflops, double* copy_gbs);

/* Second derivatives y'' (and, for quintic splines, fourth derivatives y'''')
 * at the knots of one spline from its knot record [y_0..y_{n-1}, grad_L,
 * grad_R, grad2_L, grad2_R], computed on the host with the same algorithms as
 * the device refresh.  Replaces Spline3Eq::refresh_basis
 * (src/spline3eq.cpp:192-225) and Spline5Eq::refresh_basis
 * (src/spline5eq.cpp:210-309) for callers that print splines, such as the
 * LAMMPS export (src/spline3eq.cpp:279-295, src/spline5eq.cpp:358-377).
 * basis is 3 or 5, 3 <= n <= 32, x0 < xn; natural_left / natural_right are 1
 * for a natural end.  ypp and yp4 hold n doubles (yp4 may be NULL for basis 3,
 * whose y'''' are zero).  No device is used. */
int mz_spline_refresh(int basis, int n, double x0, double xn, int natural_left, int natural_right,
                      const double* record, double* ypp, double* yp4);

#ifdef __cplusplus
}
#endif
#endif
