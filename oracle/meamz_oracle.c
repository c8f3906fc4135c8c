/* CPU restatement of the MEAM/EAM spline objective -- the oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP path is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it.  The product path never links or calls it.
 *
 * It restates, in the reference's loop order, (file:line relative to the
 * reference tree):
 *   - equidistant cubic / quintic splines: set_spline src/spline3eq.cpp:354-370,
 *     refresh_basis src/spline3eq.cpp:192-225 and src/spline5eq.cpp:210-309
 *     (own dense LU with partial pivoting in place of LAPACK dgesv_),
 *     splint by knot src/spline3eq.h:110-122, src/spline5eq.h:112-128,
 *     splint by point src/spline3eq.h:204-236, src/spline5eq.h:218-254,
 *     knot bisection src/spline.cpp:74-99 (get_knotshift), max |y| :58-67;
 *   - pair lists: Config::find_pairs src/config.cpp:143-234, Cell::find_scale
 *     src/cell.cpp:158-170, Pair::setup_pair src/pair.cpp:35-42,
 *     check_pair src/pots/pot_meam_spline/pair_meam_spline.cpp:50-99,
 *     src/pots/pot_eam_spline/pair_eam_spline.cpp:50-86,
 *     src/pots/pot_meam/pair_meam.cpp:50-96, src/pots/pot_eam/pair_eam.cpp:50-85,
 *     src/pots/pot_pair_spline/pair_pair_spline.cpp:50-73,
 *     src/pots/pot_pair/pair_pair.cpp:50-71 (pair styles: half list, j from i
 *     on, src/config.cpp:179-184);
 *   - triplets (enumerated on the fly, same j<k order): find_triplets
 *     src/config.cpp:240-285, check_triplet
 *     src/pots/pot_meam_spline/triplet_meam_spline.cpp:49-73,
 *     src/pots/pot_meam/triplet_meam.cpp:48-71;
 *   - objective: PotMEAMSpline::fast_compute
 *     src/pots/pot_meam_spline/pot_meam_spline.cpp:102-283,
 *     PotEAMSpline::fast_compute src/pots/pot_eam_spline/pot_eam_spline.cpp:100-224,
 *     PotMEAM::compute src/pots/pot_meam/pot_meam.cpp:95-275, PotEAM::compute
 *     src/pots/pot_eam/pot_eam.cpp:129-251, PotPairSpline::fast_compute
 *     src/pots/pot_pair_spline/pot_pair_spline.cpp:74-141, PotPair::compute
 *     src/pots/pot_pair/pot_pair.cpp:73-144, PotJMEAM::compute
 *     src/pots/pot_jmeam/pot_jmeam.cpp:101-300, PotJMEAMSpline::fast_compute
 *     src/pots/pot_jmeam_spline/pot_jmeam_spline.cpp:110-300, PotJMEAM2::compute
 *     src/pots/pot_jmeam2/pot_jmeam2.cpp:107-330, accumulate_error src/potential.cpp:415-491;
 *   - alloy indices src/pot_fns.cpp:132-173.
 */
#include "meamz_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAXK 64
#define DUMMY_WEIGHT 100.0 /* src/pots/pot_pair/pot_pair.h:39 */

typedef struct {
  int basis, n;
  int nat[2];
  double step, invstep;
  double x[ORC_MAXK], y[ORC_MAXK], ypp[ORC_MAXK], yp4[ORC_MAXK];
  double grad[2], grad2[2];
} spl_t;

typedef struct {
  int j;              /* global index of neighbour atom */
  double dist[3];     /* unit vector i -> j */
  double r, invr;
  int phi_idx, rho_idx, f_idx;
  int phi_knot, rho_knot, f_knot; /* -1 outside; spline styles hold the knot */
  double phi_shift, rho_shift, f_shift;
  int p_idx[2], p_knot[2];        /* jmeam styles: p (and p2) */
  double p_shift[2];
} pair_t;

struct orc_ctx {
  int style, nt, nspl, embed_extrap;
  int* basis; int* nknots; double* x0; double* xn; int* bc;
  int* koff;          /* knot-vector offset of each spline */
  int fn_off[9];      /* first spline of each function (phi rho U f g p q p2 q2) */
  int ncells, natoms;
  int* natoms_c; int* cell_of; int* types;
  double* force0; double* energy0; double* stress0; double* weight; double* eweight; double* sweight;
  double* vol;
  double energy_weight, stress_weight;
  long long* poff;    /* pair offsets per atom (CSR) */
  pair_t* pairs;
  long long npairs, ntriplets;
  long long* ntrip_atom;
  spl_t* spl;         /* refreshed splines for the current knot vector */
};

/* ------------------------------------------------------------------ splines */

/* Spline3Eq::set_spline, src/spline3eq.cpp:354-370 */
static void spl_setup(spl_t* s, int basis, int n, double left, double right, const int* bc) {
  s->basis = basis;
  s->n = n;
  s->step = (right - left) / (n - 1);
  s->invstep = 1. / s->step;
  s->x[0] = left;
  s->x[n - 1] = right;
  for (int i = 1; i < n - 1; ++i) s->x[i] = left + s->step * i;
  s->nat[0] = bc[0] == -1;
  s->nat[1] = bc[1] == -1;
}

/* dense LU with partial pivoting, solves A x = b in place (A row-major n x n) */
static void dense_solve(int n, double* A, double* b) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double m = fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[i * n + k]) > m) { m = fabs(A[i * n + k]); p = i; }
    if (p != k) {
      for (int c = 0; c < n; ++c) { double t = A[k * n + c]; A[k * n + c] = A[p * n + c]; A[p * n + c] = t; }
      double t = b[k]; b[k] = b[p]; b[p] = t;
    }
    double piv = A[k * n + k];
    for (int i = k + 1; i < n; ++i) {
      double l = A[i * n + k] / piv;
      if (l == 0.0) continue;
      A[i * n + k] = 0.0;
      for (int c = k + 1; c < n; ++c) A[i * n + c] -= l * A[k * n + c];
      b[i] -= l * b[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    double s = b[k];
    for (int c = k + 1; c < n; ++c) s -= A[k * n + c] * b[c];
    b[k] = s / A[k * n + k];
  }
}

/* Spline3Eq::refresh_basis src/spline3eq.cpp:192-225 (tridiagonal sweep) and
 * Spline5Eq::refresh_basis src/spline5eq.cpp:210-309 (2n x 2n system). */
static void spl_refresh(spl_t* s) {
  int n = s->n;
  if (s->basis == 3) {
    double u[ORC_MAXK];
    memset(u, 0, sizeof(u));
    double p, qn, un;
    if (s->nat[0]) {
      s->ypp[0] = 0;
    } else {
      s->ypp[0] = -0.5;
      u[0] = (3.0 * s->invstep) * ((s->y[1] - s->y[0]) * s->invstep - s->grad[0]);
    }
    for (int i = 1; i < n - 1; ++i) {
      p = 0.5 * s->ypp[i - 1] + 2.0;
      s->ypp[i] = (-0.5) / p;
      u[i] = (s->y[i + 1] - s->y[i]) * s->invstep - (s->y[i] - s->y[i - 1]) * s->invstep;
      u[i] = (6.0 * u[i] * 0.5 * s->invstep - 0.5 * u[i - 1]) / p;
    }
    if (s->nat[1]) {
      qn = un = 0.0;
    } else {
      qn = 0.5;
      un = (3.0 * s->invstep) * (s->grad[1] - (s->y[n - 1] - s->y[n - 2]) * s->invstep);
    }
    s->ypp[n - 1] = (un - qn * u[n - 2]) / (qn * s->ypp[n - 2] + 1.0);
    for (int k = n - 2; k >= 0; --k) s->ypp[k] = s->ypp[k] * s->ypp[k + 1] + u[k];
    for (int k = 0; k < n; ++k) s->yp4[k] = 0.0;
    return;
  }
  int nd = 2 * n;
  double* A = (double*)calloc((size_t)nd * nd, sizeof(double));
  double* b = (double*)calloc((size_t)nd, sizeof(double));
#define AA(r, c) A[(r) * nd + (c)]
  for (int i = 1; i < n - 1; ++i) {
    AA(i, i - 1) = 1. / 12; AA(i, i) = 10. / 12; AA(i, i + 1) = 1. / 12;
    AA(i, n + i - 1) = 1. / 180; AA(i, n + i) = -1. / 90; AA(i, n + i + 1) = 1. / 180;
    AA(n + i, i - 1) = 1.; AA(n + i, i) = -2.; AA(n + i, i + 1) = 1.;
    AA(n + i, n + i - 1) = 1. / 6; AA(n + i, n + i) = 2. / 3; AA(n + i, n + i + 1) = 1. / 6;
    b[i] = s->y[i + 1] - 2 * s->y[i] + s->y[i - 1];
    b[n + i] = 0.;
  }
  if (s->nat[0]) {
    AA(0, 0) = -1.; AA(0, 1) = 1.; AA(0, n) = 1. / 3; AA(0, n + 1) = 1. / 6; b[0] = 0.;
    AA(n, n) = 1.; b[n] = 0.;
  } else {
    AA(0, 0) = 1. / 3; AA(0, 1) = 1. / 6; AA(0, n) = 1. / 45; AA(0, n + 1) = 7. / 360;
    b[0] = s->y[1] - s->y[0] - s->grad[0] * s->step;
    AA(n, 0) = 1.; b[n] = s->grad2[0] * (s->step * s->step);
  }
  if (s->nat[1]) {
    AA(n - 1, n - 2) = -1.; AA(n - 1, n - 1) = 1.; AA(n - 1, 2 * n - 2) = -1. / 6; AA(n - 1, 2 * n - 1) = -1. / 3;
    b[n - 1] = 0.;
    AA(2 * n - 1, 2 * n - 1) = 1.; b[2 * n - 1] = 0.;
  } else {
    AA(n - 1, n - 2) = -1. / 6; AA(n - 1, n - 1) = -1. / 3; AA(n - 1, 2 * n - 2) = -7. / 360; AA(n - 1, 2 * n - 1) = -1. / 45;
    b[n - 1] = s->y[n - 1] - s->y[n - 2] - s->grad[1] * s->step;
    AA(2 * n - 1, n - 1) = 1.; b[2 * n - 1] = s->grad2[1] * (s->step * s->step);
  }
#undef AA
  dense_solve(nd, A, b);
  for (int i = 0; i < n; ++i) {
    s->ypp[i] = b[i] * (s->invstep * s->invstep);
    s->yp4[i] = -b[n + i] * (s->invstep * s->invstep * s->invstep * s->invstep);
  }
  free(A);
  free(b);
}

/* Spline::get_knotshift src/spline.cpp:221-246 */
static int spl_knotshift(const spl_t* s, double r, double* shift) {
  int knot = -1;
  if (r < s->x[s->n - 1] * (1. + 1.e-8)) {
    int klo = 0, khi = s->n - 1;
    while (khi - klo > 1) {
      int k = (khi + klo) >> 1;
      if (s->x[k] > r) khi = k; else klo = k;
    }
    knot = klo;
    double h = s->x[khi] - s->x[klo];
    *shift = (r - s->x[klo]) / h;
  }
  return knot;
}

/* splint_comb by knot: src/spline3eq.h:110-122, src/spline5eq.h:112-128 */
static double spl_comb_knot(const spl_t* s, int k, double shift, double* grad) {
  double a = 1.0 - shift;
  double p1 = s->y[k], d21 = s->ypp[k];
  double p2 = s->y[k + 1], d22 = s->ypp[k + 1];
  double h = s->step;
  if (s->basis == 3) {
    *grad = (p2 - p1) * s->invstep + ((3 * shift * shift - 1) * d22 - (3 * a * a - 1) * d21) * h / 6.0;
    return a * p1 + shift * p2 + ((a * a * a - a) * d21 + (shift * shift * shift - shift) * d22) * h * h / 6.0;
  }
  double d41 = s->yp4[k], d42 = s->yp4[k + 1];
  double t = shift;
  *grad = (p2 - p1) * s->invstep + ((3 * t * t - 1) * d22 - (3 * a * a - 1) * d21) * h / 6.0 +
          ((15 * t * t * t * t - 30 * t * t + 7) * d42 - (15 * a * a * a * a - 30 * a * a + 7) * d41) * (h * h * h) / 360.0;
  return a * p1 + t * p2 + ((a * a * a - a) * d21 + (t * t * t - t) * d22) * (h * h) / 6.0 +
         ((3 * a * a * a * a * a - 10 * a * a * a + 7 * a) * d41 + (3 * t * t * t * t * t - 10 * t * t * t + 7 * t) * d42) *
             (h * h * h * h) / 360.0;
}

/* splint_comb / splint_grad2 by point: src/spline3eq.h:204-266, src/spline5eq.h:218-286 */
static double spl_comb_point(const spl_t* s, double r, double* grad, double* grad2) {
  int n = s->n;
  if (r < s->x[0]) {
    double d = r - s->x[0];
    if (s->basis == 3) {
      *grad = s->grad[0];
      if (grad2) *grad2 = 0.0;
      return s->y[0] + s->grad[0] * d;
    }
    *grad = s->grad[0] + s->grad2[0] * d;
    if (grad2) *grad2 = s->grad2[0];
    return s->y[0] + s->grad[0] * d + 0.5 * s->grad2[0] * d * d;
  } else if (r > s->x[n - 1]) {
    double d = r - s->x[n - 1];
    if (s->basis == 3) {
      *grad = s->grad[1];
      if (grad2) *grad2 = 0.0;
      return s->y[n - 1] + s->grad[1] * d;
    }
    *grad = s->grad[1] + s->grad2[1] * d;
    if (grad2) *grad2 = s->grad2[1];
    return s->y[n - 1] + s->grad[1] * d + 0.5 * s->grad2[1] * d * d;
  }
  int k;
  double delta_r = r - s->x[0];
  if (delta_r < 0) {
    k = 0;
  } else {
    k = (int)(delta_r * s->invstep);
    if (k >= n - 1) k = n - 2;
  }
  double b = (delta_r - k * s->step) * s->invstep;
  double a = 1.0 - b;
  double p1 = s->y[k], d21 = s->ypp[k];
  double p2 = s->y[k + 1], d22 = s->ypp[k + 1];
  double h = s->step;
  if (s->basis == 3) {
    *grad = (p2 - p1) * s->invstep + ((3 * (b * b) - 1) * d22 - (3 * (a * a) - 1) * d21) * h / 6.0;
    if (grad2) *grad2 = a * d21 + b * d22;
    return a * p1 + b * p2 + ((a * a * a - a) * d21 + (b * b * b - b) * d22) * (h * h) / 6.0;
  }
  double d41 = s->yp4[k], d42 = s->yp4[k + 1];
  *grad = (p2 - p1) * s->invstep + ((3 * b * b - 1) * d22 - (3 * a * a - 1) * d21) * h / 6.0 +
          ((15 * b * b * b * b - 30 * b * b + 7) * d42 - (15 * a * a * a * a - 30 * a * a + 7) * d41) * (h * h * h) / 360.0;
  if (grad2) *grad2 = a * d21 + b * d22 + ((a * a * a - a) * d41 + (b * b * b - b) * d42) * (h * h) / 6.0;
  return a * p1 + b * p2 + ((a * a * a - a) * d21 + (b * b * b - b) * d22) * (h * h) / 6.0 +
         ((3 * a * a * a * a * a - 10 * a * a * a + 7 * a) * d41 + (3 * b * b * b * b * b - 10 * b * b * b + 7 * b) * d42) *
             (h * h * h * h) / 360.0;
}

/* Spline::get_max_y_mag src/spline.cpp:205-214 */
static double spl_max_y_mag(const spl_t* s) {
  double m = 0.0;
  for (int k = 0; k < s->n; ++k)
    if (fabs(m) < fabs(s->y[k])) m = s->y[k];
  return m;
}

static void load_spline(spl_t* s, int basis, int n, double x0, double xn, const int* bc, const double* rec) {
  spl_setup(s, basis, n, x0, xn, bc);
  for (int k = 0; k < n; ++k) s->y[k] = rec[k];
  s->grad[0] = rec[n];
  s->grad[1] = rec[n + 1];
  s->grad2[0] = rec[n + 2];
  s->grad2[1] = rec[n + 3];
  spl_refresh(s);
}

/* ------------------------------------------------------------ public: single spline */

void orc_spline_refresh(int basis, int n, double x0, double xn, const int* bc, const double* rec, double* ypp,
                        double* yp4) {
  spl_t s;
  load_spline(&s, basis, n, x0, xn, bc, rec);
  for (int k = 0; k < n; ++k) { ypp[k] = s.ypp[k]; yp4[k] = s.yp4[k]; }
}

int orc_knotshift(int n, double x0, double xn, double r, double* shift) {
  spl_t s;
  int bc[2] = {-1, -1};
  spl_setup(&s, 3, n, x0, xn, bc);
  return spl_knotshift(&s, r, shift);
}

double orc_splint_knot(int basis, int n, double x0, double xn, const int* bc, const double* rec, int knot,
                       double shift, double* grad) {
  spl_t s;
  load_spline(&s, basis, n, x0, xn, bc, rec);
  return spl_comb_knot(&s, knot, shift, grad);
}

double orc_splint_point(int basis, int n, double x0, double xn, const int* bc, const double* rec, double r,
                        double* grad, double* grad2) {
  spl_t s;
  load_spline(&s, basis, n, x0, xn, bc, rec);
  return spl_comb_point(&s, r, grad, grad2);
}

/* ---------------------------------------------------------------- context */

static int is_gmeam(const orc_ctx* c) { return c->style == ORC_GMEAM_SPLINE || c->style == ORC_GMEAM; }
/* extra p p q energy terms: 1 for jmeam, 2 for jmeam2 (src/pots/pot_jmeam/pot_jmeam.cpp:30-47,
 * src/pots/pot_jmeam2/pot_jmeam2.cpp:30-46) */
static int njm(const orc_ctx* c) {
  return (c->style == ORC_JMEAM_SPLINE || c->style == ORC_JMEAM) ? 1
         : (c->style == ORC_JMEAM2_SPLINE || c->style == ORC_JMEAM2) ? 2 : 0;
}
static int is_meam(const orc_ctx* c) {
  return c->style == ORC_MEAM_SPLINE || c->style == ORC_MEAM || is_gmeam(c) || njm(c) > 0;
}
static int is_pair(const orc_ctx* c) { return c->style == ORC_PAIR_SPLINE || c->style == ORC_PAIR; }
static int is_spline_style(const orc_ctx* c) {
  return c->style == ORC_EAM_SPLINE || c->style == ORC_MEAM_SPLINE || c->style == ORC_GMEAM_SPLINE ||
         c->style == ORC_PAIR_SPLINE || c->style == ORC_JMEAM_SPLINE || c->style == ORC_JMEAM2_SPLINE;
}

/* PotFns::get_2body_alloy_idx Pair_ij, src/pot_fns.cpp:141-143 */
static int pair_ij(int ti, int tj, int nt) {
  return (ti <= tj) ? ti * nt + tj - ((ti * (ti + 1)) / 2) : tj * nt + ti - ((tj * (tj + 1)) / 2);
}

/* g spline of a triplet: PotFns::get_3body_alloy_idx src/pot_fns.cpp:154-171,
 * Triplet_i for meam styles (src/pots/pot_meam/pot_meam.cpp:45-47), Triplet_i_jk
 * for gmeam styles (src/pots/pot_gmeam/pot_gmeam.cpp:31-36). */
static const spl_t* g_spline(const orc_ctx* c, int ti, int tj, int tk) {
  const int nt = c->nt;
  const int idx = is_gmeam(c) ? ti * (nt * (nt + 1) / 2) + pair_ij(tj, tk, nt) : ti;
  return &c->spl[c->fn_off[4] + idx];
}

static int check_fn(const orc_ctx* c, const spl_t* s, double r, int* knot, double* shift);

/* Terms a triplet carries: bit 0 the density term f_ij f_ik g, bit m (1, 2)
 * the jmeam energy term p_ij p_ik q (p2 p2 q2).  Spline styles: the cosine
 * lies inside the angular spline (get_knotshift); point styles: both legs lie
 * inside the radial partner.  check_triplet src/pots/pot_meam/triplet_meam.cpp:48-71,
 * src/pots/pot_meam_spline/triplet_meam_spline.cpp:49-73,
 * src/pots/pot_jmeam/triplet_jmeam.cpp:48-88,
 * src/pots/pot_jmeam_spline/triplet_jmeam_spline.cpp:49-88.  A triplet with
 * no term is dropped.  (The jmeam spline styles evaluate both angular
 * splines of a kept triplet at their stored knots, -1 included; here a term
 * whose knot is -1 contributes nothing.) */
static int triplet_terms(const orc_ctx* c, int ti, const pair_t* P, const pair_t* Q, double cs) {
  const int nt = c->nt;
  int terms = 0;
  for (int m = 0; m <= njm(c); ++m) {
    if (is_spline_style(c)) {
      const spl_t* a = m == 0 ? g_spline(c, ti, c->types[P->j], c->types[Q->j]) : &c->spl[c->fn_off[4 + 2 * m] + ti];
      double sh;
      if (spl_knotshift(a, cs, &sh) != -1) terms |= 1 << m;
    } else {
      const int f = m == 0 ? 3 : 3 + 2 * m;
      const spl_t* sj = &c->spl[c->fn_off[f] + pair_ij(ti, c->types[P->j], nt)];
      const spl_t* sk = &c->spl[c->fn_off[f] + pair_ij(ti, c->types[Q->j], nt)];
      if (!(P->r > sj->x[sj->n - 1] || Q->r > sk->x[sk->n - 1])) terms |= 1 << m;
    }
  }
  return terms;
}

/* value and d/dcos of angular spline a at cs (knot form for spline styles) */
static double ang_eval(const orc_ctx* c, const spl_t* a, double cs, double* d) {
  if (is_spline_style(c)) {
    double sh;
    int kn = spl_knotshift(a, cs, &sh);
    return spl_comb_knot(a, kn, sh, d);
  }
  return spl_comb_point(a, cs, d, NULL);
}

static void refresh_all(orc_ctx* c, const double* knots) {
  for (int s = 0; s < c->nspl; ++s)
    load_spline(&c->spl[s], c->basis[s], c->nknots[s], c->x0[s], c->xn[s], &c->bc[2 * s], knots + c->koff[s]);
}

/* Pair check for one function: spline styles store knot/shift, point styles a
 * flag (r <= x_n).  Returns nonzero if r lies below the first knot. */
static int check_fn(const orc_ctx* c, const spl_t* s, double r, int* knot, double* shift) {
  if (r < s->x[0]) return 1;
  if (is_spline_style(c)) {
    *knot = spl_knotshift(s, r, shift);
  } else {
    *knot = (r <= s->x[s->n - 1]) ? 1 : -1;
    *shift = 0.0;
  }
  return 0;
}

orc_ctx* orc_create(const orc_layout* lay, const orc_db* db, char* err) {
  orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
  c->style = lay->style;
  c->nt = lay->ntypes;
  c->nspl = lay->nspl;
  c->embed_extrap = lay->embed_extrap;
  int nt = c->nt, npij = nt * (nt + 1) / 2;
  int counts[9] = {npij, nt, nt, npij, is_gmeam(c) ? nt * npij : nt, npij, nt, npij, nt};
  int nfn = is_meam(c) ? 5 + 2 * njm(c) : (is_pair(c) ? 1 : 3);
  int tot = 0;
  for (int f = 0; f < nfn; ++f) { c->fn_off[f] = tot; tot += counts[f]; }
  if (tot != c->nspl) {
    if (err) snprintf(err, 256, "layout has %d splines, style needs %d", c->nspl, tot);
    free(c);
    return NULL;
  }
  c->basis = (int*)malloc(sizeof(int) * c->nspl);
  c->nknots = (int*)malloc(sizeof(int) * c->nspl);
  c->x0 = (double*)malloc(sizeof(double) * c->nspl);
  c->xn = (double*)malloc(sizeof(double) * c->nspl);
  c->bc = (int*)malloc(sizeof(int) * 2 * c->nspl);
  c->koff = (int*)malloc(sizeof(int) * (c->nspl + 1));
  c->spl = (spl_t*)calloc(c->nspl, sizeof(spl_t));
  int off = 0;
  for (int s = 0; s < c->nspl; ++s) {
    c->basis[s] = lay->basis[s];
    c->nknots[s] = lay->nknots[s];
    if (c->nknots[s] < 3 || c->nknots[s] > ORC_MAXK) {
      if (err) snprintf(err, 256, "spline %d: %d knots unsupported", s, c->nknots[s]);
      orc_destroy(c);
      return NULL;
    }
    c->x0[s] = lay->x0[s];
    c->xn[s] = lay->xn[s];
    c->bc[2 * s] = lay->bc[2 * s];
    c->bc[2 * s + 1] = lay->bc[2 * s + 1];
    c->koff[s] = off;
    off += c->nknots[s] + 4;
    spl_setup(&c->spl[s], c->basis[s], c->nknots[s], c->x0[s], c->xn[s], &c->bc[2 * s]);
  }
  c->koff[c->nspl] = off;

  /* max radial cutoff: phi, rho (, f (, p (, p2))) -- src/potential.cpp:73-82 */
  double max_rcut = 0.0;
  int radial[5] = {0, 1, 3, 5, 7};
  for (int q = 0; q < (is_meam(c) ? 3 + njm(c) : (is_pair(c) ? 1 : 2)); ++q) {
    int f = radial[q];
    for (int s = c->fn_off[f]; s < c->fn_off[f] + counts[f]; ++s)
      if (c->xn[s] > max_rcut) max_rcut = c->xn[s];
  }

  c->ncells = db->ncells;
  c->natoms_c = (int*)malloc(sizeof(int) * c->ncells);
  int na = 0;
  for (int k = 0; k < c->ncells; ++k) { c->natoms_c[k] = db->natoms[k]; na += db->natoms[k]; }
  c->natoms = na;
  c->cell_of = (int*)malloc(sizeof(int) * (na ? na : 1));
  c->types = (int*)malloc(sizeof(int) * (na ? na : 1));
  c->force0 = (double*)malloc(sizeof(double) * 3 * (na ? na : 1));
  c->energy0 = (double*)malloc(sizeof(double) * c->ncells);
  c->stress0 = (double*)malloc(sizeof(double) * 6 * c->ncells);
  c->weight = (double*)malloc(sizeof(double) * c->ncells);
  c->eweight = (double*)malloc(sizeof(double) * c->ncells);
  c->sweight = (double*)malloc(sizeof(double) * c->ncells);
  c->vol = (double*)malloc(sizeof(double) * c->ncells);
  c->energy_weight = db->energy_weight;
  c->stress_weight = db->stress_weight;
  memcpy(c->types, db->types, sizeof(int) * na);
  memcpy(c->force0, db->force0, sizeof(double) * 3 * na);
  for (int k = 0; k < c->ncells; ++k) {
    c->energy0[k] = db->energy0[k];
    for (int m = 0; m < 6; ++m) c->stress0[6 * k + m] = db->stress0[6 * k + m];
    c->weight[k] = db->weight[k];
    c->eweight[k] = db->eweight[k];
    c->sweight[k] = db->sweight[k];
  }
  for (int a = 0; a < na; ++a) {
    if (c->types[a] < 0 || c->types[a] >= nt) {
      if (err) snprintf(err, 256, "incorrect atom type specified - atom=%d", a);
      orc_destroy(c);
      return NULL;
    }
  }

  /* pass 1 counts, pass 2 fills (same loops) */
  c->poff = (long long*)calloc((size_t)na + 1, sizeof(long long));
  c->ntrip_atom = (long long*)calloc((size_t)na + 1, sizeof(long long));
  long long cap = 0;
  c->pairs = NULL;
  for (int pass = 0; pass < 2; ++pass) {
    long long np = 0;
    int abase = 0;
    for (int cc = 0; cc < c->ncells; ++cc) {
      const double* A = db->box + 9 * cc;
      const double* a0 = A; const double* a1 = A + 3; const double* a2 = A + 6;
      /* Cell::compute_vol src/cell.cpp:176-179 and find_scale :158-170 */
      double c12[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2], a1[0] * a2[1] - a1[1] * a2[0]};
      double c20[3] = {a2[1] * a0[2] - a2[2] * a0[1], a2[2] * a0[0] - a2[0] * a0[2], a2[0] * a0[1] - a2[1] * a0[0]};
      double c01[3] = {a0[1] * a1[2] - a0[2] * a1[1], a0[2] * a1[0] - a0[0] * a1[2], a0[0] * a1[1] - a0[1] * a1[0]};
      double vol = fabs(a0[0] * c12[0] + a0[1] * c12[1] + a0[2] * c12[2]);
      c->vol[cc] = vol;
      double sx = (int)ceil(max_rcut * (sqrt(c12[0] * c12[0] + c12[1] * c12[1] + c12[2] * c12[2]) / vol));
      double sy = (int)ceil(max_rcut * (sqrt(c20[0] * c20[0] + c20[1] * c20[1] + c20[2] * c20[2]) / vol));
      double sz = (int)ceil(max_rcut * (sqrt(c01[0] * c01[0] + c01[1] * c01[1] + c01[2] * c01[2]) / vol));
      if (sx <= 0 || sy <= 0 || sz <= 0) {
        if (err) snprintf(err, 256, "Cell #%d is too small or potential is too big", cc);
        orc_destroy(c);
        return NULL;
      }
      int nc = db->natoms[cc];
      for (int ia = 0; ia < nc; ++ia) {
        int gi = abase + ia;
        c->cell_of[gi] = cc;
        if (pass == 0) c->poff[gi] = np;
        const double* pi = db->pos + 3 * gi;
        /* pair styles count each pair once: j starts at i (src/config.cpp:179-184) */
        for (int j = is_pair(c) ? ia : 0; j < nc; ++j) {
          const double* pj = db->pos + 3 * (abase + j);
          for (int ix = -sx; ix <= sx; ++ix)
            for (int iy = -sy; iy <= sy; ++iy)
              for (int iz = -sz; iz <= sz; ++iz) {
                if (ia == j && ix == 0 && iy == 0 && iz == 0) continue;
                double d[3];
                d[0] = pj[0] + ix * a0[0] + iy * a1[0] + iz * a2[0] - pi[0];
                d[1] = pj[1] + ix * a0[1] + iy * a1[1] + iz * a2[1] - pi[1];
                d[2] = pj[2] + ix * a0[2] + iy * a1[2] + iz * a2[2] - pi[2];
                double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                if (r >= max_rcut) continue;
                pair_t P;
                P.j = abase + j;
                P.r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                P.invr = 1.0 / P.r;
                P.dist[0] = d[0] / P.r;
                P.dist[1] = d[1] / P.r;
                P.dist[2] = d[2] / P.r;
                int ti = c->types[gi], tj = c->types[abase + j];
                P.phi_idx = c->fn_off[0] + pair_ij(ti, tj, nt);
                P.rho_idx = c->fn_off[1] + tj;
                P.f_idx = is_meam(c) ? c->fn_off[3] + pair_ij(ti, tj, nt) : 0;
                P.phi_knot = P.rho_knot = P.f_knot = -1;
                P.phi_shift = P.rho_shift = P.f_shift = 0.0;
                int bad = check_fn(c, &c->spl[P.phi_idx], P.r, &P.phi_knot, &P.phi_shift);
                if (!is_pair(c)) bad |= check_fn(c, &c->spl[P.rho_idx], P.r, &P.rho_knot, &P.rho_shift);
                if (is_meam(c)) bad |= check_fn(c, &c->spl[P.f_idx], P.r, &P.f_knot, &P.f_shift);
                int any_p = 0;
                for (int m = 0; m < 2; ++m) {
                  P.p_idx[m] = 0;
                  P.p_knot[m] = -1;
                  P.p_shift[m] = 0.0;
                  if (m >= njm(c)) continue;
                  P.p_idx[m] = c->fn_off[5 + 2 * m] + pair_ij(ti, tj, nt);
                  bad |= check_fn(c, &c->spl[P.p_idx[m]], P.r, &P.p_knot[m], &P.p_shift[m]);
                  any_p |= P.p_knot[m] != -1;
                }
                if (bad) {
                  if (err)
                    snprintf(err, 256, "Pair distance too short for potential (r=%g) cell=%d atom=%d neigh=%d", P.r,
                             cc, ia, j);
                  orc_destroy(c);
                  return NULL;
                }
                if (P.phi_knot == -1 && P.rho_knot == -1 && (!is_meam(c) || P.f_knot == -1) && !any_p) continue;
                if (pass == 1) c->pairs[np] = P;
                ++np;
              }
        }
      }
      abase += nc;
    }
    if (pass == 0) {
      c->poff[na] = np;
      cap = np;
      c->pairs = (pair_t*)malloc(sizeof(pair_t) * (cap ? cap : 1));
    }
    c->npairs = np;
  }

  /* triplet counts, src/config.cpp:257-282 */
  c->ntriplets = 0;
  for (int a = 0; a < na; ++a) {
    long long nt_a = 0;
    if (is_meam(c)) {
      long long b = c->poff[a], e = c->poff[a + 1];
      int ti = c->types[a];
      for (long long p = b; p < e - 1; ++p)
        for (long long q = p + 1; q < e; ++q) {
          const pair_t* P = &c->pairs[p];
          const pair_t* Q = &c->pairs[q];
          double cs = P->dist[0] * Q->dist[0] + P->dist[1] * Q->dist[1] + P->dist[2] * Q->dist[2];
          if (triplet_terms(c, ti, P, Q, cs)) ++nt_a;
        }
    }
    c->ntrip_atom[a] = nt_a;
    c->ntriplets += nt_a;
  }
  return c;
}

void orc_destroy(orc_ctx* c) {
  if (!c) return;
  free(c->basis); free(c->nknots); free(c->x0); free(c->xn); free(c->bc); free(c->koff); free(c->spl);
  free(c->natoms_c); free(c->cell_of); free(c->types); free(c->force0); free(c->energy0); free(c->stress0);
  free(c->weight); free(c->eweight); free(c->sweight); free(c->vol); free(c->poff); free(c->pairs);
  free(c->ntrip_atom);
  free(c);
}

void orc_counts(const orc_ctx* c, long long* npairs, long long* ntriplets) {
  if (npairs) *npairs = c->npairs;
  if (ntriplets) *ntriplets = c->ntriplets;
}

int orc_errvec_size(const orc_ctx* c) {
  if (is_pair(c)) return 7 * c->ncells + 3 * c->natoms;
  return 7 * c->ncells + 3 * c->natoms + c->ncells +
         (is_gmeam(c) ? c->nt * (c->nt + 1) / 2 : is_meam(c) ? 1 + njm(c) : c->nt);
}

/* ------------------------------------------------------------ objective core */

typedef struct {
  double* energy;     /* [ncells] raw sums */
  double* stress;     /* [ncells*6] raw sums */
  double* force;      /* [natoms*3] */
  double* constraint; /* [ncells] */
  double* drho; double* f; double* df; /* per-pair scratch */
  double* p[2]; double* dp[2];         /* jmeam styles: p (and p2) per pair */
} accum_t;

static double eval_fn(const orc_ctx* c, const spl_t* s, int knot, double shift, double r, double* grad) {
  if (is_spline_style(c)) return spl_comb_knot(s, knot, shift, grad);
  return spl_comb_point(s, r, grad, NULL);
}

/* Pair styles: one pass over the half list; a pair whose neighbour is the atom
 * itself (a periodic image) counts half, src/pots/pot_pair_spline/
 * pot_pair_spline.cpp:98-134 (knot form), src/pots/pot_pair/pot_pair.cpp:99-135
 * (point form). */
static void pair_body(const orc_ctx* c, int a, accum_t* acc) {
  const int cc = c->cell_of[a];
  for (long long p = c->poff[a]; p < c->poff[a + 1]; ++p) {
    const pair_t* P = &c->pairs[p];
    if (P->phi_knot == -1) continue;
    double phigrad;
    double phival = eval_fn(c, &c->spl[P->phi_idx], P->phi_knot, P->phi_shift, P->r, &phigrad);
    if (P->j == a) {
      phival *= 0.5;
      phigrad *= 0.5;
    }
    acc->energy[cc] += phival;
    double t[3] = {P->dist[0] * phigrad, P->dist[1] * phigrad, P->dist[2] * phigrad};
    double* fi = acc->force + 3 * a; double* fj = acc->force + 3 * P->j;
    fi[0] += t[0]; fi[1] += t[1]; fi[2] += t[2];
    fj[0] -= t[0]; fj[1] -= t[1]; fj[2] -= t[2];
    t[0] *= P->r; t[1] *= P->r; t[2] *= P->r;
    double* s = acc->stress + 6 * cc;
    s[0] -= P->dist[0] * t[0]; s[1] -= P->dist[1] * t[1]; s[2] -= P->dist[2] * t[2];
    s[3] -= P->dist[0] * t[1]; s[4] -= P->dist[1] * t[2]; s[5] -= P->dist[2] * t[0];
  }
}

/* The per-atom body of fast_compute / compute (pair loop, triplet loop,
 * embedding, pair force loop, triplet force loop). */
static void atom_body(const orc_ctx* c, int a, accum_t* acc) {
  if (is_pair(c)) {
    pair_body(c, a, acc);
    return;
  }
  const int meam = is_meam(c);
  const int cc = c->cell_of[a];
  const double w = c->weight[cc];
  long long b = c->poff[a], e = c->poff[a + 1];
  double rho_val = 0.0, dF = 0.0;
  for (long long p = b; p < e; ++p) {
    const pair_t* P = &c->pairs[p];
    if (P->phi_knot != -1) {
      double phigrad;
      double phival = 0.5 * eval_fn(c, &c->spl[P->phi_idx], P->phi_knot, P->phi_shift, P->r, &phigrad);
      phigrad *= 0.5;
      acc->energy[cc] += phival;
      double t[3] = {P->dist[0] * phigrad, P->dist[1] * phigrad, P->dist[2] * phigrad};
      double* fi = acc->force + 3 * a; double* fj = acc->force + 3 * P->j;
      fi[0] += t[0]; fi[1] += t[1]; fi[2] += t[2];
      fj[0] -= t[0]; fj[1] -= t[1]; fj[2] -= t[2];
      t[0] *= P->r; t[1] *= P->r; t[2] *= P->r;
      double* s = acc->stress + 6 * cc;
      s[0] -= P->dist[0] * t[0]; s[1] -= P->dist[1] * t[1]; s[2] -= P->dist[2] * t[2];
      s[3] -= P->dist[0] * t[1]; s[4] -= P->dist[1] * t[2]; s[5] -= P->dist[2] * t[0];
    }
    if (P->rho_knot != -1) {
      rho_val += eval_fn(c, &c->spl[P->rho_idx], P->rho_knot, P->rho_shift, P->r, &acc->drho[p]);
    } else {
      acc->drho[p] = 0.0;
    }
    if (meam) {
      if (P->f_knot != -1) {
        acc->f[p] = eval_fn(c, &c->spl[P->f_idx], P->f_knot, P->f_shift, P->r, &acc->df[p]);
      } else {
        acc->f[p] = 0.0;
        acc->df[p] = 0.0;
      }
    }
    for (int m = 0; m < njm(c); ++m) {
      if (P->p_knot[m] != -1) {
        acc->p[m][p] = eval_fn(c, &c->spl[P->p_idx[m]], P->p_knot[m], P->p_shift[m], P->r, &acc->dp[m][p]);
      } else {
        acc->p[m][p] = 0.0;
        acc->dp[m][p] = 0.0;
      }
    }
  }
  const int ti = c->types[a];
  if (meam) {
    for (long long p = b; p < e - 1; ++p)
      for (long long q = p + 1; q < e; ++q) {
        const pair_t* P = &c->pairs[p];
        const pair_t* Q = &c->pairs[q];
        double cs = P->dist[0] * Q->dist[0] + P->dist[1] * Q->dist[1] + P->dist[2] * Q->dist[2];
        const int terms = triplet_terms(c, ti, P, Q, cs);
        double dg;
        if (terms & 1) rho_val += acc->f[p] * acc->f[q] * ang_eval(c, g_spline(c, ti, c->types[P->j], c->types[Q->j]), cs, &dg);
        for (int m = 0; m < njm(c); ++m)
          if (terms >> (m + 1) & 1)
            acc->energy[cc] += acc->p[m][p] * acc->p[m][q] * ang_eval(c, &c->spl[c->fn_off[6 + 2 * m] + ti], cs, &dg);
      }
  }
  const spl_t* U = &c->spl[c->fn_off[2] + c->types[a]];
  double lo = U->x[0], hi = U->x[U->n - 1];
  if (rho_val < lo) {
    acc->constraint[cc] += w * DUMMY_WEIGHT * 10. * (lo - rho_val) * (lo - rho_val);
    if (!c->embed_extrap) rho_val = lo;
  } else if (rho_val > hi) {
    acc->constraint[cc] += w * DUMMY_WEIGHT * 10. * (rho_val - hi) * (rho_val - hi);
    if (!c->embed_extrap) rho_val = hi;
  }
  acc->energy[cc] += spl_comb_point(U, rho_val, &dF, NULL);
  for (long long p = b; p < e; ++p) {
    const pair_t* P = &c->pairs[p];
    double t[3] = {P->dist[0] * acc->drho[p] * dF, P->dist[1] * acc->drho[p] * dF, P->dist[2] * acc->drho[p] * dF};
    double* fi = acc->force + 3 * a; double* fj = acc->force + 3 * P->j;
    fi[0] += t[0]; fi[1] += t[1]; fi[2] += t[2];
    fj[0] -= t[0]; fj[1] -= t[1]; fj[2] -= t[2];
    t[0] *= P->r; t[1] *= P->r; t[2] *= P->r;
    double* s = acc->stress + 6 * cc;
    s[0] -= P->dist[0] * t[0]; s[1] -= P->dist[1] * t[1]; s[2] -= P->dist[2] * t[2];
    s[3] -= P->dist[0] * t[1]; s[4] -= P->dist[1] * t[2]; s[5] -= P->dist[2] * t[0];
  }
  if (meam) {
    for (long long p = b; p < e - 1; ++p)
      for (long long q = p + 1; q < e; ++q) {
        const pair_t* P = &c->pairs[p];
        const pair_t* Q = &c->pairs[q];
        double cs = P->dist[0] * Q->dist[0] + P->dist[1] * Q->dist[1] + P->dist[2] * Q->dist[2];
        const int terms = triplet_terms(c, ti, P, Q, cs);
        if (!terms) continue;
        double gv = 0.0, dg = 0.0;
        if (terms & 1) gv = ang_eval(c, g_spline(c, ti, c->types[P->j], c->types[Q->j]), cs, &dg);
        double dV3j = gv * acc->df[p] * acc->f[q] * dF;
        double dV3k = gv * acc->f[p] * acc->df[q] * dF;
        double V3 = dg * acc->f[p] * acc->f[q] * dF;
        /* jmeam: + q p'_j p_k, + q p_j p'_k, + q' p_j p_k (src/pots/pot_jmeam/pot_jmeam.cpp:263-265) */
        for (int m = 0; m < njm(c); ++m) {
          double qv = 0.0, dq = 0.0;
          if (terms >> (m + 1) & 1) qv = ang_eval(c, &c->spl[c->fn_off[6 + 2 * m] + ti], cs, &dq);
          dV3j += qv * acc->dp[m][p] * acc->p[m][q];
          dV3k += qv * acc->p[m][p] * acc->dp[m][q];
          V3 += dq * acc->p[m][p] * acc->p[m][q];
        }
        double vlj = V3 * P->invr;
        double vlk = V3 * Q->invr;
        double vv3j = dV3j - vlj * cs;
        double vv3k = dV3k - vlk * cs;
        double dfj[3], dfk[3];
        for (int m = 0; m < 3; ++m) {
          dfj[m] = P->dist[m] * vv3j + Q->dist[m] * vlj;
          dfk[m] = Q->dist[m] * vv3k + P->dist[m] * vlk;
        }
        double* fi = acc->force + 3 * a; double* fj = acc->force + 3 * P->j; double* fk = acc->force + 3 * Q->j;
        for (int m = 0; m < 3; ++m) {
          fi[m] += dfj[m] + dfk[m];
          fj[m] -= dfj[m];
          fk[m] -= dfk[m];
        }
        for (int m = 0; m < 3; ++m) { dfj[m] *= P->r; dfk[m] *= Q->r; }
        double* s = acc->stress + 6 * cc;
        s[0] -= P->dist[0] * dfj[0]; s[1] -= P->dist[1] * dfj[1]; s[2] -= P->dist[2] * dfj[2];
        s[3] -= P->dist[0] * dfj[1]; s[4] -= P->dist[1] * dfj[2]; s[5] -= P->dist[2] * dfj[0];
        s[0] -= Q->dist[0] * dfk[0]; s[1] -= Q->dist[1] * dfk[1]; s[2] -= Q->dist[2] * dfk[2];
        s[3] -= Q->dist[0] * dfk[1]; s[4] -= Q->dist[1] * dfk[2]; s[5] -= Q->dist[2] * dfk[0];
      }
  }
}

/* accumulate_error src/potential.cpp:415-491 plus the style penalty
 * (MEAM f-max src/pots/pot_meam_spline/pot_meam_spline.cpp:274-278, EAM U'(mid)
 * src/pots/pot_eam_spline/pot_eam_spline.cpp:211-218). */
static double finish(const orc_ctx* c, const accum_t* acc, double* errvec, double* cell_energy, double* cell_stress,
                     double* forces) {
  double F = 0.0;
  long long cnt = 0;
  int a = 0;
  for (int cc = 0; cc < c->ncells; ++cc) {
    int n = c->natoms_c[cc];
    double E = acc->energy[cc] / n;
    double e = c->weight[cc] * c->eweight[cc] * c->energy_weight * n * (E - c->energy0[cc]);
    F += e * e;
    if (errvec) errvec[cnt] = e;
    ++cnt;
    if (cell_energy) cell_energy[cc] = E;
    for (int m = 0; m < 6; ++m) {
      double S = acc->stress[6 * cc + m] / c->vol[cc];
      double es = c->weight[cc] * c->sweight[cc] * c->stress_weight * (S - c->stress0[6 * cc + m]);
      F += es * es;
      if (errvec) errvec[cnt] = es;
      ++cnt;
      if (cell_stress) cell_stress[6 * cc + m] = S;
    }
    for (int k = 0; k < n; ++k, ++a)
      for (int m = 0; m < 3; ++m) {
        double fv = acc->force[3 * a + m];
        double ef = c->weight[cc] * (fv - c->force0[3 * a + m]);
        F += ef * ef;
        if (errvec) errvec[cnt] = ef;
        ++cnt;
        if (forces) forces[3 * a + m] = fv;
      }
  }
  if (is_pair(c)) return F; /* no constraint rows, no penalty (extra_error is empty) */
  for (int cc = 0; cc < c->ncells; ++cc) {
    double v = acc->constraint[cc];
    F += v * v;
    if (errvec) errvec[cnt] = v;
    ++cnt;
  }
  if (is_gmeam(c)) {
    /* one f-max term per f spline, src/pots/pot_gmeam_spline/pot_gmeam_spline.cpp:275-280 */
    int npij = c->nt * (c->nt + 1) / 2;
    for (int s = c->fn_off[3]; s < c->fn_off[3] + npij; ++s) {
      double m = fabs(spl_max_y_mag(&c->spl[s]));
      double pe = DUMMY_WEIGHT * 25. * (1.0 - m) * (1.0 - m);
      F += pe * pe;
      if (errvec) errvec[cnt] = pe;
      ++cnt;
    }
  } else if (is_meam(c)) {
    /* PotFns::get_max_y_mag src/pot_fns.cpp:100-111 over the f splines, then
     * (jmeam) over the p and p2 splines, src/pots/pot_jmeam2/pot_jmeam2.cpp:308-324 */
    int npij = c->nt * (c->nt + 1) / 2;
    for (int grp = 0; grp <= njm(c); ++grp) {
      const int f0 = c->fn_off[grp == 0 ? 3 : 3 + 2 * grp];
      double mx = 0.0;
      for (int s = f0; s < f0 + npij; ++s) {
        double v = spl_max_y_mag(&c->spl[s]);
        if (fabs(mx) < fabs(v)) mx = v;
      }
      double m = fabs(mx);
      double pe = DUMMY_WEIGHT * 25. * (1.0 - m) * (1.0 - m);
      F += pe * pe;
      if (errvec) errvec[cnt + grp] = pe;
    }
  } else {
    for (int t = 0; t < c->nt; ++t) {
      const spl_t* U = &c->spl[c->fn_off[2] + t];
      double g;
      spl_comb_point(U, 0.5 * (U->x[0] + U->x[U->n - 1]), &g, NULL);
      double pe = DUMMY_WEIGHT * g;
      F += pe * pe;
      if (errvec) errvec[cnt + t] = pe;
    }
  }
  return F;
}

static void accum_alloc(const orc_ctx* c, accum_t* acc) {
  acc->energy = (double*)calloc(c->ncells, sizeof(double));
  acc->stress = (double*)calloc(6 * (size_t)c->ncells, sizeof(double));
  acc->force = (double*)calloc(3 * (size_t)(c->natoms ? c->natoms : 1), sizeof(double));
  acc->constraint = (double*)calloc(c->ncells, sizeof(double));
  size_t np = (size_t)(c->npairs ? c->npairs : 1);
  acc->drho = (double*)calloc(np, sizeof(double));
  acc->f = (double*)calloc(np, sizeof(double));
  acc->df = (double*)calloc(np, sizeof(double));
  for (int m = 0; m < 2; ++m) {
    acc->p[m] = (double*)calloc(np, sizeof(double));
    acc->dp[m] = (double*)calloc(np, sizeof(double));
  }
}

static void accum_free(accum_t* acc) {
  free(acc->energy); free(acc->stress); free(acc->force); free(acc->constraint);
  free(acc->drho); free(acc->f); free(acc->df);
  for (int m = 0; m < 2; ++m) { free(acc->p[m]); free(acc->dp[m]); }
}

double orc_eval(orc_ctx* c, const double* knots, double* errvec, double* cell_energy, double* cell_stress,
                double* forces) {
  refresh_all(c, knots);
  accum_t acc;
  accum_alloc(c, &acc);
  for (int a = 0; a < c->natoms; ++a) atom_body(c, a, &acc);
  double F = finish(c, &acc, errvec, cell_energy, cell_stress, forces);
  accum_free(&acc);
  return F;
}

/* ----------------------------------------------------------- threaded form */

typedef struct {
  const orc_ctx* c;
  const int* atoms;
  int count;
  accum_t acc;
} worker_t;

static void* worker_run(void* arg) {
  worker_t* w = (worker_t*)arg;
  for (int k = 0; k < w->count; ++k) atom_body(w->c, w->atoms[k], &w->acc);
  return NULL;
}

static int cmp_trip_desc(const void* x, const void* y, void* ctx) {
  const long long* t = (const long long*)ctx;
  int a = *(const int*)x, b = *(const int*)y;
  return (t[a] < t[b]) - (t[a] > t[b]);
}

static const long long* g_sort_key;
static int cmp_trip_desc_g(const void* x, const void* y) { return cmp_trip_desc(x, y, (void*)g_sort_key); }

double orc_eval_mt(orc_ctx* c, const double* knots, int nthreads) {
  return orc_eval_mt_out(c, knots, nthreads, NULL, NULL, NULL, NULL);
}

/* orc_eval_mt with the outputs of orc_eval (any may be NULL). */
double orc_eval_mt_out(orc_ctx* c, const double* knots, int nthreads, double* errvec, double* cell_energy,
                       double* cell_stress, double* forces) {
  if (nthreads < 1) nthreads = 1;
  refresh_all(c, knots);
  /* greedy balance by per-atom work (triplets + pairs), largest first */
  int* order = (int*)malloc(sizeof(int) * (c->natoms ? c->natoms : 1));
  long long* key = (long long*)malloc(sizeof(long long) * (c->natoms ? c->natoms : 1));
  for (int a = 0; a < c->natoms; ++a) { order[a] = a; key[a] = c->ntrip_atom[a] + (c->poff[a + 1] - c->poff[a]); }
  g_sort_key = key;
  qsort(order, c->natoms, sizeof(int), cmp_trip_desc_g);
  int* owner = (int*)malloc(sizeof(int) * (c->natoms ? c->natoms : 1));
  long long* load = (long long*)calloc(nthreads, sizeof(long long));
  int* cnt = (int*)calloc(nthreads, sizeof(int));
  for (int k = 0; k < c->natoms; ++k) {
    int best = 0;
    for (int t = 1; t < nthreads; ++t) if (load[t] < load[best]) best = t;
    owner[k] = best;
    load[best] += key[order[k]] + 1;
    cnt[best]++;
  }
  int** lists = (int**)malloc(sizeof(int*) * nthreads);
  int* fill = (int*)calloc(nthreads, sizeof(int));
  for (int t = 0; t < nthreads; ++t) lists[t] = (int*)malloc(sizeof(int) * (cnt[t] ? cnt[t] : 1));
  for (int k = 0; k < c->natoms; ++k) lists[owner[k]][fill[owner[k]]++] = order[k];
  worker_t* ws = (worker_t*)calloc(nthreads, sizeof(worker_t));
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
  for (int t = 0; t < nthreads; ++t) {
    ws[t].c = c;
    ws[t].atoms = lists[t];
    ws[t].count = cnt[t];
    accum_alloc(c, &ws[t].acc);
    pthread_create(&th[t], NULL, worker_run, &ws[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  /* the single sum-reduction (MPI_Reduce SUM in the reference) */
  accum_t* r = &ws[0].acc;
  for (int t = 1; t < nthreads; ++t) {
    accum_t* o = &ws[t].acc;
    for (int k = 0; k < c->ncells; ++k) { r->energy[k] += o->energy[k]; r->constraint[k] += o->constraint[k]; }
    for (int k = 0; k < 6 * c->ncells; ++k) r->stress[k] += o->stress[k];
    for (int k = 0; k < 3 * c->natoms; ++k) r->force[k] += o->force[k];
  }
  double F = finish(c, r, errvec, cell_energy, cell_stress, forces);
  for (int t = 0; t < nthreads; ++t) { accum_free(&ws[t].acc); free(lists[t]); }
  free(ws); free(th); free(lists); free(fill); free(cnt); free(load); free(owner); free(order); free(key);
  return F;
}

/* PotMEAMSpline::fast_compute_densities src/pots/pot_meam_spline/pot_meam_spline.cpp:289-349
 * (and the EAM counterpart): unclamped densities. */
void orc_densities(orc_ctx* c, const double* knots, double* dens) {
  refresh_all(c, knots);
  const int meam = is_meam(c);
  double* f = (double*)calloc((size_t)(c->npairs ? c->npairs : 1), sizeof(double));
  for (int a = 0; a < c->natoms; ++a) {
    long long b = c->poff[a], e = c->poff[a + 1];
    double d = 0.0, tmp;
    for (long long p = b; p < e; ++p) {
      const pair_t* P = &c->pairs[p];
      if (P->rho_knot != -1) d += eval_fn(c, &c->spl[P->rho_idx], P->rho_knot, P->rho_shift, P->r, &tmp);
      if (meam) f[p] = (P->f_knot != -1) ? eval_fn(c, &c->spl[P->f_idx], P->f_knot, P->f_shift, P->r, &tmp) : 0.0;
    }
    if (meam) {
      for (long long p = b; p < e - 1; ++p)
        for (long long q = p + 1; q < e; ++q) {
          const pair_t* P = &c->pairs[p];
          const pair_t* Q = &c->pairs[q];
          const spl_t* g = g_spline(c, c->types[a], c->types[P->j], c->types[Q->j]);
          double cs = P->dist[0] * Q->dist[0] + P->dist[1] * Q->dist[1] + P->dist[2] * Q->dist[2];
          if (!(triplet_terms(c, c->types[a], P, Q, cs) & 1)) continue;
          d += f[p] * f[q] * ang_eval(c, g, cs, &tmp);
        }
    }
    dens[a] = d;
  }
  free(f);
}
