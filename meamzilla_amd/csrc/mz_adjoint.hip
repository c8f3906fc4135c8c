// Reverse sweep: gradient of the objective with respect to the knot vector.
//
// With the residuals of the forward pass, the objective's sensitivities are
//   lambda_a = dF/dF_a = 2 w (F_a - F0_a)       (forces)
//   mu_c     = dF/dS_c = 2 e_S W_S / V          (raw stress sums)
//   lamE_c   = dF/dE_c = 2 e_E W_E              (raw energy sums)
//   kappa_c  = dF/dP_c = 2 P_c                  (density-range penalty)
// Every pair p = (i -> j) carries a total force vector T_p that enters
// F_i (+), F_j (-) and the virial -d_p (x) T_p r_p, so the force and stress
// terms of dF reduce to sum_p w_p . dT_p with the fixed dual vector
//   w_p = lambda_i - lambda_j - r_p v_c(d_p),  v = (mu_xx dx + mu_zx dz,
//         mu_yy dy + mu_xy dx, mu_zz dz + mu_yz dy).
// T_p = (phi'/2 + rho' U'_i) d_p + U'_i G_p, G_p the three-body leg sum kept by
// the forward pass.  Per atom, S_i = sum_p (a_p rho'_p + w_p . G_p) (a_p = w_p . d_p)
// multiplies dU'_i, and psi_i = dF/drho_i collects lamE U' + S U'' and the
// penalty.  One sweep over the ordered triplets (j,k) of each atom then
// accumulates, per own pair j, the adjoints of f_j and f'_j and, per g
// interval, the adjoints of the interval's power-basis coefficients; a last
// kernel maps coefficient adjoints through the (linear) spline refresh back to
// the knot vector.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <cstring>
#include <string>
#include <vector>

#include "mz_ctx.h"
#include "mz_device.h"

namespace mz {


struct AdjArgs {
  PotDesc pot;
  const Chunk* chunks;
  const int* atom_poff;
  const int* owner;        // [npairs] atom that owns the pair
  const uint2* pairs;
  const double* pos;
  const double* box;
  const int* types;
  const double* weight;
  const char* tables;
  int table_bytes;
  int natoms, ncells;
  long long npairs;
  const double* rho;       // [cand][natoms] unclamped densities
  const double4* G;        // [cand][npairs]: three-body leg sum G_p and S1_p = sum_k g f_k
  const double* T1;        // jmeam: [cand][npairs][2] T1_p = sum_k q p_k per p p q term
  const double* lam;       // [cand][natoms*3]
  const double* cell_adj;  // [cand][ncells*8]
  double* slots;           // [cand][ntab][gridDim.x]: one partial table per workgroup, slot-major
  int ntab;                // ncoef*6 + nspl*6
  int gco0, ngco;          // first g interval record, number of g interval records
  int grep_log2;           // log2 of the lane replicas of the g-coefficient adjoints
  int nrad;                // interval records below the g records (phi, rho, U, f)
  int rrep_log2;           // log2 of the lane replicas of their adjoints, -1: none
  int rrep_stride;         // doubles per replica (nrad NC padded to 1 mod 32: replicas on distinct banks)
  int nchunks;             // chunks; workgroup x takes x, x + gridDim.x, ...
};

struct ChainArgs {
  int nspl, ncoef, ntab, nknot, meam, nt;
  int f0, nf, u0;
  int fper;                // one f-max penalty per f spline (gmeam)
  int nodens;              // pair styles: no U'(mid) penalty
  int npgrp;               // max-|y| penalty groups: f, then (jmeam) p and p2
  int pgrp0[3];            // first spline of each group (nf splines each)
  const int* nknots;
  const int* basis;
  const int* koff;
  const int* coef_off;
  const double* x0;
  const double* xn;
  const double* J;         // per spline [2n][n+4]: d(ypp, yp4)/d(knot record)
  const int* joff;
  const double* tab;       // [cand][ntab] coefficient and head adjoints, summed over workgroups
  const double* knots;
  const char* tables;
  int table_bytes;
  double* grad;            // [cand][nknot]
};

// Partner data of every pair, planes of field pairs (consecutive partners of
// consecutive lanes read consecutive 16-byte entries, one 16-byte LDS read
// per plane): unit vector d (3), f, a f',
// V = (w - a d) / r (3), the part of the dual normal to the pair over its
// length, with which B_kj = (w_k . d_j - a_k cos) / r_k = V_k . d_j.
constexpr int kPartnerFields = 8;

// Where interval-coefficient adjoints go: coefficient m of interval record rec
// at base[rec * rs + m * cs] (the workgroup table: rs 6, cs 1; a lane replica
// of the radial records: rs NC, cs 1, base at the replica's start).
struct CoefSink {
  double* base;
  int rs, cs;
};

// Interval-coefficient adjoint: weight al on the value, bt on d(value)/dt.
template <bool Q>
__device__ __forceinline__ void acc_interval(const CoefSink& sk, int rec, double t, double al, double bt) {
  double* c = sk.base + rec * sk.rs;
  const int cs = sk.cs;
  const double t2 = t * t, t3 = t2 * t;
  atomicAdd(c, al);
  atomicAdd(c + cs, al * t + bt);
  atomicAdd(c + 2 * cs, al * t2 + 2.0 * bt * t);
  atomicAdd(c + 3 * cs, al * t3 + 3.0 * bt * t2);
  if (Q) {
    const double t4 = t3 * t;
    atomicAdd(c + 4 * cs, al * t4 + 4.0 * bt * t3);
    atomicAdd(c + 5 * cs, al * t4 * t + 5.0 * bt * t4);
  }
}

// Adjoint of a point evaluation (value weight al, d/dr weight be) including
// the end extrapolation; head slots per spline: y0, yn, gL, gR, g2L, g2R.
template <bool Q>
__device__ __forceinline__ void acc_point(double* tab, const CoefSink& sk, int ncoef, int s, const SplHead& H,
                                          double r, double al, double be) {
  double* hd = tab + ncoef * 6 + s * 6;
  if (r < H.x0) {
    const double e = r - H.x0;
    atomicAdd(hd + 0, al);
    atomicAdd(hd + 2, al * e + be);
    if (H.basis == 5) atomicAdd(hd + 4, 0.5 * al * e * e + be * e);
    return;
  }
  if (r > H.xn) {
    const double e = r - H.xn;
    atomicAdd(hd + 1, al);
    atomicAdd(hd + 3, al * e + be);
    if (H.basis == 5) atomicAdd(hd + 5, 0.5 * al * e * e + be * e);
    return;
  }
  const double delta = r - H.x0;
  int k = (int)__dmul_rn(delta, H.invstep);
  if (k >= H.n - 1) k = H.n - 2;
  if (k < 0) k = 0;
  const double t = __dmul_rn(__dadd_rn(delta, -__dmul_rn((double)k, H.step)), H.invstep);
  acc_interval<Q>(sk, H.coef_off + k, t, al, be * H.invstep);
}

// Adjoint of one radial function of a pair (value weight al, d/dr weight be).
template <bool POINT, bool Q>
__device__ __forceinline__ void acc_radial(double* tab, const CoefSink& sk, int ncoef, int s, const SplHead& H,
                                           const SplInterval* IV, double r, double al, double be) {
  if (POINT) {
    if (!(r <= H.xn)) return;
    acc_point<Q>(tab, sk, ncoef, s, H, r, al, be);
    return;
  }
  int k;
  double t;
  if (!locate_knot(H, IV + H.coef_off, r, k, t)) return;
  acc_interval<Q>(sk, H.coef_off + k, t, al, be * H.invstep);
}

// What the sweep needs of one g spline.
struct GSpl {
  double x0, inv, xcut, xn, kmax, y0, yn, gL, gR, g2L, g2R;
  int rec;        // first interval record, counted from the first g record
  bool quintic;
};

__device__ __forceinline__ GSpl gspl_of(const SplHead& H, int gco0) {
  return GSpl{H.x0, H.invstep, H.xcut, H.xn, (double)(H.n - 2), H.y0, H.yn, H.gL, H.gR, H.g2L, H.g2R,
              H.coef_off - gco0, H.basis == 5};
}

// Value and d/dt of an angular spline at cs from its compact records: the
// interval (fl, t) or, for point styles beyond the ends, the quadratic
// extrapolation (extrap; fl 0 left, 1 right; t the offset); keep is false for
// a knot-style cosine at or beyond x_n (1 + 1e-8), whose term is absent.
template <bool POINT, bool Q, int NC>
__device__ __forceinline__ void ang_at(const GSpl& G, const double2* gt2, int ps, double cs, double& v, double& dvt,
                                       double& t, int& fl, bool& extrap, bool& keep) {
  extrap = false;
  keep = true;
  if (POINT && (cs < G.x0 || cs > G.xn)) {
    extrap = true;
    const bool left = cs < G.x0;
    const double e = cs - (left ? G.x0 : G.xn);
    const double g1 = left ? G.gL : G.gR, g2 = left ? G.g2L : G.g2R;
    v = (left ? G.y0 : G.yn) + g1 * e + 0.5 * g2 * e * e;
    dvt = (g1 + g2 * e) / G.inv;
    t = e;
    fl = left ? 0 : 1;
    return;
  }
  const double u = (cs - G.x0) * G.inv;
  const double flo = fmin(fmax(floor(u), 0.0), G.kmax);
  t = u - flo;
  fl = (int)flo;
  const double2* C = gt2 + G.rec + fl;
  const double2 c01 = C[0], c23 = C[ps];
  if (Q) {
    const double2 c45 = C[2 * ps];
    v = c01.x + t * (c01.y + t * (c23.x + t * (c23.y + t * (c45.x + t * c45.y))));
    dvt = c01.y + t * (2.0 * c23.x + t * (3.0 * c23.y + t * (4.0 * c45.x + t * 5.0 * c45.y)));
  } else {
    v = c01.x + t * (c01.y + t * (c23.x + t * c23.y));
    dvt = c01.y + t * (2.0 * c23.x + t * 3.0 * c23.y);
  }
  if (!POINT) keep = cs < G.xcut;
}

// One visit of the triplet sweep: partner pair k and its record, the cosine,
// the g spline (gi), its interval fl and shift t (or the extrapolation side
// and offset), whether the term is kept, and the interval's coefficients.
struct AdjVis {
  int k, gi, fl;
  bool extrap, keep;
  double4 Pa, Pb;
  double cs, t;
  double2 c01, c23, c45;
};

// GR: lane replicas of the g-coefficient adjoints (32 keeps the 32 lanes that
// one LDS cycle serves on distinct replicas; fewer when LDS is short).
// GP: the g spline of a triplet depends on both neighbours' types (gmeam),
// chosen per triplet from the heads staged in LDS.
// JM: jmeam styles, JM p p q terms: the partner record gains p and a p' per
// term, the sweep a q evaluation, a push and the q-coefficient adjoints
// (the g term with psi -> lambda_E, U' -> 1, f -> p).
template <bool MEAM, bool POINT, bool Q, int GR, bool GP, int JM, int BLK>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu((MEAM && !GP && JM == 0) ? 4 : 1))) k_adjoint(const AdjArgs A) {
  extern __shared__ __attribute__((aligned(64))) char dyn[];
  constexpr int NJ = JM > 0 ? JM : 1;
  constexpr int NA = BLK / 4;  // atoms per chunk: phase B takes four lanes per atom
  __shared__ double2 pd[(kPartnerFields + 2 * JM) / 2][BLK];  // partner data, planes of field pairs
  __shared__ double push[2 + JM][BLK];
  __shared__ double psi[NA], Up[NA];
  __shared__ int poff[NA + 1];
  __shared__ int ptyp[GP ? BLK : 1];  // neighbour type of every pair (GP)
  constexpr int NC = Q ? 6 : 4;  // coefficients per g interval that carry adjoints
  const int cand = blockIdx.y;
  const int tid = threadIdx.x;
  const int ncoef = A.pot.ncoef;
  constexpr int R = GR;
  // the paired sweep (two own pairs per lane) for styles whose g depends on
  // the centre atom only and that carry no p p q terms; its instantiations
  // are held to 128 VGPRs (four waves per SIMD, the kernel attribute), a few
  // values live across the sweep spill around it, none inside
  constexpr bool kCouples = MEAM && !GP && JM == 0;
  const int nt = A.pot.nt;

  // The candidate's interval records stay in global memory (cache-resident;
  // read per pair and per atom, not per triplet), so LDS goes to the
  // accumulators; staged are the spline heads (every evaluation starts from
  // one) and the g coefficients, read per triplet.
  const char* tbl = A.tables + (size_t)cand * A.table_bytes;
  const SplInterval* IV = reinterpret_cast<const SplInterval*>(tbl + sizeof(SplHead) * A.pot.nspl);
  const int ghb = A.pot.nspl * (int)sizeof(SplHead);              // all spline heads, staged
  const SplHead* HD = reinterpret_cast<const SplHead*>(dyn);
  const SplHead* GH = HD + A.pot.fn_off[4];
  // g coefficients as planes of pairs: plane m holds (c_2m, c_2m+1) of every record
  double2* gt2 = reinterpret_cast<double2*>(dyn + ghb);
  // knot positions of every interval record, for locating r (LDS reads only)
  KnotRec* kg = reinterpret_cast<KnotRec*>(gt2 + (MEAM ? A.ngco * (NC / 2) : 0));
  double* atab = reinterpret_cast<double*>(kg + ncoef);  // coefficient + head adjoints
  double* gacc = atab + A.ntab;                                   // [g record][coef][replica]
  for (int q = tid; q < ghb / 16; q += BLK) reinterpret_cast<int4*>(dyn)[q] = reinterpret_cast<const int4*>(tbl)[q];
  if (MEAM) {
    const SplInterval* src = reinterpret_cast<const SplInterval*>(A.tables + (size_t)cand * A.table_bytes +
                                                                  sizeof(SplHead) * A.pot.nspl);
    for (int q = tid; q < A.ngco * (NC / 2); q += BLK) {
      const int m = q / A.ngco;
      const SplInterval& I = src[A.gco0 + q - m * A.ngco];
      gt2[q] = make_double2(I.c[2 * m], I.c[2 * m + 1]);
    }
  }
  for (int q = tid; q < ncoef; q += BLK) kg[q] = KnotRec{IV[q].xk, IV[q].invh};
  const int nacc = A.ntab + (MEAM ? A.ngco * NC * R : 0);
  for (int q = tid; q < nacc; q += BLK) atab[q] = 0.0;

  // The grid holds as many workgroups as the GPU keeps resident; workgroup x
  // of a candidate takes chunks x, x + gridDim.x, ... of the decreasing-work
  // order, so every workgroup gets about the same work.  Table staging,
  // zeroing, the g-replica reduction and the partial-table store are paid
  // once per workgroup, and each candidate's partial tables are summed by
  // k_chain in a fixed order (no global atomics).
  for (int ci = blockIdx.x; ci < A.nchunks; ci += gridDim.x) {
  // the lane index, opaque to the compiler inside each chunk: addresses
  // derived from it are recomputed per chunk instead of held (and spilled)
  // across the whole loop
  int tidc = threadIdx.x;
  asm volatile("" : "+v"(tidc));
  const int tid = tidc;
  const Chunk ch = A.chunks[ci];
  const int p = tid;
  const bool has = p < ch.np;
  const int ba = tid >> 2;  // atom of this lane in phase B

  // Every global load of phases A and B is issued here, beside the table
  // staging, so that one barrier covers their latency: pair record, owner and
  // leg sums, then positions, types and force duals of both ends.
  uint2 pr = make_uint2(0u, 0u);
  int ig = ch.a0;
  double G0 = 0, G1 = 0, G2 = 0, S1 = 0;
  double T1v[NJ];
#pragma unroll
  for (int m = 0; m < NJ; ++m) T1v[m] = 0.0;
  if (has) {
    pr = A.pairs[ch.p0 + p];
    ig = A.owner[ch.p0 + p];
#pragma unroll
    for (int m = 0; m < JM; ++m) T1v[m] = A.T1[((size_t)cand * A.npairs + ch.p0 + p) * 2 + m];
    if (MEAM) {
      const double4 Gq = A.G[(size_t)cand * A.npairs + ch.p0 + p];
      G0 = Gq.x;
      G1 = Gq.y;
      G2 = Gq.z;
      S1 = Gq.w;
    }
  }
  double rho_b = 0.0;
  int type_b = 0;
  if (ba < ch.na) {
    rho_b = A.rho[(size_t)cand * A.natoms + ch.a0 + ba];
    type_b = A.types[ch.a0 + ba];
  }
  const double* box = A.box + 9 * ch.cell;
  const double* lam = A.lam + (size_t)cand * A.natoms * 3;
  const double* cadj = A.cell_adj + ((size_t)cand * A.ncells + ch.cell) * 8;
  const double lamE = cadj[0];
  const double mxx = cadj[1], myy = cadj[2], mzz = cadj[3], mxy = cadj[4], myz = cadj[5], mzx = cadj[6];
  const double kappa = cadj[7];
  const double wcell = A.weight[ch.cell];
  const int j = (int)pr.x;
  int ti = 0, tj = 0;
  double d[3] = {0.0, 0.0, 0.0}, lw[3] = {0.0, 0.0, 0.0};
  if (has) {
    const int ix = (int)(int8_t)(pr.y & 0xff), iy = (int)(int8_t)((pr.y >> 8) & 0xff),
              iz = (int)(int8_t)((pr.y >> 16) & 0xff);
    const double* pj = A.pos + 3 * (size_t)j;
    const double* pi = A.pos + 3 * (size_t)ig;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      double t = __dadd_rn(pj[m], __dmul_rn((double)ix, box[m]));
      t = __dadd_rn(t, __dmul_rn((double)iy, box[3 + m]));
      t = __dadd_rn(t, __dmul_rn((double)iz, box[6 + m]));
      d[m] = __dadd_rn(t, -pi[m]);
      lw[m] = lam[3 * (size_t)ig + m] - lam[3 * (size_t)j + m];
    }
    ti = A.types[ig];
    tj = A.types[j];
  }
  for (int a = tid; a <= ch.na; a += BLK) poff[a] = A.atom_poff[ch.a0 + a] - ch.p0;
  __syncthreads();

  // ---- phase A: pair duals, S_i, partner data
  const int ia = ig - ch.a0;
  int pij = 0;
  double r = 0, invr = 0, ux = 0, uy = 0, uz = 0, fj = 0, dfj = 0, drho = 0, aj = 0, wx = 0, wy = 0, wz = 0;
  int rrec = -1, frec = -1, phrec = -1;  // located rho, f and phi interval records (knot style)
  double trho = 0, tf = 0, tphi = 0, rinv = 0, finv = 0, phinv = 0;
  // jmeam: p and p' of the pair per term, and the located p interval (knot style)
  double pvj[NJ], dpj[NJ], ptt[NJ], pinv[NJ];
  int prec[NJ];
#pragma unroll
  for (int m = 0; m < NJ; ++m) {
    pvj[m] = dpj[m] = ptt[m] = pinv[m] = 0.0;
    prec[m] = -1;
  }
  if (has) {
    r = __dsqrt_rn(__dadd_rn(__dadd_rn(__dmul_rn(d[0], d[0]), __dmul_rn(d[1], d[1])), __dmul_rn(d[2], d[2])));
    invr = 1.0 / r;
    ux = d[0] / r;
    uy = d[1] / r;
    uz = d[2] / r;
    pij = pair_ij(ti, tj, nt);
    double v, dv;
    if (!POINT) {
      // phi is located here too, beside rho and f, for its adjoint in phase D
      const SplHead& Hp = HD[A.pot.fn_off[0] + pij];
      int k;
      if (locate_knot(Hp, kg + Hp.coef_off, r, k, tphi)) {
        phrec = Hp.coef_off + k;
        phinv = Hp.invstep;
      }
    }
    if (A.pot.nodens) {
      // pair styles: phi only
    } else if (POINT) {
      if (eval_radial<POINT, Q>(HD[A.pot.fn_off[1] + tj], IV, r, v, dv)) drho = dv;
      if (MEAM && eval_radial<POINT, Q>(HD[A.pot.fn_off[3] + pij], IV, r, v, dv)) {
        fj = v;
        dfj = dv;
      }
    } else {
      // knot style: the located intervals are kept for the coefficient
      // adjoints of phase D
      const SplHead& Hr = HD[A.pot.fn_off[1] + tj];
      int k;
      if (locate_knot(Hr, kg + Hr.coef_off, r, k, trho)) {
        rrec = Hr.coef_off + k;
        rinv = Hr.invstep;
        poly<Q>(IV[rrec], trho, rinv, v, dv);
        drho = dv;
      }
      if (MEAM) {
        const SplHead& Hf = HD[A.pot.fn_off[3] + pij];
        if (locate_knot(Hf, kg + Hf.coef_off, r, k, tf)) {
          frec = Hf.coef_off + k;
          finv = Hf.invstep;
          poly<Q>(IV[frec], tf, finv, v, dv);
          fj = v;
          dfj = dv;
        }
      }
    }
#pragma unroll
    for (int m = 0; m < JM; ++m) {
      const SplHead& Hp = HD[A.pot.fn_off[5 + 2 * m] + pij];
      if (POINT) {
        if (eval_radial<POINT, Q>(Hp, IV, r, v, dv)) {
          pvj[m] = v;
          dpj[m] = dv;
        }
      } else {
        int k;
        if (locate_knot(Hp, kg + Hp.coef_off, r, k, ptt[m])) {
          prec[m] = Hp.coef_off + k;
          pinv[m] = Hp.invstep;
          poly<Q>(IV[prec[m]], ptt[m], pinv[m], v, dv);
          pvj[m] = v;
          dpj[m] = dv;
        }
      }
    }
    const double vx = mxx * ux + mzx * uz, vy = myy * uy + mxy * ux, vz = mzz * uz + myz * uy;
    wx = lw[0] - r * vx;
    wy = lw[1] - r * vy;
    wz = lw[2] - r * vz;
    aj = wx * ux + wy * uy + wz * uz;
    double sv = aj * drho;
    if (MEAM) {
      sv += wx * G0 + wy * G1 + wz * G2;
      const double pv[kPartnerFields] = {ux,      uy, uz, fj, aj * dfj, invr * (wx - aj * ux), invr * (wy - aj * uy),
                                         invr * (wz - aj * uz)};
#pragma unroll
      for (int q = 0; q < kPartnerFields / 2; ++q) pd[q][p] = make_double2(pv[2 * q], pv[2 * q + 1]);
      push[0][p] = 0.0;
#pragma unroll
      for (int m = 0; m < JM; ++m) {
        pd[kPartnerFields / 2 + m][p] = make_double2(pvj[m], aj * dpj[m]);
        push[2 + m][p] = 0.0;
      }
    }
    push[1][p] = sv;  // staged for S_i; cleared again below
    if (GP) ptyp[p] = tj;
  }
  __syncthreads();

  // ---- phase B: S_i summed by four lanes per atom in a fixed order (each
  // staged entry read and cleared by one lane), then the atom's density
  // sensitivity and embedding adjoint.
  double Sa = 0.0;
  if (ba < ch.na)
    for (int q = poff[ba] + (tid & 3); q < poff[ba + 1]; q += 4) {
      Sa += push[1][q];
      push[1][q] = 0.0;
    }
  Sa += dpp_move<kDppXor1>(Sa);
  Sa += dpp_move<kDppXor2>(Sa);
  if (ba < ch.na && (tid & 3) == 0 && A.pot.nodens) {
    psi[ba] = 0.0;  // pair styles: no density, no embedding
    Up[ba] = 0.0;
  } else if (ba < ch.na && (tid & 3) == 0) {
    const int a = ba;
    const double rho = rho_b;
    const int su = A.pot.fn_off[2] + type_b;
    const SplHead& HU = HD[su];
    const bool below = rho < HU.x0, above = rho > HU.xn;
    const bool out = below || above;
    const double bound = below ? HU.x0 : HU.xn;
    const double rt = (out && !A.pot.embed_extrap) ? bound : rho;
    double dU, d2U;
    (void)eval_point<Q>(HU, IV, rt, dU, d2U);
    double ps = 0.0;
    if (!out || A.pot.embed_extrap) ps = lamE * dU + Sa * d2U;
    if (out) ps += kappa * wcell * kDummyWeight * 10. * 2.0 * (rho - bound);
    psi[a] = ps;
    Up[a] = dU;
    acc_point<Q>(atab, CoefSink{atab, 6, 1}, ncoef, su, HU, rt, lamE, Sa);
  }
  __syncthreads();

  // ---- phase C: unordered triplets {j, k} of each atom.
  // With B_jk = (w_j . d_k - a_j cos) / r_j = V_j . d_k and
  // X = psi g + U' g' (B_jk + B_kj),
  // one visit adds f_k X + U' g A_k to the f-value adjoint of pair j and
  // f_j X + U' g A_j to that of pair k (A = a f'), and the g-coefficient
  // adjoint of weights
  // psi f_j f_k + U' (A_j f_k + A_k f_j) (value) and U' f_j f_k (B_jk + B_kj)
  // (slope).  Styles whose g depends on the centre atom alone and that have
  // no p p q terms take the paired sweep below (two own pairs per lane, one
  // partner record per step for both).  The others: the lane of pair j visits
  // partners j+1 .. j+(n-1)/2 (mod n), plus j+n/2 for the first n/2 lanes
  // when n is even (the forward pass's round-robin); the partner sums go
  // through LDS atomics on distinct k.
  double Phi = 0.0, Phip = 0.0;
  double Rq[NJ];  // jmeam: p-value adjoint of the lane's pair per term, own visits
#pragma unroll
  for (int m = 0; m < NJ; ++m) Rq[m] = 0.0;
  if (MEAM && has) {
    // knot style: the rho, f and phi intervals located in phase A wait in the
    // staging slot (free from phase B to the next chunk) for phase D, three
    // 16-bit record indices (0xffff: none; the host keeps ncoef below it)
    if (!POINT)
      reinterpret_cast<int2*>(&push[1][0])[p] =
          make_int2((int)(((unsigned)rrec & 0xffffu) | ((unsigned)frec << 16)), phrec);
    const double ps = psi[ia], up = Up[ia];
    const int kb = poff[ia], ke = poff[ia + 1];
    const int n = ke - kb, l = p - kb;
    const int gfix = GP ? ti * (nt * (nt + 1) / 2) : ti;  // g of the atom's type (and first pair with GP)
    const GSpl G0 = gspl_of(GP ? GH[gfix] : HD[A.pot.fn_off[4] + ti], A.gco0);
    const double Aj = aj * dfj;
    const double Vjx = invr * (wx - aj * ux), Vjy = invr * (wy - aj * uy), Vjz = invr * (wz - aj * uz);
    // lane constants: U' A_j, U' f_j, and U' / step of the atom's g
    const double upAj = up * Aj, upfj = up * fj, upi0 = up * G0.inv;
    const int kmax0 = (int)G0.kmax;
    if constexpr (kCouples) {
    // One triplet {j, k} of the sweep in two steps: locate (the cosine of own
    // pair j's direction u with partner k's, the g interval and shift, the
    // interval's coefficients) and apply (g and g' at the cosine, pair j's
    // f-value adjoint share added to Racc, the g-coefficient adjoints; returns
    // pair k's share).  SPAN: the atom's g spline spans [-1, 1] (knot style),
    // so no term is cut and the interval is min(trunc(u), n - 2); otherwise
    // the interval is the clamped floor and, knot style, a term is absent for
    // a cosine at or beyond x_n (1 + 1e-8) unless every g spline reaches
    // beyond 1.
    struct TriV {
      double t;
      int fl;
      bool ex, keep;
      double2 c01, c23, c45;
    };
    const bool gcut = !POINT && !A.pot.g_keep_all;
    auto tri_locate = [&](double oux, double ouy, double ouz, const double4& Pa, TriV& v, auto span) {
      const double cs = oux * Pa.x + ouy * Pa.y + ouz * Pa.z;
      v.ex = false;
      v.keep = true;
      if (decltype(span)::value) {
        const double u = (cs - G0.x0) * G0.inv;
        const int k = (int)u;
        v.fl = k < kmax0 ? k : kmax0;
        v.t = u - (double)v.fl;
        const double2* C = gt2 + G0.rec + v.fl;
        v.c01 = C[0];
        v.c23 = C[A.ngco];
        if (Q) v.c45 = C[2 * A.ngco];
        return;
      }
      if (POINT && (cs < G0.x0 || cs > G0.xn)) {
        v.ex = true;
        const bool left = cs < G0.x0;
        v.t = cs - (left ? G0.x0 : G0.xn);
        v.fl = left ? 0 : 1;
        return;
      }
      const double u = (cs - G0.x0) * G0.inv;
      int fl = (int)floor(u);
      fl = fl < 0 ? 0 : (fl > kmax0 ? kmax0 : fl);
      v.fl = fl;
      v.t = u - (double)fl;
      if (gcut) v.keep = cs < G0.xcut;
      const double2* C = gt2 + G0.rec + fl;
      v.c01 = C[0];
      v.c23 = C[A.ngco];
      if (Q) v.c45 = C[2 * A.ngco];
    };
    auto tri_apply = [&](double oux, double ouy, double ouz, double of, double oA, double oVx, double oVy,
                         double oVz, const double4& Pa, const double4& Pb, const TriV& v, double& Racc,
                         auto span) -> double {
      const double t = v.t;
      double gv, dgt;
      if (POINT && v.ex) {
        const bool left = v.fl == 0;
        const double g1 = left ? G0.gL : G0.gR, g2 = left ? G0.g2L : G0.g2R;
        gv = (left ? G0.y0 : G0.yn) + g1 * t + 0.5 * g2 * t * t;
        dgt = (g1 + g2 * t) / G0.inv;
      } else {
        horner_d<Q>(v.c01, v.c23, v.c45, t, gv, dgt);
      }
      double fk = Pa.w, Ak = Pb.x;
      if (!decltype(span)::value && !v.keep) {
        gv = 0.0;
        dgt = 0.0;
        fk = 0.0;
        Ak = 0.0;
      }
      const double Bs = fma(oVx, Pa.x, fma(oVy, Pa.y, fma(oVz, Pa.z, fma(Pb.y, oux, fma(Pb.z, ouy, Pb.w * ouz)))));
      const double X = fma(ps, gv, upi0 * dgt * Bs);
      const double ugv = up * gv;
      const double ff = of * fk;
      Racc = fma(fk, X, fma(ugv, Ak, Racc));
      const double al = fma(ps, ff, up * fma(oA, fk, of * Ak));
      if (POINT && v.ex) {
        const double be = up * ff * Bs;
        double* ghead = atab + ncoef * 6 + (A.pot.fn_off[4] + ti) * 6;
        atomicAdd(ghead + v.fl, al);
        atomicAdd(ghead + 2 + v.fl, al * t + be);
        if (G0.quintic) atomicAdd(ghead + 4 + v.fl, 0.5 * al * t * t + be * t);
      } else {
        const double bt = upi0 * (ff * Bs);
        double* cp = gacc + G0.rec * NC * R + (tid & (R - 1)) + v.fl * (NC * R);
        const double m1 = fma(al, t, bt);
        const double b2 = bt * t;
        const double m2 = fma(t, m1, b2);
        const double b3 = b2 * t;
        const double m3 = fma(t, m2, b3);
        atomicAdd(cp, al);
        atomicAdd(cp + R, m1);
        atomicAdd(cp + 2 * R, m2);
        atomicAdd(cp + 3 * R, m3);
        if (Q) {
          const double b4 = b3 * t;
          const double m4 = fma(t, m3, b4);
          atomicAdd(cp + 4 * R, m4);
          atomicAdd(cp + 5 * R, fma(t, m4, b4 * t));
        }
      }
      return fma(of, X, ugv * oA);
    };
    // Couples: the atom's local pairs 2c and 2c + 1.  Both lanes of a couple
    // hold both pairs and visit one partner pair per step for the two of them,
    // so one partner-record read and one push serve two triplets.  The couples
    // of an atom run the round-robin among themselves (couple c visits couples
    // c + 1 .. c + (mc - 1) / 2, plus c + mc / 2 for the first mc / 2 when mc
    // is even): the even lane takes the visited couple's first pair, the odd
    // lane its second.  The even lane also takes the couple's own triplet
    // {2c, 2c + 1} and, when n is odd, the odd lane visits the last pair,
    // which is in no couple and visits nobody.
    auto sweep = [&](auto span) {
      const int mc = n >> 1;
      double Rs = 0.0;
      if (l < 2 * mc) {
        const int cpl = l >> 1, e = l & 1;
        const int bud = e ? p - 1 : p + 1;
        const double2 b0 = pd[0][bud], b1 = pd[1][bud], b2 = pd[2][bud], b3 = pd[3][bud];
        const double4 Ba = make_double4(b0.x, b0.y, b1.x, b1.y), Bb = make_double4(b2.x, b2.y, b3.x, b3.y);
        double Rb = 0.0;
        if (e == 0) {  // the couple's own triplet: the buddy is the partner
          TriV v;
          tri_locate(ux, uy, uz, Ba, v, span);
          Rb = tri_apply(ux, uy, uz, fj, Aj, Vjx, Vjy, Vjz, Ba, Bb, v, Rs, span);
        }
        const bool odd = e == 1 && (n & 1);
        const int trips = ((mc - 1) >> 1) + (((mc & 1) == 0 && cpl < (mc >> 1)) ? 1 : 0) + (odd ? 1 : 0);
        int cn = odd ? cpl : (cpl + 1 == mc ? 0 : cpl + 1);
        int kq = odd ? kb + n - 1 : kb + 2 * cn + e;
        for (int m = 0; m < trips; ++m) {
          const double2 q0 = pd[0][kq], q1 = pd[1][kq], q2 = pd[2][kq], q3 = pd[3][kq];
          const double4 Pa = make_double4(q0.x, q0.y, q1.x, q1.y), Pb = make_double4(q2.x, q2.y, q3.x, q3.y);
          // both triplets' g coefficients are read before either's atomics
          TriV v0, v1;
          tri_locate(ux, uy, uz, Pa, v0, span);
          tri_locate(Ba.x, Ba.y, Ba.z, Pa, v1, span);
          const double sh = tri_apply(ux, uy, uz, fj, Aj, Vjx, Vjy, Vjz, Pa, Pb, v0, Rs, span) +
                            tri_apply(Ba.x, Ba.y, Ba.z, Ba.w, Bb.x, Bb.y, Bb.z, Bb.w, Pa, Pb, v1, Rb, span);
          atomicAdd(&push[0][kq], sh);
          cn = (cn + 1 == mc) ? 0 : cn + 1;
          kq = kb + 2 * cn + e;
        }
        atomicAdd(&push[0][bud], Rb);
      }
      Phi = Rs;
    };
    if (!POINT && A.pot.g_span)
      sweep(std::true_type{});
    else
      sweep(std::false_type{});
    } else {
    const int trips = ((n - 1) >> 1) + (((n & 1) == 0 && l < (n >> 1)) ? 1 : 0);
    double Rj = 0.0;
    GSpl Gq[NJ];  // q (q2) of the atom's type
#pragma unroll
    for (int m = 0; m < JM; ++m) Gq[m] = gspl_of(HD[A.pot.fn_off[6 + 2 * m] + ti], A.gco0);
    // Each visit locates its cosine and reads its g coefficients at the top;
    // the partner record of the next visit is read at the end, behind the
    // atomics (reads stay inside the atom's pairs).  Prefetching a whole
    // visit ahead needs registers that cost a wave per SIMD, and measured
    // slower.
    auto fetch_partner = [&](int kq, AdjVis& v) {
      v.k = kq;
      const double2 p0 = pd[0][kq], p1 = pd[1][kq], p2 = pd[2][kq], p3 = pd[3][kq];
      v.Pa = make_double4(p0.x, p0.y, p1.x, p1.y);
      v.Pb = make_double4(p2.x, p2.y, p3.x, p3.y);
      v.cs = ux * v.Pa.x + uy * v.Pa.y + uz * v.Pa.z;
      v.gi = GP ? gfix + pair_ij(tj, ptyp[kq], nt) : gfix;
    };
    auto locate_load = [&](AdjVis& v) {
      const GSpl G = GP ? gspl_of(GH[v.gi], A.gco0) : G0;
      v.extrap = false;
      v.keep = true;
      if (POINT && (v.cs < G.x0 || v.cs > G.xn)) {
        v.extrap = true;
        const bool left = v.cs < G.x0;
        v.t = v.cs - (left ? G.x0 : G.xn);
        v.fl = left ? 0 : 1;
      } else {
        // interval index clamped as an integer (the cosine lies in [-1, 1])
        const double u = (v.cs - G.x0) * G.inv;
        const int kmax = GP ? (int)G.kmax : kmax0;
        int fl = (int)floor(u);
        fl = fl < 0 ? 0 : (fl > kmax ? kmax : fl);
        v.fl = fl;
        v.t = u - (double)fl;
        if (!POINT) v.keep = v.cs < G.xcut;
      }
      const double2* C = gt2 + G.rec + v.fl;
      v.c01 = C[0];
      v.c23 = C[A.ngco];
      if (Q) v.c45 = C[2 * A.ngco];
    };
    int kl = (l + 1 == n) ? 0 : l + 1;
    AdjVis cur;
    fetch_partner(kb + kl, cur);
    for (int m = 0; m < trips; ++m) {
      kl = (kl + 1 == n) ? 0 : kl + 1;
      locate_load(cur);
      const int k = cur.k;
      const double4 Pa = cur.Pa, Pb = cur.Pb;
      const double cs = cur.cs;
      const int gi = cur.gi;
      const GSpl G = GP ? gspl_of(GH[gi], A.gco0) : G0;
      const double ginv = G.inv;
      const double upi = GP ? up * ginv : upi0;
      const double t = cur.t;
      const int fl = cur.fl;
      const bool extrap = POINT && cur.extrap, keep = cur.keep;
      double gv, dgt;
      if (POINT && extrap) {
        const bool left = fl == 0;
        const double e = t;
        const double g1 = left ? G.gL : G.gR, g2 = left ? G.g2L : G.g2R;
        gv = (left ? G.y0 : G.yn) + g1 * e + 0.5 * g2 * e * e;
        dgt = (g1 + g2 * e) / ginv;
      } else {
        const double2 c01 = cur.c01, c23 = cur.c23;
        if (Q) {
          const double2 c45 = cur.c45;
          gv = c01.x + t * (c01.y + t * (c23.x + t * (c23.y + t * (c45.x + t * c45.y))));
          dgt = c01.y + t * (2.0 * c23.x + t * (3.0 * c23.y + t * (4.0 * c45.x + t * 5.0 * c45.y)));
        } else {
          gv = c01.x + t * (c01.y + t * (c23.x + t * c23.y));
          dgt = c01.y + t * (2.0 * c23.x + t * 3.0 * c23.y);
        }
      }
      if (!keep) {
        gv = 0.0;
        dgt = 0.0;
      }
      const double fk = keep ? Pa.w : 0.0;
      const double Ak = keep ? Pb.x : 0.0;
      // B_jk + B_kj, then X = psi g + U' g' (B_jk + B_kj) with g' = dgt / step
      const double Bs = fma(Vjx, Pa.x, fma(Vjy, Pa.y, fma(Vjz, Pa.z, fma(Pb.y, ux, fma(Pb.z, uy, Pb.w * uz)))));
      const double X = fma(ps, gv, upi * dgt * Bs);
      const double ugv = up * gv;
      const double ff = fj * fk;
      Rj = fma(fk, X, fma(ugv, Ak, Rj));
      atomicAdd(&push[0][k], fma(fj, X, ugv * Aj));
      const double al = fma(ps, ff, fma(upAj, fk, upfj * Ak));
      if (POINT && extrap) {
        const double be = up * ff * Bs;
        const bool left = fl == 0;
        const double e = t;
        double* ghead = atab + ncoef * 6 + (A.pot.fn_off[4] + gi) * 6;
        atomicAdd(ghead + (left ? 0 : 1), al);
        atomicAdd(ghead + (left ? 2 : 3), al * e + be);
        if (G.quintic) atomicAdd(ghead + (left ? 4 : 5), 0.5 * al * e * e + be * e);
      } else {
        // power-basis moments al t^m + m bt t^(m-1) by Horner steps,
        // m_k = t m_(k-1) + bt t^(k-1), bt the slope weight per unit t
        const double bt = upi * (ff * Bs);
        double* cp = gacc + G.rec * NC * R + (tid & (R - 1)) + fl * (NC * R);
        const double m1 = fma(al, t, bt);
        const double b2 = bt * t;
        const double m2 = fma(t, m1, b2);
        const double b3 = b2 * t;
        const double m3 = fma(t, m2, b3);
        atomicAdd(cp, al);
        atomicAdd(cp + R, m1);
        atomicAdd(cp + 2 * R, m2);
        atomicAdd(cp + 3 * R, m3);
        if (Q) {
          const double b4 = b3 * t;
          const double m4 = fma(t, m3, b4);
          atomicAdd(cp + 4 * R, m4);
          atomicAdd(cp + 5 * R, fma(t, m4, b4 * t));
        }
      }
      // jmeam: X_q = lambda_E q + q' (B_jk + B_kj); pair j gains p_k X_q + q A_k,
      // pair k gains p_j X_q + q A_j (A = a p'); q-coefficient adjoint of weights
      // lambda_E p_j p_k + A_j p_k + A_k p_j (value) and p_j p_k (B_jk + B_kj) (slope)
#pragma unroll
      for (int m = 0; m < JM; ++m) {
        double qv, dqt, tq;
        int flq;
        bool exq, kq;
        ang_at<POINT, Q, NC>(Gq[m], gt2, A.ngco, cs, qv, dqt, tq, flq, exq, kq);
        if (!kq) continue;
        const double2 pq = pd[kPartnerFields / 2 + m][k];
        const double pk = pq.x, Apk = pq.y;
        const double pj = pvj[m], Apj = aj * dpj[m];
        const double Xq = lamE * qv + dqt * Gq[m].inv * Bs;
        Rq[m] += pk * Xq + qv * Apk;
        atomicAdd(&push[2 + m][k], pj * Xq + qv * Apj);
        const double alq = lamE * pj * pk + (Apj * pk + Apk * pj);
        const double beq = pj * pk * Bs;
        if (POINT && exq) {
          const bool left = flq == 0;
          double* qhead = atab + ncoef * 6 + (A.pot.fn_off[6 + 2 * m] + ti) * 6;
          atomicAdd(qhead + (left ? 0 : 1), alq);
          atomicAdd(qhead + (left ? 2 : 3), alq * tq + beq);
          if (Gq[m].quintic) atomicAdd(qhead + (left ? 4 : 5), 0.5 * alq * tq * tq + beq * tq);
        } else {
          const double bt = beq * Gq[m].inv;
          const double t2 = tq * tq, t3 = t2 * tq;
          double* cp = gacc + Gq[m].rec * NC * R + (tid & (R - 1)) + flq * (NC * R);
          atomicAdd(cp, alq);
          atomicAdd(cp + R, alq * tq + bt);
          atomicAdd(cp + 2 * R, alq * t2 + 2.0 * bt * tq);
          atomicAdd(cp + 3 * R, alq * t3 + 3.0 * bt * t2);
          if (Q) {
            const double t4 = t3 * tq;
            atomicAdd(cp + 4 * R, alq * t4 + 4.0 * bt * t3);
            atomicAdd(cp + 5 * R, alq * t4 * tq + 5.0 * bt * t4);
          }
        }
      }
      fetch_partner(kb + kl, cur);  // the next visit's partner, read behind this visit's atomics
    }
    Phi = Rj;
    }  // kCouples
  }
  // ---- phase D: pair functions.  Pairs of similar length hit the same
  // interval, so their coefficient adjoints go to RR lane replicas of the
  // radial records, laid out [replica][record][coef] in the partner-data
  // area, which the sweep no longer reads.
  __syncthreads();
  const int RR = A.rrep_log2 >= 0 ? 1 << A.rrep_log2 : 0;
  const int RS = A.rrep_stride;
  double* rrep = &pd[0][0].x;
  for (int q = tid; q < RS * RR; q += BLK) rrep[q] = 0.0;
  if (MEAM && has) {
    Phi += push[0][p];
    if (POINT) Phip = Up[ia] * aj * S1;
#pragma unroll
    for (int m = 0; m < JM; ++m) Rq[m] += push[2 + m][p];
  }
  __syncthreads();
  if (has) {
    const CoefSink sk = RR ? CoefSink{rrep + (tid & (RR - 1)) * RS, NC, 1} : CoefSink{atab, 6, 1};
    const double ps = psi[ia], up = Up[ia];
    const int sphi = A.pot.fn_off[0] + pij, srho = A.pot.fn_off[1] + tj;
    if (POINT) {
      acc_radial<POINT, Q>(atab, sk, ncoef, sphi, HD[sphi], IV, r, 0.5 * lamE, 0.5 * aj);
      if (!A.pot.nodens) acc_radial<POINT, Q>(atab, sk, ncoef, srho, HD[srho], IV, r, ps, aj * up);
      if (MEAM) {
        const int sf = A.pot.fn_off[3] + pij;
        acc_radial<POINT, Q>(atab, sk, ncoef, sf, HD[sf], IV, r, Phi, Phip);
      }
    } else {
      if (MEAM) {
        // phi, rho and f at the intervals phase A located, their shifts
        // recomputed from r (fewer values live across the sweep)
        const int2 rf = reinterpret_cast<const int2*>(&push[1][0])[p];
        const int rr = (int)((unsigned)rf.x & 0xffffu), fr = (int)((unsigned)rf.x >> 16), pr = rf.y;
        if (pr >= 0) {
          const KnotRec K = kg[pr];
          acc_interval<Q>(sk, pr, (r - K.xk) * K.invh, 0.5 * lamE, 0.5 * aj * HD[sphi].invstep);
        }
        if (rr != 0xffff) {
          const KnotRec K = kg[rr];
          acc_interval<Q>(sk, rr, (r - K.xk) * K.invh, ps, aj * up * HD[srho].invstep);
        }
        if (fr != 0xffff) {
          const KnotRec K = kg[fr];
          acc_interval<Q>(sk, fr, (r - K.xk) * K.invh, Phi, up * aj * S1 * HD[A.pot.fn_off[3] + pij].invstep);
        }
      } else {
        if (phrec >= 0) acc_interval<Q>(sk, phrec, tphi, 0.5 * lamE, 0.5 * aj * phinv);
        if (rrec >= 0) acc_interval<Q>(sk, rrec, trho, ps, aj * up * rinv);
      }
    }
    // jmeam p splines: value adjoint Rq, p' adjoint a T1 (no U' factor)
#pragma unroll
    for (int m = 0; m < JM; ++m) {
      if (POINT) {
        const int sp = A.pot.fn_off[5 + 2 * m] + pij;
        acc_radial<POINT, Q>(atab, sk, ncoef, sp, HD[sp], IV, r, Rq[m], aj * T1v[m]);
      } else if (prec[m] >= 0) {
        acc_interval<Q>(sk, prec[m], ptt[m], Rq[m], aj * T1v[m] * pinv[m]);
      }
    }
  }
  __syncthreads();
  for (int q = tid; q < A.nrad * NC * (RR ? 1 : 0); q += BLK) {
    double sacc = 0.0;
    for (int rep = 0; rep < RR; ++rep) sacc += rrep[rep * RS + q];
    atab[(q / NC) * 6 + q % NC] += sacc;
  }
  // The next chunk overwrites the partner-data area only after its first
  // barrier, which every lane reaches after this reduction.
  }  // chunks
  if (MEAM) {
    for (int q = tid; q < A.ngco * NC; q += BLK) {
      double sacc = 0.0;
      for (int rep = 0; rep < R; ++rep) sacc += gacc[(size_t)q * R + rep];
      atab[(A.gco0 + q / NC) * 6 + q % NC] += sacc;
    }
  }
  __syncthreads();
  double* slot = A.slots + (size_t)cand * A.ntab * gridDim.x + blockIdx.x;
  for (int q = tid; q < A.ntab; q += BLK) slot[(size_t)q * gridDim.x] = atab[q];
}

// Sum of every candidate's per-workgroup partial tables: one 256-lane
// workgroup per table entry, lane l adds slots l, l + 256, ..., a fixed
// shuffle tree adds each wave's lanes and lane 0 the four waves in order, so
// the order never changes between runs.
__global__ void __launch_bounds__(256) k_slotsum(const double* slots, int ntab, int nslot, double* out) {
  __shared__ double red[4];
  const int q = blockIdx.x, cand = blockIdx.y, tid = threadIdx.x;
  const double* s = slots + ((size_t)cand * ntab + q) * nslot;
  double v = 0.0;
  for (int sl = tid; sl < nslot; sl += 256) v += s[sl];
  const double t = block_sum<256>(v, red);
  if (tid == 0) out[(size_t)cand * ntab + q] = t;
}

// One 64-lane workgroup per (spline, candidate): summed adjoint table, penalties,
// interval coefficients -> (y, y'', y'''') -> knot record.
__global__ void __launch_bounds__(64) k_chain(const ChainArgs A) {
  const int s = blockIdx.x, cand = blockIdx.y, lane = threadIdx.x;
  const int n = A.nknots[s], ni = n - 1, off = A.coef_off[s];
  __shared__ double cb[(kMaxKnots - 1) * 6], hb[6], yb[kMaxKnots], db[kMaxKnots], qb[kMaxKnots],
      rb[kMaxKnots + 4];
  const double* tab = A.tab + (size_t)cand * A.ntab;
  for (int q = lane; q < ni * 6 + 6; q += 64) {
    const int idx = q < ni * 6 ? off * 6 + q : A.ncoef * 6 + s * 6 + (q - ni * 6);
    const double sacc = tab[idx];
    if (q < ni * 6) cb[q] = sacc; else hb[q - ni * 6] = sacc;
  }
  __syncthreads();
  if (lane == 0) {
    const char* tb = A.tables + (size_t)cand * A.table_bytes;
    const SplHead* HD = reinterpret_cast<const SplHead*>(tb);
    const SplInterval* IV = reinterpret_cast<const SplInterval*>(tb + sizeof(SplHead) * A.nspl);
    const SplHead& H = HD[s];
    const double* kv = A.knots + (size_t)cand * A.nknot;
    if (!A.meam && !A.nodens && s >= A.u0 && s < A.u0 + A.nt) {
      // U'(midpoint) penalty, src/pots/pot_eam_spline/pot_eam_spline.cpp:211-218
      const double mid = 0.5 * (H.x0 + H.xn);
      double d, d2;
      if (H.basis == 5) eval_point<true>(H, IV, mid, d, d2);
      else eval_point<false>(H, IV, mid, d, d2);
      const double be = 2.0 * (kDummyWeight * d) * kDummyWeight;
      const double delta = mid - H.x0;
      int k = (int)__dmul_rn(delta, H.invstep);
      if (k >= H.n - 1) k = H.n - 2;
      if (k < 0) k = 0;
      const double t = __dmul_rn(__dadd_rn(delta, -__dmul_rn((double)k, H.step)), H.invstep);
      const double bt = be * H.invstep;
      double tp = 1.0;
      for (int m = 1; m < 6; ++m) {
        cb[k * 6 + m] += m * bt * tp;
        tp *= t;
      }
    }
    const double h = H.step;
    const double H2 = h * h / 6.0, H4 = (h * h * h * h) / 360.0;
    for (int k = 0; k < n; ++k) yb[k] = db[k] = qb[k] = 0.0;
    for (int k = 0; k < ni; ++k) {
      const double* c = cb + 6 * k;
      yb[k] += c[0] - c[1];
      yb[k + 1] += c[1];
      db[k] += H2 * (-2.0 * c[1] + 3.0 * c[2] - c[3]);
      db[k + 1] += H2 * (-c[1] + c[3]);
      qb[k] += H4 * (8.0 * c[1] - 20.0 * c[3] + 15.0 * c[4] - 3.0 * c[5]);
      qb[k + 1] += H4 * (7.0 * c[1] - 10.0 * c[3] + 3.0 * c[5]);
    }
    yb[0] += hb[0];
    yb[n - 1] += hb[1];
    for (int k = 0; k < n; ++k) rb[k] = yb[k];
    rb[n] = hb[2];
    rb[n + 1] = hb[3];
    rb[n + 2] = A.basis[s] == 5 ? hb[4] : 0.0;
    rb[n + 3] = A.basis[s] == 5 ? hb[5] : 0.0;
    if (A.meam && A.fper && s >= A.f0 && s < A.f0 + A.nf) {
      // per-spline f-max penalty (src/pots/pot_gmeam_spline/pot_gmeam_spline.cpp:275-280):
      // the knot holding this spline's signed max |y| moves its term
      const double* y = kv + A.koff[s];
      double m = 0.0;
      int km = -1;
      for (int k = 0; k < n; ++k)
        if (fabs(m) < fabs(y[k])) {
          m = y[k];
          km = k;
        }
      if (km >= 0) {
        const double am = fabs(m);
        const double pe = kDummyWeight * 25. * (1.0 - am) * (1.0 - am);
        rb[km] += -2.0 * pe * kDummyWeight * 25. * 2.0 * (1.0 - am) * (m > 0 ? 1.0 : -1.0);
      }
    } else if (A.meam && !A.fper) {
      // f-max penalty (src/pots/pot_meam_spline/pot_meam_spline.cpp:274-278),
      // and the p-max ones of the jmeam styles (src/pots/pot_jmeam2/pot_jmeam2.cpp:308-324):
      // only the knot holding the signed max |y| over the group moves its term
      int g0 = -1;
      for (int g = 0; g < A.npgrp; ++g)
        if (s >= A.pgrp0[g] && s < A.pgrp0[g] + A.nf) g0 = A.pgrp0[g];
      double mx = 0.0;
      int ms = -1, mk = -1;
      for (int q = 0; q < (g0 >= 0 ? A.nf : 0); ++q) {
        const int s2 = g0 + q;
        const double* y = kv + A.koff[s2];
        double m = 0.0;
        int km = -1;
        for (int k = 0; k < A.nknots[s2]; ++k)
          if (fabs(m) < fabs(y[k])) {
            m = y[k];
            km = k;
          }
        if (fabs(mx) < fabs(m)) {
          mx = m;
          ms = s2;
          mk = km;
        }
      }
      if (ms == s && mk >= 0) {
        const double m = fabs(mx);
        const double pe = kDummyWeight * 25. * (1.0 - m) * (1.0 - m);
        rb[mk] += -2.0 * pe * kDummyWeight * 25. * 2.0 * (1.0 - m) * (mx > 0 ? 1.0 : -1.0);
      }
    }
  }
  __syncthreads();
  const double* J = A.J + A.joff[s];
  for (int rr = lane; rr < n + 4; rr += 64) {
    double v = rb[rr];
    for (int q = 0; q < n; ++q) v += J[(size_t)q * (n + 4) + rr] * db[q];
    if (A.basis[s] == 5)
      for (int q = 0; q < n; ++q) v += J[(size_t)(n + q) * (n + 4) + rr] * qb[q];
    A.grad[(size_t)cand * A.nknot + A.koff[s] + rr] = v;
  }
}

}  // namespace mz

// ------------------------------------------------------------------------ host

namespace {

// y'' (and y'''' for quintic) of one spline from its knot record; the same
// algorithms as the refresh kernel (Thomas sweep / Gaussian elimination).
void host_refresh(int basis, int n, double left, double right, int natL, int natR, const double* rec, double* ypp,
                  double* yp4) {
  const double step = (right - left) / (n - 1), invstep = 1. / step;
  const double* y = rec;
  const double gL = rec[n], gR = rec[n + 1], g2L = rec[n + 2], g2R = rec[n + 3];
  if (basis == 3) {
    std::vector<double> u(n, 0.0);
    if (natL) {
      ypp[0] = 0;
    } else {
      ypp[0] = -0.5;
      u[0] = (3.0 * invstep) * ((y[1] - y[0]) * invstep - gL);
    }
    for (int i = 1; i < n - 1; ++i) {
      const double p = 0.5 * ypp[i - 1] + 2.0;
      ypp[i] = (-0.5) / p;
      u[i] = (y[i + 1] - y[i]) * invstep - (y[i] - y[i - 1]) * invstep;
      u[i] = (6.0 * u[i] * 0.5 * invstep - 0.5 * u[i - 1]) / p;
    }
    double qn = 0.0, un = 0.0;
    if (!natR) {
      qn = 0.5;
      un = (3.0 * invstep) * (gR - (y[n - 1] - y[n - 2]) * invstep);
    }
    ypp[n - 1] = (un - qn * u[n - 2]) / (qn * ypp[n - 2] + 1.0);
    for (int k = n - 2; k >= 0; --k) ypp[k] = ypp[k] * ypp[k + 1] + u[k];
    for (int k = 0; k < n; ++k) yp4[k] = 0.0;
    return;
  }
  const int nd = 2 * n;
  std::vector<double> A((size_t)nd * nd, 0.0), b(nd, 0.0);
  auto a = [&](int r, int c) -> double& { return A[(size_t)r * nd + c]; };
  for (int i = 1; i < n - 1; ++i) {
    a(i, i - 1) = 1. / 12; a(i, i) = 10. / 12; a(i, i + 1) = 1. / 12;
    a(i, n + i - 1) = 1. / 180; a(i, n + i) = -1. / 90; a(i, n + i + 1) = 1. / 180;
    a(n + i, i - 1) = 1.; a(n + i, i) = -2.; a(n + i, i + 1) = 1.;
    a(n + i, n + i - 1) = 1. / 6; a(n + i, n + i) = 2. / 3; a(n + i, n + i + 1) = 1. / 6;
    b[i] = y[i + 1] - 2 * y[i] + y[i - 1];
  }
  if (natL) {
    a(0, 0) = -1.; a(0, 1) = 1.; a(0, n) = 1. / 3; a(0, n + 1) = 1. / 6;
    a(n, n) = 1.;
  } else {
    a(0, 0) = 1. / 3; a(0, 1) = 1. / 6; a(0, n) = 1. / 45; a(0, n + 1) = 7. / 360;
    b[0] = y[1] - y[0] - gL * step;
    a(n, 0) = 1.; b[n] = g2L * (step * step);
  }
  if (natR) {
    a(n - 1, n - 2) = -1.; a(n - 1, n - 1) = 1.; a(n - 1, 2 * n - 2) = -1. / 6; a(n - 1, 2 * n - 1) = -1. / 3;
    a(2 * n - 1, 2 * n - 1) = 1.;
  } else {
    a(n - 1, n - 2) = -1. / 6; a(n - 1, n - 1) = -1. / 3; a(n - 1, 2 * n - 2) = -7. / 360; a(n - 1, 2 * n - 1) = -1. / 45;
    b[n - 1] = y[n - 1] - y[n - 2] - gR * step;
    a(2 * n - 1, n - 1) = 1.; b[2 * n - 1] = g2R * (step * step);
  }
  for (int k = 0; k < nd; ++k) {
    int p = k;
    double m = std::fabs(a(k, k));
    for (int i = k + 1; i < nd; ++i)
      if (std::fabs(a(i, k)) > m) { m = std::fabs(a(i, k)); p = i; }
    if (p != k) {
      for (int c = 0; c < nd; ++c) std::swap(a(k, c), a(p, c));
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < nd; ++i) {
      const double l = a(i, k) / a(k, k);
      if (l == 0.0) continue;
      a(i, k) = 0.0;
      for (int c = k + 1; c < nd; ++c) a(i, c) -= l * a(k, c);
      b[i] -= l * b[k];
    }
  }
  for (int k = nd - 1; k >= 0; --k) {
    double sacc = b[k];
    for (int c = k + 1; c < nd; ++c) sacc -= a(k, c) * b[c];
    b[k] = sacc / a(k, k);
  }
  for (int i = 0; i < n; ++i) {
    ypp[i] = b[i] * (invstep * invstep);
    yp4[i] = -b[n + i] * (invstep * invstep * invstep * invstep);
  }
}

#define ADJ_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) return mz_fail(MZ_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

const void* adjoint_fn(const mz_ctx* c, int grep_log2, int blk);

int setup_adjoint(mz_ctx* c) {
  mz::AdjointExtra& X = c->adjx;
  if (X.ready) return MZ_OK;
  // refresh Jacobians (linear maps knot record -> y'', y'''')
  std::vector<double> J;
  std::vector<int> joff(c->nspl);
  for (int s = 0; s < c->nspl; ++s) {
    const int n = c->nknots[s];
    joff[s] = (int)J.size();
    std::vector<double> col(2 * (size_t)n * (n + 4), 0.0);
    std::vector<double> rec(n + 4), ypp(n), yp4(n);
    for (int r = 0; r < n + 4; ++r) {
      std::fill(rec.begin(), rec.end(), 0.0);
      rec[r] = 1.0;
      host_refresh(c->basis[s], n, c->x0[s], c->xn[s], c->bc[2 * s] == -1, c->bc[2 * s + 1] == -1, rec.data(),
                   ypp.data(), yp4.data());
      for (int q = 0; q < n; ++q) {
        col[(size_t)q * (n + 4) + r] = ypp[q];
        col[(size_t)(n + q) * (n + 4) + r] = yp4[q];
      }
    }
    J.insert(J.end(), col.begin(), col.end());
  }
  if (X.d_J.upload(J.data(), J.size()) || X.d_joff.upload(joff.data(), joff.size()))
    return mz_fail(MZ_ERR_NOMEM, "adjoint Jacobian upload failed");
  // LDS: compact g coefficients, coefficient adjoints and the lane replicas
  // of the g-coefficient adjoints beside the static arrays of the workgroup
  // size.  Each built (workgroup size, replica count) pair that fits the
  // device's workgroup LDS is scored by its waves resident per CU times a
  // replica weight, 1 for 16 or 32 replicas (the lanes of a wave that share a
  // bank then add into one replica column: no bank conflicts), 3/4 for 8 and
  // 1/4 for 1; ties go to the smaller workgroup, then to more replicas.  The
  // weights reproduce the measured order: ternary 2000-cell workload, 32
  // candidates per launch: 512 x 16 (16 waves) 402 ms, 256 x 8 (12 waves)
  // 515 ms, 256 x 1 (16 waves) and 512 x 32 (8 waves) 920 ms; one-element
  // 256-cell workload: 256 x 32 about 8 % ahead of 512 x 32 (both 16 waves).
  // phase D finds the located interval records as 16-bit indices
  if (c->ncoef >= 0xffff) return mz_fail(MZ_ERR_CAPACITY, "more than 65534 spline intervals");
  X.ntab = c->ncoef * 6 + c->nspl * 6;
  X.gco0 = c->meam ? c->coef_off[c->pot.fn_off[4]] : 0;
  X.ngco = c->meam ? c->ncoef - X.gco0 : 0;
  const int nc_g = c->pot.quintic ? 6 : 4;
  const int fixed = (X.ngco * nc_g + X.ntab) * (int)sizeof(double) + c->nspl * (int)sizeof(mz::SplHead) +
                    c->ncoef * (int)sizeof(mz::KnotRec);
  int maxlds = 0, cus = 0;
  if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess || cus < 1)
    return mz_fail(MZ_ERR_HIP, "device attribute query failed");
  struct Cfg {
    int blk, lg;
  };
  const Cfg cfgs[] = {{256, 5}, {256, 3}, {256, 0}, {512, 5}, {512, 4}};
  int best = 0;  // score in quarter waves
  X.fn = nullptr;
  for (const Cfg& f : cfgs) {
    if (!c->meam && f.lg != 0) continue;  // no g spline: no replicas
    const void* fn = adjoint_fn(c, f.lg, f.blk);
    if (!fn) continue;
    hipFuncAttributes fa{};
    if (hipFuncGetAttributes(&fa, fn) != hipSuccess) continue;
    const int dyn = fixed + (c->meam ? X.ngco * nc_g * (1 << f.lg) * 8 : 0);
    if ((long long)fa.sharedSizeBytes + dyn > maxlds) continue;
    int wpc = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&wpc, fn, f.blk, dyn) != hipSuccess || wpc < 1) continue;
    const int weight = !c->meam || f.lg >= 4 ? 4 : (f.lg == 3 ? 3 : 1);
    const int score = wpc * f.blk / 64 * weight;
    if (score > best) {  // strict: earlier entries win ties
      best = score;
      X.fn = fn;
      X.blk = f.blk;
      X.grep_log2 = f.lg;
      X.dyn_bytes = dyn;
      X.resident = wpc * cus;
    }
  }
  if (!X.fn) return mz_fail(MZ_ERR_CAPACITY, "no adjoint workgroup fits the device's workgroup LDS");
  // chunks of at most blk pairs and blk / 4 atoms for the adjoint workgroups
  const int natom_max = X.blk / 4;
  std::vector<mz::Chunk> ch;
  for (int cc = 0; cc < c->ncells; ++cc) {
    int a = c->abase[cc];
    const int aend = a + c->natoms_c[cc];
    while (a < aend) {
      mz::Chunk k{};
      k.cell = cc;
      k.a0 = a;
      k.p0 = c->atom_poff[a];
      int na = 0;
      while (a < aend && na < natom_max) {
        const int np_a = c->atom_poff[a + 1] - c->atom_poff[a];
        if (np_a > X.blk)
          return mz_fail(MZ_ERR_CAPACITY,
                         "atom " + std::to_string(a) + " has more than " + std::to_string(X.blk) + " neighbours");
        if (c->atom_poff[a + 1] - k.p0 > X.blk) break;
        ++a;
        ++na;
      }
      k.na = na;
      k.np = c->atom_poff[a] - k.p0;
      ch.push_back(k);
    }
  }
  std::vector<long long> work(ch.size());
  for (size_t q = 0; q < ch.size(); ++q) {
    long long w = ch[q].np;
    for (int a = ch[q].a0; a < ch[q].a0 + ch[q].na; ++a) w += 2 * c->ntrip_atom[a];
    work[q] = w;
  }
  std::vector<size_t> order(ch.size());
  for (size_t q = 0; q < order.size(); ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work[x] > work[y]; });
  std::vector<mz::Chunk> sorted(ch.size());
  for (size_t q = 0; q < ch.size(); ++q) sorted[q] = ch[order[q]];
  if (X.d_chunks.upload(sorted.data(), sorted.size())) return mz_fail(MZ_ERR_NOMEM, "adjoint chunk upload failed");
  std::vector<int> owner((size_t)c->npairs);
  for (int a = 0; a < c->natoms; ++a)
    for (int q = c->atom_poff[a]; q < c->atom_poff[a + 1]; ++q) owner[q] = a;
  if (X.d_owner.upload(owner.data(), owner.size())) return mz_fail(MZ_ERR_NOMEM, "adjoint owner upload failed");
  X.nchunks = (int)sorted.size();

  // Lane replicas of the radial interval records reuse the partner-data area.
  X.nrad = c->meam ? X.gco0 : c->ncoef;
  X.rrep_stride = ((X.nrad * nc_g + 31) & ~31) + 1;
  X.rrep_log2 = -1;
  for (int lg = 4; lg >= 0 && X.rrep_log2 < 0; --lg)
    if (X.rrep_stride * (1 << lg) <= mz::kPartnerFields * X.blk) X.rrep_log2 = lg;
  X.ready = true;
  return MZ_OK;
}

int ensure_adjoint_buffers(mz_ctx* c, int ncand) {
  mz::AdjointExtra& X = c->adjx;
  if (ncand <= X.cap) return MZ_OK;
  const size_t nc = (size_t)ncand;
  bool bad = X.d_rho.alloc(nc * c->natoms) || X.d_dF.alloc(nc * c->natoms) || X.d_G.alloc(nc * c->npairs * 4) ||
             X.d_lam.alloc(nc * c->natoms * 3) || X.d_cadj.alloc(nc * c->ncells * 8) ||
             X.d_slots.alloc(std::max<size_t>(nc, 8 * (size_t)X.resident) * X.ntab) || X.d_tabsum.alloc(nc * X.ntab) ||
             X.d_grad.alloc(nc * c->nknot) ||
             (c->pot.njm && X.d_T1.alloc(nc * c->npairs * 2));
  if (bad) return mz_fail(MZ_ERR_NOMEM, "device allocation of gradient buffers failed");
  if (X.h_grad) (void)hipHostFree(X.h_grad);
  X.h_grad = nullptr;
  if (hipHostMalloc((void**)&X.h_grad, nc * c->nknot * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return mz_fail(MZ_ERR_NOMEM, "pinned gradient allocation failed");
  X.cap = ncand;
  for (auto& g : X.graphs)
    if (g.ex) (void)hipGraphExecDestroy(g.ex);
  X.graphs.clear();
  return MZ_OK;
}

// The k_adjoint instantiation of a potential: style flags and the lane
// replica count of the g-coefficient adjoints.
// Built pairs (workgroup size, g replicas): 512 with 32 or 16, 256 with 32, 8
// or 1 (the selection list of setup_adjoint).
template <bool M, bool P, bool Q, bool GP, int JM>
const void* adjoint_fn_g(int grep_log2, int blk) {
  if (blk == 512) {
    if (grep_log2 == 5) return reinterpret_cast<const void*>(&mz::k_adjoint<M, P, Q, 32, GP, JM, 512>);
    if (grep_log2 == 4) return reinterpret_cast<const void*>(&mz::k_adjoint<M, P, Q, 16, GP, JM, 512>);
    return nullptr;
  }
  if (grep_log2 == 5) return reinterpret_cast<const void*>(&mz::k_adjoint<M, P, Q, 32, GP, JM, 256>);
  if (grep_log2 == 3) return reinterpret_cast<const void*>(&mz::k_adjoint<M, P, Q, 8, GP, JM, 256>);
  if (grep_log2 == 0) return reinterpret_cast<const void*>(&mz::k_adjoint<M, P, Q, 1, GP, JM, 256>);
  return nullptr;
}

template <bool M, bool P, bool Q>
const void* adjoint_fn_t(const mz::PotDesc& pot, int grep_log2, int blk) {
  if (!M) {
    if (grep_log2 != 0) return nullptr;
    return blk == 512 ? reinterpret_cast<const void*>(&mz::k_adjoint<M, P, Q, 1, false, 0, 512>)
                      : reinterpret_cast<const void*>(&mz::k_adjoint<M, P, Q, 1, false, 0, 256>);
  }
  if (pot.gpair) return adjoint_fn_g<M, P, Q, true, 0>(grep_log2, blk);
  if (pot.njm == 1) return adjoint_fn_g<M, P, Q, false, 1>(grep_log2, blk);
  if (pot.njm == 2) return adjoint_fn_g<M, P, Q, false, 2>(grep_log2, blk);
  return adjoint_fn_g<M, P, Q, false, 0>(grep_log2, blk);
}

const void* adjoint_fn(const mz_ctx* c, int g, int b) {
  const bool q = c->pot.quintic != 0;
  if (c->meam) {
    if (c->point) return q ? adjoint_fn_t<true, true, true>(c->pot, g, b) : adjoint_fn_t<true, true, false>(c->pot, g, b);
    return q ? adjoint_fn_t<true, false, true>(c->pot, g, b) : adjoint_fn_t<true, false, false>(c->pot, g, b);
  }
  if (c->point) return q ? adjoint_fn_t<false, true, true>(c->pot, g, b) : adjoint_fn_t<false, true, false>(c->pot, g, b);
  return q ? adjoint_fn_t<false, false, true>(c->pot, g, b) : adjoint_fn_t<false, false, false>(c->pot, g, b);
}

// Adjoint workgroups per candidate.  The launch holds eight times the
// workgroups the GPU keeps resident (dispatch then evens out the progress of
// the statically assigned chunk sequences: measured against one and four
// times, 4.86 / 5.27 / 4.96 ms on the one-element 256-cell workload at 64
// candidates), at most twice that for a single candidate (the partial tables
// grow with the grid) and no more than there are chunks.  A multiple of eight
// puts workgroup x of every candidate, which walks the same chunks, on the
// same XCD (workgroups go to the eight XCDs round-robin).
int adjoint_nslot(const mz::AdjointExtra& X, int ncand) {
  const int nc = std::max(ncand, 1);
  int n = std::min(8 * X.resident / nc, 2 * X.resident);
  if (n >= 8) n &= ~7;
  return std::max(1, std::min(X.nchunks, n));
}

}  // namespace

int mz_run_forward(mz_ctx* c, int ncand, double* d_errvec, double* d_ce, double* d_cs, double* d_rho, double* d_dF,
                   double* d_G, double* d_lam, double* d_cadj, double* d_T1);
mz::AtomArgs mz_atom_args(const mz_ctx* c);
void mz_timing_add(mz_ctx* c, mz_timing* acc, bool adjoint);
int mz_sub_batch(int ncand, size_t fit);
int mz_ensure_buffers(mz_ctx* c, int ncand);
size_t mz_candidate_budget(const mz_ctx* c);
size_t mz_forward_bytes(const mz_ctx* c);

// Launch sequence of objective + gradient for ncand staged knot vectors.
static int run_grad(mz_ctx* c, int ncand) {
  mz::AdjointExtra& X = c->adjx;
  hipStream_t st = c->stream;
  int rc = mz_run_forward(c, ncand, nullptr, nullptr, nullptr, X.d_rho.p, X.d_dF.p, c->meam ? X.d_G.p : nullptr,
                          X.d_lam.p, X.d_cadj.p, c->pot.njm ? X.d_T1.p : nullptr);
  if (rc) return rc;
  if (c->timing) ADJ_HIP(hipEventRecord(c->ev[5], st));
  mz::AdjArgs a{};
  a.pot = c->pot;
  a.chunks = X.d_chunks.p;
  a.atom_poff = c->d_atom_poff.p;
  a.owner = X.d_owner.p;
  a.pairs = c->d_pairs.p;
  a.pos = c->d_pos.p;
  a.box = c->d_box.p;
  a.types = c->d_types.p;
  a.weight = c->d_weight.p;
  a.tables = c->d_tables.p;
  a.table_bytes = c->table_bytes;
  a.natoms = c->natoms;
  a.ncells = c->ncells;
  a.npairs = c->npairs;
  a.rho = X.d_rho.p;
  a.G = reinterpret_cast<const double4*>(X.d_G.p);
  a.T1 = X.d_T1.p;
  a.lam = X.d_lam.p;
  a.cell_adj = X.d_cadj.p;
  a.slots = X.d_slots.p;
  a.ntab = X.ntab;
  a.gco0 = X.gco0;
  a.ngco = X.ngco;
  a.grep_log2 = X.grep_log2;
  a.nrad = X.nrad;
  a.rrep_log2 = X.rrep_log2;
  a.rrep_stride = X.rrep_stride;
  a.nchunks = X.nchunks;
  const int nslot = adjoint_nslot(X, ncand);
  void* kargs[] = {&a};
  ADJ_HIP(hipLaunchKernel(X.fn, dim3(nslot, ncand), dim3(X.blk), kargs, X.dyn_bytes, st));
  if (c->timing) ADJ_HIP(hipEventRecord(c->ev[6], st));
  hipLaunchKernelGGL(mz::k_slotsum, dim3(X.ntab, ncand), dim3(256), 0, st, X.d_slots.p, X.ntab, nslot, X.d_tabsum.p);
  ADJ_HIP(hipGetLastError());
  mz::ChainArgs ca{};
  ca.nspl = c->nspl;
  ca.ncoef = c->ncoef;
  ca.ntab = X.ntab;
  ca.nknot = c->nknot;
  ca.meam = c->meam;
  ca.nt = c->nt;
  ca.f0 = c->pot.fn_off[3];
  ca.nf = c->nt * (c->nt + 1) / 2;
  ca.fper = c->meam && c->nfpen == ca.nf && (c->style == mz::kGMeamSpline || c->style == mz::kGMeam);
  ca.u0 = c->pot.fn_off[2];
  ca.nodens = c->pot.nodens;
  ca.npgrp = c->meam && !ca.fper ? 1 + c->pot.njm : 0;
  for (int g = 0; g < 3; ++g) ca.pgrp0[g] = c->pot.fn_off[g == 0 ? 3 : 3 + 2 * g];
  ca.nknots = c->d_nknots.p;
  ca.basis = c->d_basis.p;
  ca.koff = c->d_koff.p;
  ca.coef_off = c->d_coef_off.p;
  ca.x0 = c->d_x0.p;
  ca.xn = c->d_xn.p;
  ca.J = X.d_J.p;
  ca.joff = X.d_joff.p;
  ca.tab = X.d_tabsum.p;
  ca.knots = c->d_knots.p;
  ca.tables = c->d_tables.p;
  ca.table_bytes = c->table_bytes;
  ca.grad = X.d_grad.p;
  hipLaunchKernelGGL(mz::k_chain, dim3(c->nspl, ncand), dim3(64), 0, st, ca);
  ADJ_HIP(hipGetLastError());
  if (c->timing) ADJ_HIP(hipEventRecord(c->ev[3], st));
  return MZ_OK;
}

int mz_grad_impl(mz_ctx* c, int ncand, const double* knots, double* F, double* G) {
  if (!c || !knots || !F || !G || ncand < 1) return mz_fail(MZ_ERR_ARG, "null argument or ncand < 1");
  int rc = setup_adjoint(c);
  if (rc) return rc;
  mz::AdjointExtra& X = c->adjx;
  // sub-batches of candidates whose forward and gradient buffers (the stored
  // leg vectors G_p above all) fit the device memory that is free
  const size_t per_cand = mz_forward_bytes(c) + (size_t)c->npairs * (c->pot.njm ? 6 : 4) * sizeof(double) +
                          (size_t)c->natoms * 40 + (size_t)c->ncells * 64 + (size_t)c->nknot * 8 +
                          (size_t)X.ntab * 8 * std::max(1, 8 * X.resident / std::max(ncand, 1));
  const int sub = mz_sub_batch(ncand, mz_candidate_budget(c) / per_cand);
  mz_timing acc{};  // kernel times summed over the sub-batches of a timed call
  for (int c0 = 0; c0 < ncand; c0 += sub) {
    const int nb = std::min(sub, ncand - c0);
    rc = mz_ensure_buffers(c, nb);
    if (rc) return rc;
    rc = ensure_adjoint_buffers(c, nb);
    if (rc) return rc;
    const double* kv = knots + (size_t)c0 * c->nknot;
    const size_t n = (size_t)nb * c->nknot;
    for (size_t q = 0; q < n; ++q)
      if (!std::isfinite(kv[q])) return mz_fail(MZ_ERR_ARG, "knot vector holds a non-finite value");
    std::memcpy(c->h_stage, kv, n * sizeof(double));
    hipStream_t st = c->stream;
    if (c->timing) {
      // stream launches with the kernel events (graph replays keep no event timestamps)
      ADJ_HIP(hipMemcpyAsync(c->d_knots.p, c->h_stage, n * sizeof(double), hipMemcpyHostToDevice, st));
      rc = run_grad(c, nb);
      if (rc) return rc;
      ADJ_HIP(hipMemcpyAsync(c->h_F, c->d_F.p, nb * sizeof(double), hipMemcpyDeviceToHost, st));
      ADJ_HIP(hipMemcpyAsync(X.h_grad, X.d_grad.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
      ADJ_HIP(hipStreamSynchronize(st));
      mz_timing_add(c, &acc, true);
    } else {
      hipGraphExec_t ex = nullptr;
      for (auto& g : X.graphs)
        if (g.ncand == nb) ex = g.ex;
      if (!ex) {
        hipGraph_t graph = nullptr;
        ADJ_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t e1 = hipMemcpyAsync(c->d_knots.p, c->h_stage, n * sizeof(double), hipMemcpyHostToDevice, st);
        int rc2 = run_grad(c, nb);
        hipError_t e2 = hipMemcpyAsync(c->h_F, c->d_F.p, nb * sizeof(double), hipMemcpyDeviceToHost, st);
        hipError_t e3 = hipMemcpyAsync(X.h_grad, X.d_grad.p, n * sizeof(double), hipMemcpyDeviceToHost, st);
        hipError_t e4 = hipStreamEndCapture(st, &graph);
        if (rc2) return rc2;
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess)
          return mz_fail(MZ_ERR_HIP, "graph capture of the gradient evaluation failed");
        hipError_t e5 = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e5 != hipSuccess) return mz_fail(MZ_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e5));
        mz::AdjointExtra::Graph g;
        g.ex = ex;
        g.ncand = nb;
        X.graphs.push_back(g);
      }
      ADJ_HIP(hipGraphLaunch(ex, st));
      ADJ_HIP(hipStreamSynchronize(st));
    }
    std::memcpy(F + c0, c->h_F, nb * sizeof(double));
    std::memcpy(G + (size_t)c0 * c->nknot, X.h_grad, n * sizeof(double));
  }
  if (c->timing) c->last = acc;
  return MZ_OK;
}

void mz_adjoint_drop(mz_ctx* c) {
  mz::AdjointExtra& X = c->adjx;
  for (auto& g : X.graphs)
    if (g.ex) (void)hipGraphExecDestroy(g.ex);
  X.graphs.clear();
  X.d_rho.release(); X.d_dF.release(); X.d_G.release(); X.d_lam.release(); X.d_cadj.release(); X.d_T1.release();
  X.d_slots.release(); X.d_tabsum.release(); X.d_grad.release(); X.d_J.release(); X.d_joff.release(); X.d_chunks.release();
  X.d_owner.release();
  if (X.h_grad) (void)hipHostFree(X.h_grad);
  X.h_grad = nullptr;
  X.cap = 0;
  X.ready = false;
}

extern "C" int mz_objective_grad(mz_ctx* c, const double* knots, double* F, double* dF) {
  return mz_grad_impl(c, 1, knots, F, dF);
}

extern "C" int mz_spline_refresh(int basis, int n, double x0, double xn, int natural_left, int natural_right,
                                 const double* record, double* ypp, double* yp4) {
  if (!record || !ypp || (basis == 5 && !yp4)) return mz_fail(MZ_ERR_ARG, "null argument");
  if ((basis != 3 && basis != 5) || n < 3 || n > mz::kMaxKnots || !(xn > x0))
    return mz_fail(MZ_ERR_ARG, "unsupported spline (basis 3 or 5, 3 <= n <= 32, x0 < xn)");
  std::vector<double> four(n, 0.0);
  host_refresh(basis, n, x0, xn, natural_left != 0, natural_right != 0, record, ypp, yp4 ? yp4 : four.data());
  return MZ_OK;
}

