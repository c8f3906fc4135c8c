// gfx950 kernels of the MEAM/EAM spline objective.
//
//   k_refresh   knot vector -> per-interval power-basis spline tables
//               (Spline3Eq::refresh_basis src/spline3eq.cpp:192-225,
//                Spline5Eq::refresh_basis src/spline5eq.cpp:210-309)
//   k_zero      clears the per-evaluation accumulators
//   k_atoms     one workgroup per chunk of consecutive atoms of a cell: pair
//               pass, on-the-fly triplet pass, embedding, pair and three-body
//               forces and virial (PotMEAMSpline::fast_compute
//               src/pots/pot_meam_spline/pot_meam_spline.cpp:143-270 and its
//               EAM / point-evaluated siblings)
//   k_resid     force assembly from the stored pair force vectors, residual
//               vector and per-block sums of squares
//               (Potential::accumulate_error src/potential.cpp:415-491)
//   k_final     deterministic final sum plus the style penalty
//               (pot_meam_spline.cpp:274-278, pot_eam_spline.cpp:211-218)
//
// FP64 throughout, no global atomics in the forward pass.  The triplet pass is a
// balanced round-robin over each atom's neighbour pairs: every (j,k) pair of
// neighbours is visited once by the lane owning j, which keeps its own leg in
// registers and pushes the partner leg into an LDS accumulator.  Each pair
// stores its force vector T_p once; an atom's force is its own sum minus the
// sum over the pairs that point at it, gathered in k_resid in a fixed order.
#include <hip/hip_runtime.h>

#include "mz_common.h"
#include "mz_kernels.h"
#include "mz_device.h"

namespace mz {

// ---------------------------------------------------------------- k_refresh

// One 64-lane workgroup per (spline, candidate).
__global__ void __launch_bounds__(64) k_refresh(const double* __restrict__ knots, int nknot, const RefreshMeta meta,
                                                char* __restrict__ tables, int table_bytes) {
  const int s = blockIdx.x;
  const int cand = blockIdx.y;
  const int lane = threadIdx.x;
  const int n = meta.nknots[s];
  const int basis = meta.basis[s];
  const double* rec = knots + (size_t)cand * nknot + meta.koff[s];
  SplHead* heads = reinterpret_cast<SplHead*>(tables + (size_t)cand * table_bytes);
  SplInterval* ivs = reinterpret_cast<SplInterval*>(tables + (size_t)cand * table_bytes + sizeof(SplHead) * meta.nspl);

  __shared__ double y[kMaxKnots], x[kMaxKnots], ypp[kMaxKnots], yp4[kMaxKnots];
  __shared__ double A[2 * kMaxKnots][2 * kMaxKnots + 1];
  __shared__ double b[2 * kMaxKnots];
  __shared__ int piv;

  const double left = meta.x0[s], right = meta.xn[s];
  const double step = (right - left) / (n - 1);  // set_spline src/spline3eq.cpp:362-369
  const double invstep = 1. / step;
  const int natL = meta.bc[2 * s] == -1, natR = meta.bc[2 * s + 1] == -1;
  const double gL = rec[n], gR = rec[n + 1], g2L = rec[n + 2], g2R = rec[n + 3];
  if (lane < n) {
    y[lane] = rec[lane];
    x[lane] = (lane == 0) ? left : (lane == n - 1 ? right : __dadd_rn(left, __dmul_rn(step, (double)lane)));
  }
  __syncthreads();
  if (basis == 3) {
    if (lane == 0) {
      double u[kMaxKnots];
      for (int i = 0; i < n; ++i) u[i] = 0.0;
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
      double qn, un;
      if (natR) {
        qn = un = 0.0;
      } else {
        qn = 0.5;
        un = (3.0 * invstep) * (gR - (y[n - 1] - y[n - 2]) * invstep);
      }
      ypp[n - 1] = (un - qn * u[n - 2]) / (qn * ypp[n - 2] + 1.0);
      for (int k = n - 2; k >= 0; --k) ypp[k] = ypp[k] * ypp[k + 1] + u[k];
      for (int k = 0; k < n; ++k) yp4[k] = 0.0;
    }
  } else {
    // 2n x 2n system of the quintic spline, solved by Gaussian elimination
    // with partial pivoting (rows spread over the lanes).
    const int nd = 2 * n;
    for (int r = lane; r < nd; r += 64) {
      for (int c = 0; c < nd; ++c) A[r][c] = 0.0;
      b[r] = 0.0;
    }
    __syncthreads();
    if (lane >= 1 && lane < n - 1) {
      const int i = lane;
      A[i][i - 1] = 1. / 12; A[i][i] = 10. / 12; A[i][i + 1] = 1. / 12;
      A[i][n + i - 1] = 1. / 180; A[i][n + i] = -1. / 90; A[i][n + i + 1] = 1. / 180;
      A[n + i][i - 1] = 1.; A[n + i][i] = -2.; A[n + i][i + 1] = 1.;
      A[n + i][n + i - 1] = 1. / 6; A[n + i][n + i] = 2. / 3; A[n + i][n + i + 1] = 1. / 6;
      b[i] = y[i + 1] - 2 * y[i] + y[i - 1];
      b[n + i] = 0.;
    }
    if (lane == 0) {
      if (natL) {
        A[0][0] = -1.; A[0][1] = 1.; A[0][n] = 1. / 3; A[0][n + 1] = 1. / 6; b[0] = 0.;
        A[n][n] = 1.; b[n] = 0.;
      } else {
        A[0][0] = 1. / 3; A[0][1] = 1. / 6; A[0][n] = 1. / 45; A[0][n + 1] = 7. / 360;
        b[0] = y[1] - y[0] - gL * step;
        A[n][0] = 1.; b[n] = g2L * (step * step);
      }
      if (natR) {
        A[n - 1][n - 2] = -1.; A[n - 1][n - 1] = 1.; A[n - 1][2 * n - 2] = -1. / 6; A[n - 1][2 * n - 1] = -1. / 3;
        b[n - 1] = 0.;
        A[2 * n - 1][2 * n - 1] = 1.; b[2 * n - 1] = 0.;
      } else {
        A[n - 1][n - 2] = -1. / 6; A[n - 1][n - 1] = -1. / 3; A[n - 1][2 * n - 2] = -7. / 360;
        A[n - 1][2 * n - 1] = -1. / 45;
        b[n - 1] = y[n - 1] - y[n - 2] - gR * step;
        A[2 * n - 1][n - 1] = 1.; b[2 * n - 1] = g2R * (step * step);
      }
    }
    __syncthreads();
    for (int k = 0; k < nd; ++k) {
      if (lane == 0) {
        int p = k;
        double m = fabs(A[k][k]);
        for (int i = k + 1; i < nd; ++i)
          if (fabs(A[i][k]) > m) { m = fabs(A[i][k]); p = i; }
        piv = p;
      }
      __syncthreads();
      const int p = piv;
      if (p != k) {
        for (int c = lane; c < nd; c += 64) { double t = A[k][c]; A[k][c] = A[p][c]; A[p][c] = t; }
        if (lane == 0) { double t = b[k]; b[k] = b[p]; b[p] = t; }
      }
      __syncthreads();
      const double pv = A[k][k];
      for (int i = k + 1 + lane; i < nd; i += 64) {
        const double l = A[i][k] / pv;
        if (l != 0.0) {
          A[i][k] = 0.0;
          for (int c = k + 1; c < nd; ++c) A[i][c] -= l * A[k][c];
          b[i] -= l * b[k];
        }
      }
      __syncthreads();
    }
    if (lane == 0) {
      for (int k = nd - 1; k >= 0; --k) {
        double sacc = b[k];
        for (int c = k + 1; c < nd; ++c) sacc -= A[k][c] * b[c];
        b[k] = sacc / A[k][k];
      }
      for (int i = 0; i < n; ++i) {
        ypp[i] = b[i] * (invstep * invstep);
        yp4[i] = -b[n + i] * (invstep * invstep * invstep * invstep);
      }
    }
  }
  __syncthreads();
  // interval records in power basis of the shift
  const int off = meta.coef_off[s];
  const double h = step;
  const double H2 = h * h / 6.0, H4 = (h * h * h * h) / 360.0;
  if (lane < n - 1) {
    const int k = lane;
    const double y0 = y[k], y1 = y[k + 1], d0 = ypp[k], d1 = ypp[k + 1], q0 = yp4[k], q1 = yp4[k + 1];
    SplInterval I;
    I.c[0] = y0;
    I.c[1] = (y1 - y0) + H2 * (-2.0 * d0 - d1) + H4 * (8.0 * q0 + 7.0 * q1);
    I.c[2] = 3.0 * H2 * d0;
    I.c[3] = H2 * (d1 - d0) + H4 * (-20.0 * q0 - 10.0 * q1);
    I.c[4] = 15.0 * H4 * q0;
    I.c[5] = H4 * (3.0 * q1 - 3.0 * q0);
    I.xk = x[k];
    I.invh = 1.0 / (x[k + 1] - x[k]);
    ivs[off + k] = I;
  }
  if (lane == 0) {
    SplHead Hd;
    Hd.x0 = left;
    Hd.xn = right;
    Hd.step = step;
    Hd.invstep = invstep;
    Hd.xcut = x[n - 1] * (1. + 1.e-8);
    Hd.y0 = y[0];
    Hd.yn = y[n - 1];
    Hd.gL = gL;
    Hd.gR = gR;
    Hd.g2L = basis == 5 ? g2L : 0.0;
    Hd.g2R = basis == 5 ? g2R : 0.0;
    Hd.invstep2 = invstep * invstep;
    Hd.n = n;
    Hd.basis = basis;
    Hd.coef_off = off;
    Hd.pad = 0;
    Hd.pad2[0] = Hd.pad2[1] = 0.0;
    heads[s] = Hd;
  }
}

// ------------------------------------------------------------------- k_zero

__global__ void k_zero(double* __restrict__ p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.0;
}

// ------------------------------------------------------------------ k_atoms

// LDS layout of one workgroup (static part).
template <int BLOCK, int PPL, int MAXA>
struct AtomSmem {
  double kd[BLOCK * PPL][6];   // per pair: unit vector, f, f', 1/r (48-byte records)
  double gl[3][BLOCK * PPL];   // three-body leg sums pushed by the partner lanes
  double rho[MAXA];
  double dF[MAXA];
  double fi[MAXA][3];
  int poff[MAXA + 1];
  double red8[8][BLOCK / 64];
};

template <int BLOCK, int PPL, int MAXA, bool MEAM, bool POINT, bool Q>
__global__ void __launch_bounds__(BLOCK) k_atoms(const AtomArgs args) {
  extern __shared__ __attribute__((aligned(64))) char dyn[];
  __shared__ __attribute__((aligned(16))) AtomSmem<BLOCK, PPL, MAXA> sm;
  const int cand = blockIdx.y;
  const Chunk ch = args.chunks[blockIdx.x];
  const int tid = threadIdx.x;

  // stage this candidate's spline table in LDS, plus the g intervals again as
  // 48-byte records (6 coefficients) so that lanes reading different
  // intervals hit different banks
  {
    const int4* src = reinterpret_cast<const int4*>(args.tables + (size_t)cand * args.table_bytes);
    int4* dst = reinterpret_cast<int4*>(dyn);
    for (int q = tid; q < args.table_bytes / 16; q += BLOCK) dst[q] = src[q];
  }
  const SplHead* HD = reinterpret_cast<const SplHead*>(dyn);
  const SplInterval* IV = reinterpret_cast<const SplInterval*>(dyn + sizeof(SplHead) * args.pot.nspl);
  double* gt = reinterpret_cast<double*>(dyn + args.table_bytes);
  if (MEAM) {
    const SplInterval* src = reinterpret_cast<const SplInterval*>(args.tables + (size_t)cand * args.table_bytes +
                                                                  sizeof(SplHead) * args.pot.nspl);
    for (int q = tid; q < args.pot.ngco * 6; q += BLOCK) gt[q] = src[args.pot.gco0 + q / 6].c[q % 6];
  }
  for (int a = tid; a <= ch.na; a += BLOCK) sm.poff[a] = args.atom_poff[ch.a0 + a] - ch.p0;
  for (int a = tid; a < ch.na; a += BLOCK) {
    sm.rho[a] = 0.0;
    sm.fi[a][0] = sm.fi[a][1] = sm.fi[a][2] = 0.0;
  }
  __syncthreads();

  const int nt = args.pot.nt;
  const double* box = args.box + 9 * ch.cell;

  // per-lane pair state
  int ia[PPL], jg[PPL];
  double r[PPL], invr[PPL], ux[PPL], uy[PPL], uz[PPL];
  double aphi[PPL], brho[PPL], fv[PPL], dfv[PPL], Gx[PPL], Gy[PPL], Gz[PPL];
  double e_acc = 0.0;

  // ---- phase 1: geometry and radial functions
#pragma unroll
  for (int s = 0; s < PPL; ++s) {
    const int p = tid + s * BLOCK;
    ia[s] = -1;
    Gx[s] = Gy[s] = Gz[s] = 0.0;
    if (p < ch.np) {
      int lo = 0, hi = ch.na - 1;
      while (lo < hi) {  // last atom with poff <= p
        const int mid = (lo + hi + 1) >> 1;
        if (sm.poff[mid] <= p) lo = mid; else hi = mid - 1;
      }
      ia[s] = lo;
      const int ig = ch.a0 + lo;
      const uint2 pr = args.pairs[ch.p0 + p];
      const int j = (int)pr.x;
      jg[s] = j;
      const int ix = (int)(int8_t)(pr.y & 0xff), iy = (int)(int8_t)((pr.y >> 8) & 0xff),
                iz = (int)(int8_t)((pr.y >> 16) & 0xff);
      const double* pj = args.pos + 3 * (size_t)j;
      const double* pi = args.pos + 3 * (size_t)ig;
      double d[3];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        // Config::find_pairs src/config.cpp:195-197, evaluated left to right
        double t = __dadd_rn(pj[m], __dmul_rn((double)ix, box[m]));
        t = __dadd_rn(t, __dmul_rn((double)iy, box[3 + m]));
        t = __dadd_rn(t, __dmul_rn((double)iz, box[6 + m]));
        d[m] = __dadd_rn(t, -pi[m]);
      }
      const double rr = __dsqrt_rn(__dadd_rn(__dadd_rn(__dmul_rn(d[0], d[0]), __dmul_rn(d[1], d[1])),
                                             __dmul_rn(d[2], d[2])));
      r[s] = rr;
      invr[s] = 1.0 / rr;
      ux[s] = d[0] / rr;
      uy[s] = d[1] / rr;
      uz[s] = d[2] / rr;
      const int ti = args.types[ig], tj = args.types[j];
      const int pij = pair_ij(ti, tj, nt);
      double v, dv;
      const SplHead& Hphi = HD[args.pot.fn_off[0] + pij];
      aphi[s] = 0.0;
      if (eval_radial<POINT, Q>(Hphi, IV, rr, v, dv)) {
        e_acc += 0.5 * v;
        aphi[s] = 0.5 * dv;
      }
      const SplHead& Hrho = HD[args.pot.fn_off[1] + tj];
      brho[s] = 0.0;
      if (eval_radial<POINT, Q>(Hrho, IV, rr, v, dv)) {
        atomicAdd(&sm.rho[lo], v);
        brho[s] = dv;
      }
      fv[s] = 0.0;
      dfv[s] = 0.0;
      if (MEAM) {
        const SplHead& Hf = HD[args.pot.fn_off[3] + pij];
        if (eval_radial<POINT, Q>(Hf, IV, rr, v, dv)) {
          fv[s] = v;
          dfv[s] = dv;
        }
        double2* K = reinterpret_cast<double2*>(sm.kd[p]);
        K[0] = make_double2(ux[s], uy[s]);
        K[1] = make_double2(uz[s], fv[s]);
        K[2] = make_double2(dfv[s], invr[s]);
        sm.gl[0][p] = sm.gl[1][p] = sm.gl[2][p] = 0.0;
      }
    }
  }
  __syncthreads();

  // ---- phase 2: triplets, each unordered (j, k) once (dF factored out)
  // Lane j of an atom with n pairs visits partners k = j+1 .. j+(n-1)/2
  // (mod n), plus k = j+n/2 for the first n/2 lanes when n is even: every
  // lane gets the same number of partners.  The own leg accumulates in
  // registers, the partner leg is pushed into LDS (distinct k per lane and
  // step, so the atomics never collide).
  if (MEAM) {
#pragma unroll
    for (int s = 0; s < PPL; ++s) {
      const int p = tid + s * BLOCK;
      if (ia[s] < 0) continue;
      const int kb = sm.poff[ia[s]], ke = sm.poff[ia[s] + 1];
      const int n = ke - kb, l = p - kb;
      const int trips = ((n - 1) >> 1) + (((n & 1) == 0 && l < (n >> 1)) ? 1 : 0);
      const int ti = args.types[ch.a0 + ia[s]];
      const SplHead& Hg = HD[args.pot.fn_off[4] + ti];
      const double gx0 = Hg.x0, ginv = Hg.invstep, gxcut = Hg.xcut, gxn = Hg.xn;
      const double gkmax = (double)(Hg.n - 2);
      const double gy0 = Hg.y0, gyn = Hg.yn, ggL = Hg.gL, ggR = Hg.gR, gg2L = Hg.g2L, gg2R = Hg.g2R;
      const double* gtg = gt + (size_t)(Hg.coef_off - args.pot.gco0) * 6;
      const double fj = fv[s], dfj = dfv[s], ivj = invr[s];
      const double xj = ux[s], yj = uy[s], zj = uz[s];
      double rho_t = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
      int kl = l;
      for (int m = 0; m < trips; ++m) {
        kl = (kl + 1 == n) ? 0 : kl + 1;
        const int k = kb + kl;
        const double2* K = reinterpret_cast<const double2*>(sm.kd[k]);
        const double2 ka = K[0], kbv = K[1], kc = K[2];
        const double xk = ka.x, yk = ka.y, zk = kbv.x;
        const double cs = xj * xk + yj * yk + zj * zk;
        double gv, dgt;  // g and dg/dt (t = shift within the interval)
        bool keep = true;
        if (POINT && (cs < gx0 || cs > gxn)) {
          const bool left = cs < gx0;
          const double e = cs - (left ? gx0 : gxn);
          const double g1 = left ? ggL : ggR, g2 = left ? gg2L : gg2R;
          gv = (left ? gy0 : gyn) + g1 * e + 0.5 * g2 * e * e;
          dgt = (g1 + g2 * e) / ginv;
        } else {
          const double u = (cs - gx0) * ginv;
          const double fl = fmin(fmax(floor(u), 0.0), gkmax);
          const double t = u - fl;
          const double2* C = reinterpret_cast<const double2*>(gtg + 6 * (int)fl);
          const double2 c01 = C[0], c23 = C[1];
          if (Q) {
            const double2 c45 = C[2];
            gv = c01.x + t * (c01.y + t * (c23.x + t * (c23.y + t * (c45.x + t * c45.y))));
            dgt = c01.y + t * (2.0 * c23.x + t * (3.0 * c23.y + t * (4.0 * c45.x + t * 5.0 * c45.y)));
          } else {
            gv = c01.x + t * (c01.y + t * (c23.x + t * c23.y));
            dgt = c01.y + t * (2.0 * c23.x + t * 3.0 * c23.y);
          }
          if (!POINT) keep = cs < gxcut;
        }
        const double fk = keep ? kbv.y : 0.0;
        const double dfk = keep ? kc.x : 0.0;
        const double ivk = kc.y;
        const double ff = fj * fk;
        rho_t += ff * gv;
        const double V3 = dgt * ginv * ff;          // V3 / dF
        const double vlj = V3 * ivj, vlk = V3 * ivk;
        const double vvj = gv * dfj * fk - vlj * cs;  // (dV3j - vlj cos) / dF
        const double vvk = gv * fj * dfk - vlk * cs;  // (dV3k - vlk cos) / dF
        gx += xj * vvj + xk * vlj;
        gy += yj * vvj + yk * vlj;
        gz += zj * vvj + zk * vlj;
        atomicAdd(&sm.gl[0][k], xk * vvk + xj * vlk);
        atomicAdd(&sm.gl[1][k], yk * vvk + yj * vlk);
        atomicAdd(&sm.gl[2][k], zk * vvk + zj * vlk);
      }
      Gx[s] = gx;
      Gy[s] = gy;
      Gz[s] = gz;
      atomicAdd(&sm.rho[ia[s]], rho_t);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < PPL; ++s) {
      const int p = tid + s * BLOCK;
      if (ia[s] < 0) continue;
      Gx[s] += sm.gl[0][p];
      Gy[s] += sm.gl[1][p];
      Gz[s] += sm.gl[2][p];
    }
  }

  // ---- phase 3: embedding
  double c_acc = 0.0;
  const double wcell = args.weight[ch.cell];
  for (int a = tid; a < ch.na; a += BLOCK) {
    double rho = sm.rho[a];
    if (args.rho_out) args.rho_out[(size_t)cand * args.natoms + ch.a0 + a] = rho;
    const SplHead& HU = HD[args.pot.fn_off[2] + args.types[ch.a0 + a]];
    const double lo = HU.x0, hi = HU.xn;
    if (rho < lo) {
      c_acc += wcell * kDummyWeight * 10. * (lo - rho) * (lo - rho);
      if (!args.pot.embed_extrap) rho = lo;
    } else if (rho > hi) {
      c_acc += wcell * kDummyWeight * 10. * (rho - hi) * (rho - hi);
      if (!args.pot.embed_extrap) rho = hi;
    }
    double dU, d2U;
    e_acc += eval_point<Q>(HU, IV, rho, dU, d2U);
    sm.dF[a] = dU;
    if (args.dF_out) args.dF_out[(size_t)cand * args.natoms + ch.a0 + a] = dU;
  }
  __syncthreads();

  // ---- phase 4: pair forces, three-body forces, virial
  double sxx = 0.0, syy = 0.0, szz = 0.0, sxy = 0.0, syz = 0.0, szx = 0.0;
#pragma unroll
  for (int s = 0; s < PPL; ++s) {
    if (ia[s] < 0) continue;
    const double dFi = sm.dF[ia[s]];
    const double sc = aphi[s] + brho[s] * dFi;
    const double tx = ux[s] * sc + dFi * Gx[s];
    const double ty = uy[s] * sc + dFi * Gy[s];
    const double tz = uz[s] * sc + dFi * Gz[s];
    if (args.G_out) {
      double* g = args.G_out + ((size_t)cand * args.npairs + ch.p0 + tid + s * BLOCK) * 3;
      g[0] = Gx[s];
      g[1] = Gy[s];
      g[2] = Gz[s];
    }
    double* tv = args.tvec + ((size_t)cand * args.npairs + ch.p0 + tid + s * BLOCK) * 3;
    tv[0] = tx;
    tv[1] = ty;
    tv[2] = tz;
    atomicAdd(&sm.fi[ia[s]][0], tx);
    atomicAdd(&sm.fi[ia[s]][1], ty);
    atomicAdd(&sm.fi[ia[s]][2], tz);
    const double rx = tx * r[s], ry = ty * r[s], rz = tz * r[s];
    sxx -= ux[s] * rx;
    syy -= uy[s] * ry;
    szz -= uz[s] * rz;
    sxy -= ux[s] * ry;
    syz -= uy[s] * rz;
    szx -= uz[s] * rx;
  }
  __syncthreads();
  for (int a = tid; a < ch.na; a += BLOCK) {
    double* fa = args.fown + ((size_t)cand * args.natoms + ch.a0 + a) * 3;
    fa[0] = sm.fi[a][0];
    fa[1] = sm.fi[a][1];
    fa[2] = sm.fi[a][2];
  }
  // chunk partials of energy, virial and penalty: one wave-level reduction
  // per value, then the per-wave sums in fixed order
  const double vals[8] = {e_acc, sxx, syy, szz, sxy, syz, szx, c_acc};
  const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double v = wave_sum(vals[q]);
    if (lane == 0) sm.red8[q][wid] = v;
  }
  __syncthreads();
  if (tid < 8) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) t += sm.red8[tid][w];
    args.chunk_acc[((size_t)cand * args.nchunks + blockIdx.x) * 8 + tid] = t;
  }
}

// ------------------------------------------------------------------ k_resid

// Eight lanes per item; items are the atoms, then the cells.  An atom's lanes
// gather the force vectors of the pairs pointing at it (F_a = own sum minus
// incoming sum, fixed order, no atomics) and form its three force residuals;
// a cell's lanes each reduce one of the eight per-chunk accumulators (energy,
// six virial components, density constraint) over the cell's chunks.
// Residuals follow Potential::accumulate_error (src/potential.cpp:415-491).
constexpr int kResidLanes = 8;

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_resid(const ResidArgs a) {
  __shared__ double red[BLOCK / 64];
  const int cand = blockIdx.y;
  const int item = (blockIdx.x * BLOCK + threadIdx.x) / kResidLanes;
  const int l = threadIdx.x & (kResidLanes - 1);
  double sq = 0.0;
  double* ev = a.errvec ? a.errvec + (size_t)cand * a.nerr : nullptr;
  if (item < a.natoms) {
    const int i = item;
    const double* tv = a.tvec + (size_t)cand * a.npairs * 3;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    const int b = a.in_off[i], e = a.in_off[i + 1];
    for (int q = b + l; q < e; q += kResidLanes) {
      const double* t = tv + 3 * (size_t)a.in_list[q];
      sx += t[0];
      sy += t[1];
      sz += t[2];
    }
#pragma unroll
    for (int o = kResidLanes / 2; o > 0; o >>= 1) {
      sx += __shfl_xor(sx, o, kResidLanes);
      sy += __shfl_xor(sy, o, kResidLanes);
      sz += __shfl_xor(sz, o, kResidLanes);
    }
    if (l < 3) {
      const int c = a.cell_of[i];
      const double w = a.weight[c];
      const size_t ai = ((size_t)cand * a.natoms + i) * 3 + l;
      const double F = a.fown[ai] - (l == 0 ? sx : (l == 1 ? sy : sz));
      const double er = w * (F - a.force0[3 * (size_t)i + l]);
      sq = er * er;
      a.forces[ai] = F;
      if (ev) ev[7 * (size_t)(c + 1) + 3 * (size_t)i + l] = er;
      if (a.lam) a.lam[ai] = 2.0 * er * w;
    }
  } else if (item < a.natoms + a.ncells) {
    const int c = item - a.natoms;
    const double* cacc = a.chunk_acc + (size_t)cand * a.nchunks * 8;
    double acc = 0.0;
    for (int q = a.cell_chunk_off[c]; q < a.cell_chunk_off[c + 1]; ++q) acc += cacc[8 * (size_t)a.cell_chunks[q] + l];
    const int n = a.natoms_c[c];
    const size_t e0 = 7 * (size_t)c + 3 * (size_t)a.abase[c];
    double* adj = a.cell_adj ? a.cell_adj + ((size_t)cand * a.ncells + c) * 8 : nullptr;
    double er, dadj;
    if (l == 0) {
      const double E = acc / n;
      const double WE = a.weight[c] * a.eweight[c] * a.energy_weight;
      er = WE * n * (E - a.energy0[c]);
      dadj = 2.0 * er * WE;
      if (ev) ev[e0] = er;
      if (a.cell_energy) a.cell_energy[(size_t)cand * a.ncells + c] = E;
    } else if (l < 7) {
      const double S = acc / a.vol[c];
      const double WS = a.weight[c] * a.sweight[c] * a.stress_weight;
      er = WS * (S - a.stress0[6 * c + l - 1]);
      dadj = 2.0 * er * WS / a.vol[c];
      if (ev) ev[e0 + l] = er;
      if (a.cell_stress) a.cell_stress[((size_t)cand * a.ncells + c) * 6 + l - 1] = S;
    } else {
      er = acc;
      dadj = 2.0 * er;
      if (ev) ev[7 * (size_t)a.ncells + 3 * (size_t)a.natoms + c] = er;
    }
    sq = er * er;
    if (adj) adj[l] = dadj;
  }
  const double t = block_sum<BLOCK>(sq, red);
  if (threadIdx.x == 0) a.partial[(size_t)cand * gridDim.x + blockIdx.x] = t;
}

// ------------------------------------------------------------------ k_final

__global__ void __launch_bounds__(256) k_final(const FinalArgs a) {
  __shared__ double red[4];
  const int cand = blockIdx.x;
  double s = 0.0;
  for (int q = threadIdx.x; q < a.npartial; q += 256) s += a.partial[(size_t)cand * a.npartial + q];
  const double tot = block_sum<256>(s, red);
  if (threadIdx.x != 0) return;
  double F = tot;
  double* ev = a.errvec ? a.errvec + (size_t)cand * a.nerr : nullptr;
  const size_t tail = 8 * (size_t)a.ncells + 3 * (size_t)a.natoms;
  const double* kv = a.knots + (size_t)cand * a.nknot;
  if (a.meam) {
    // PotFns::get_max_y_mag over the f splines (src/pot_fns.cpp:100-111)
    double mx = 0.0;
    for (int q = 0; q < a.nf; ++q) {
      const int s2 = a.f0 + q;
      const double* y = kv + a.koff[s2];
      double m = 0.0;
      for (int k = 0; k < a.nknots[s2]; ++k)
        if (fabs(m) < fabs(y[k])) m = y[k];
      if (fabs(mx) < fabs(m)) mx = m;
    }
    const double m = fabs(mx);
    const double pe = kDummyWeight * 25. * (1.0 - m) * (1.0 - m);
    F += pe * pe;
    if (ev) ev[tail] = pe;
  } else {
    const char* tb = a.tables + (size_t)cand * a.table_bytes;
    const SplHead* HD = reinterpret_cast<const SplHead*>(tb);
    const SplInterval* IV = reinterpret_cast<const SplInterval*>(tb + sizeof(SplHead) * a.nspl);
    for (int t = 0; t < a.nt; ++t) {
      const SplHead& H = HD[a.u0 + t];
      double d, d2;
      if (a.quintic) eval_point<true>(H, IV, 0.5 * (H.x0 + H.xn), d, d2);
      else eval_point<false>(H, IV, 0.5 * (H.x0 + H.xn), d, d2);
      const double pe = kDummyWeight * d;
      F += pe * pe;
      if (ev) ev[tail + t] = pe;
    }
  }
  a.F[cand] = F;
}

// ---------------------------------------------------------------- launchers

void launch_refresh(hipStream_t st, const double* knots, int nknot, const RefreshMeta& meta, char* tables,
                    int table_bytes, int ncand) {
  hipLaunchKernelGGL(k_refresh, dim3(meta.nspl, ncand), dim3(64), 0, st, knots, nknot, meta, tables, table_bytes);
}

void launch_zero(hipStream_t st, double* p, size_t n) {
  if (n == 0) return;
  size_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
}

template <int BLOCK, int PPL, int MAXA>
static void launch_atoms_t(hipStream_t st, const AtomArgs& a, int nchunks, int ncand) {
  const int style = a.pot.style;
  const bool meam = style == kMeamSpline || style == kMeam;
  const bool point = style == kEam || style == kMeam;
  const bool q = a.pot.quintic != 0;
  const size_t dyn = (size_t)a.table_bytes + (size_t)a.pot.ngco * 48;
  dim3 grid(nchunks, ncand), block(BLOCK);
#define MZ_LAUNCH(M, P, Q_)                                                                          \
  hipLaunchKernelGGL((k_atoms<BLOCK, PPL, MAXA, M, P, Q_>), grid, block, dyn, st, a)
  if (meam) {
    if (point) { if (q) MZ_LAUNCH(true, true, true); else MZ_LAUNCH(true, true, false); }
    else { if (q) MZ_LAUNCH(true, false, true); else MZ_LAUNCH(true, false, false); }
  } else {
    if (point) { if (q) MZ_LAUNCH(false, true, true); else MZ_LAUNCH(false, true, false); }
    else { if (q) MZ_LAUNCH(false, false, true); else MZ_LAUNCH(false, false, false); }
  }
#undef MZ_LAUNCH
}

int launch_atoms(hipStream_t st, const AtomArgs& a, int nchunks, int ncand, int block, int ppl) {
  if (block == 256 && ppl == 2) launch_atoms_t<256, 2, 64>(st, a, nchunks, ncand);
  else if (block == 256 && ppl == 1) launch_atoms_t<256, 1, 64>(st, a, nchunks, ncand);
  else if (block == 128 && ppl == 2) launch_atoms_t<128, 2, 64>(st, a, nchunks, ncand);
  else if (block == 128 && ppl == 4) launch_atoms_t<128, 4, 64>(st, a, nchunks, ncand);
  else if (block == 512 && ppl == 1) launch_atoms_t<512, 1, 64>(st, a, nchunks, ncand);
  else if (block == 64 && ppl == 4) launch_atoms_t<64, 4, 64>(st, a, nchunks, ncand);
  else return -1;
  return 0;
}

int resid_blocks(int natoms, int ncells) {
  return (int)(((long long)(natoms + ncells) * kResidLanes + 255) / 256);
}

void launch_resid(hipStream_t st, const ResidArgs& a, int nblocks, int ncand) {
  hipLaunchKernelGGL(k_resid<256>, dim3(nblocks, ncand), dim3(256), 0, st, a);
}

void launch_final(hipStream_t st, const FinalArgs& a, int ncand) {
  hipLaunchKernelGGL(k_final, dim3(ncand), dim3(256), 0, st, a);
}

}  // namespace mz
