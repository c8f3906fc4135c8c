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

#include <type_traits>

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

// Location of r on a radial spline, shared by every candidate of a workgroup
// (the knots of all candidates sit at the same x).  reg: 0 interval k with
// shift t, 1 left of x_0 (point style, t = r - x_0).  Returns false when the
// function is cut off: knot style beyond x_n (1 + 1e-8) (src/spline.cpp:221-246),
// point style beyond x_n.
template <bool POINT>
__device__ __forceinline__ bool locate_radial(const SplHead& H, const SplInterval* IV, double r, int& k, double& t,
                                              int& reg) {
  reg = 0;
  if (POINT) {
    if (!(r <= H.xn)) return false;
    if (r < H.x0) {
      reg = 1;
      t = r - H.x0;
      k = 0;
      return true;
    }
    const double delta = r - H.x0;
    k = (int)__dmul_rn(delta, H.invstep);
    if (k >= H.n - 1) k = H.n - 2;
    if (k < 0) k = 0;
    t = __dmul_rn(__dadd_rn(delta, -__dmul_rn((double)k, H.step)), H.invstep);
    return true;
  }
  return locate_knot(H, IV + H.coef_off, r, k, t);
}

// Value and d/dr of one candidate's spline at a located r.
template <bool Q>
__device__ __forceinline__ void eval_located(const SplHead& H, const SplInterval* IV, int k, double t, int reg,
                                             double& v, double& d) {
  if (reg == 1) {
    d = H.gL + H.g2L * t;
    v = H.y0 + H.gL * t + 0.5 * H.g2L * t * t;
    return;
  }
  poly<Q>(IV[H.coef_off + k], t, H.invstep, v, d);
}

// Shared part of a g evaluation: interval and shift of cos (knot positions are
// common to all candidates), or the extrapolation side (point style).
struct GLoc {
  int off;    // interval (record) index in the spline's compact g records
  int side;   // 0 inside, 1 left of x_0, 2 right of x_n (point style)
  double t;   // shift in the interval, or cos - end when extrapolating
};

template <bool POINT, bool Q>
__device__ __forceinline__ GLoc g_locate(const SplHead& Hg, double cs) {
  GLoc L;
  if (POINT && (cs < Hg.x0 || cs > Hg.xn)) {
    L.side = cs < Hg.x0 ? 1 : 2;
    L.t = cs - (L.side == 1 ? Hg.x0 : Hg.xn);
    L.off = 0;
    return L;
  }
  const double u = (cs - Hg.x0) * Hg.invstep;
  const double fl = fmin(fmax(floor(u), 0.0), (double)(Hg.n - 2));
  L.t = u - fl;
  L.off = (int)fl;
  L.side = 0;
  return L;
}

// g and dg/dcos of one candidate at a located cosine.
template <bool POINT, bool Q>
__device__ __forceinline__ void g_value(const SplHead& Hg, const double2* gtg, int ps, const GLoc& L, double& gv,
                                        double& gp) {
  if (POINT && L.side != 0) {
    const bool left = L.side == 1;
    const double e = L.t;
    const double g1 = left ? Hg.gL : Hg.gR, g2 = left ? Hg.g2L : Hg.g2R;
    gv = (left ? Hg.y0 : Hg.yn) + g1 * e + 0.5 * g2 * e * e;
    gp = g1 + g2 * e;
    return;
  }
  const double2* C = gtg + L.off;
  const double2 c01 = C[0], c23 = C[ps], c45 = Q ? C[2 * ps] : c23;
  double dgt;
  horner_d<Q>(c01, c23, c45, L.t, gv, dgt);
  gp = dgt * Hg.invstep;
}

// LDS layout of one workgroup (static part).  CT candidates share the pair
// geometry; f of candidate 0 rides in the partner record, the others in fx.
template <int BLOCK, int PPL, int MAXA, int CT, bool GP, int JM>
struct AtomSmem {
  double4 rec[BLOCK * PPL];                 // per pair: unit vector, f (candidate 0); later T
  double fx[CT > 1 ? CT - 1 : 1][BLOCK * PPL];  // f of candidates 1..CT-1
  double rp[CT][BLOCK * PPL];               // per pair: its density contribution to the owning atom
  double rho[CT][MAXA];
  double dF[CT][MAXA];
  int poff[MAXA + 1];
  double red8[CT][8][BLOCK / 64];
  int ptype[GP ? BLOCK * PPL : 1];          // per pair: type of its neighbour (g by pair types)
  double px[JM > 0 ? JM * CT : 1][JM > 0 ? BLOCK * PPL : 1];  // jmeam: p (p2) of every pair, per candidate
};

// One workgroup per (chunk, group of CT candidates).  TG: the spline tables
// are read from global memory instead of being staged in LDS (large tables).
// GP: the g spline of a triplet depends on the types of both neighbours
// (gmeam styles, src/pot_fns.cpp:154-171 Triplet_i_jk), chosen per triplet.
// JM: jmeam styles, JM extra energy terms p_ij p_ik q(cos) per triplet
// (src/pots/pot_jmeam/pot_jmeam.cpp:221,263-265; jmeam2 adds p2 p2 q2,
// src/pots/pot_jmeam2/pot_jmeam2.cpp:240-241,284-286), walked beside g.
// The 256-thread, one-pair-per-lane, four-candidate tile is held to three
// waves per SIMD (168 VGPRs, no spill; 169 would leave two): its 28 KB of LDS
// then fit three workgroups per CU.
template <int BLOCK, int PPL, int MAXA, int CT, bool MEAM, bool POINT, bool Q, bool TG, bool GP, int JM>
__global__ void __launch_bounds__(BLOCK)
    __attribute__((amdgpu_waves_per_eu((CT == 4 && BLOCK == 256 && PPL == 1) ? 3 : 1))) k_atoms(const AtomArgs args) {
  extern __shared__ __attribute__((aligned(64))) char dyn[];
  __shared__ __attribute__((aligned(16))) AtomSmem<BLOCK, PPL, MAXA, CT, GP, JM> sm;
  constexpr int NJ = JM > 0 ? JM : 1;
  constexpr int NG = g_record_doubles(Q);
  const Chunk ch = args.chunks[blockIdx.x];
  const int tid = threadIdx.x;
  const int TB = args.table_bytes;
  int cand[CT];
  bool valid[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const int cc = blockIdx.y * CT + c;
    valid[c] = cc < args.ncand;
    cand[c] = valid[c] ? cc : args.ncand - 1;
  }

  // LDS per candidate: its spline table (with TG only the heads; the
  // interval records are then read from global memory, cache-resident, per
  // pair and per atom), then the g intervals again as compact power-of-two
  // records (c0..c3, or c0..c5 padded to 8 doubles), which the triplet walk
  // reads per triplet.
  const int gbytes = MEAM ? args.pot.ngco * NG * 8 : 0;
  const int ghb = TG ? args.pot.nspl * (int)sizeof(SplHead) : 0;
  const int tbytes = TG ? 0 : TB;
  const int cbytes = tbytes + ghb + gbytes;
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const char* tb = args.tables + (size_t)cand[c] * TB;
    char* cb = dyn + (size_t)c * cbytes;
    if (!TG) {
      const int4* src = reinterpret_cast<const int4*>(tb);
      for (int q = tid; q < TB / 16; q += BLOCK) reinterpret_cast<int4*>(cb)[q] = src[q];
    }
    if (TG) {
      const int4* hsrc = reinterpret_cast<const int4*>(tb);
      for (int q = tid; q < ghb / 16; q += BLOCK) reinterpret_cast<int4*>(cb)[q] = hsrc[q];
    }
    if (MEAM) {
      const SplInterval* gsrc = reinterpret_cast<const SplInterval*>(tb + sizeof(SplHead) * args.pot.nspl);
      double2* gdst = reinterpret_cast<double2*>(cb + tbytes + ghb);
      const int ng = args.pot.ngco;
      for (int q = tid; q < ng * (NG / 2); q += BLOCK) {
        const int m = q / ng;
        const SplInterval& I = gsrc[args.pot.gco0 + q - m * ng];
        gdst[q] = make_double2(I.c[2 * m], I.c[2 * m + 1]);
      }
    }
  }
  auto HDc = [&](int c) { return reinterpret_cast<const SplHead*>(dyn + (size_t)c * cbytes); };
  auto IVc = [&](int c) {
    return reinterpret_cast<const SplInterval*>((TG ? args.tables + (size_t)cand[c] * TB : dyn + (size_t)c * cbytes) +
                                                sizeof(SplHead) * args.pot.nspl);
  };
  auto GHc = [&](int c) { return HDc(c) + args.pot.fn_off[4]; };  // g heads by element type (type, pair of types with GP)
  auto GTc = [&](int c) { return reinterpret_cast<const double2*>(dyn + (size_t)c * cbytes + tbytes + ghb); };
  const int ngc = args.pot.ngco;  // plane stride of the compact g records
  for (int a = tid; a <= ch.na; a += BLOCK) sm.poff[a] = args.atom_poff[ch.a0 + a] - ch.p0;
  __syncthreads();

  const int nt = args.pot.nt;
  const double* box = args.box + 9 * ch.cell;
  const SplHead* HD0 = HDc(0);
  // knots sit at the same x for all candidates: candidate 0's records locate r
  // (staging a shared copy of them in LDS per workgroup measured slower)
  const SplInterval* IV0 = IVc(0);

  // per-lane pair state; ipos: the pair's slot in its neighbour's incoming
  // run, read here with the pair record so that phase 4's scattered force
  // store does not wait on it
  int ia[PPL], tjs[PPL], ipos[PPL];
  double r[PPL], invr[PPL], ux[PPL], uy[PPL], uz[PPL];
  double aphi[PPL][CT], brho[PPL][CT], fv[PPL][CT], dfv[PPL][CT], Gx[PPL][CT], Gy[PPL][CT], Gz[PPL][CT],
      S1v[PPL][CT];
  // jmeam: p and p' of the lane's pairs, and the legs H of the p p q terms
  double pv[NJ][PPL][CT], dpv[NJ][PPL][CT], Hx[PPL][CT], Hy[PPL][CT], Hz[PPL][CT];
  double e_acc[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) e_acc[c] = 0.0;

  // ---- phase 1: geometry and radial functions
#pragma unroll
  for (int s = 0; s < PPL; ++s) {
    const int p = tid + s * BLOCK;
    ia[s] = -1;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      aphi[s][c] = brho[s][c] = fv[s][c] = dfv[s][c] = 0.0;
      Gx[s][c] = Gy[s][c] = Gz[s][c] = S1v[s][c] = 0.0;
      Hx[s][c] = Hy[s][c] = Hz[s][c] = 0.0;
#pragma unroll
      for (int m = 0; m < NJ; ++m) pv[m][s][c] = dpv[m][s][c] = 0.0;
    }
    ipos[s] = 0;
    if (p < ch.np) {
      ipos[s] = args.in_pos[ch.p0 + p];
      int lo = 0, hi = ch.na - 1;
      while (lo < hi) {  // last atom with poff <= p
        const int mid = (lo + hi + 1) >> 1;
        if (sm.poff[mid] <= p) lo = mid; else hi = mid - 1;
      }
      ia[s] = lo;
      const int ig = ch.a0 + lo;
      const uint2 pr = args.pairs[ch.p0 + p];
      const int j = (int)pr.x;
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
      tjs[s] = tj;
      if (GP) sm.ptype[p] = tj;
      const int pij = pair_ij(ti, tj, nt);
      int k, reg;
      double t;
      const int sphi = args.pot.fn_off[0] + pij, srho = args.pot.fn_off[1] + tj;
      if (locate_radial<POINT>(HD0[sphi], IV0, rr, k, t, reg)) {
#pragma unroll
        for (int c = 0; c < CT; ++c) {
          double v, dv;
          eval_located<Q>(HDc(c)[sphi], IVc(c), k, t, reg, v, dv);
          e_acc[c] += 0.5 * v;
          aphi[s][c] = 0.5 * dv;
        }
      }
      double rv[CT];
#pragma unroll
      for (int c = 0; c < CT; ++c) rv[c] = 0.0;
      if (!args.pot.nodens && locate_radial<POINT>(HD0[srho], IV0, rr, k, t, reg)) {
#pragma unroll
        for (int c = 0; c < CT; ++c) eval_located<Q>(HDc(c)[srho], IVc(c), k, t, reg, rv[c], brho[s][c]);
      }
#pragma unroll
      for (int c = 0; c < CT; ++c) sm.rp[c][p] = rv[c];
      if (MEAM) {
        const int sf = args.pot.fn_off[3] + pij;
        if (locate_radial<POINT>(HD0[sf], IV0, rr, k, t, reg)) {
#pragma unroll
          for (int c = 0; c < CT; ++c) eval_located<Q>(HDc(c)[sf], IVc(c), k, t, reg, fv[s][c], dfv[s][c]);
        }
        sm.rec[p] = make_double4(ux[s], uy[s], uz[s], fv[s][0]);
#pragma unroll
        for (int c = 1; c < CT; ++c) sm.fx[c - 1][p] = fv[s][c];
#pragma unroll
        for (int m = 0; m < JM; ++m) {
          const int sp = args.pot.fn_off[5 + 2 * m] + pij;
          if (locate_radial<POINT>(HD0[sp], IV0, rr, k, t, reg)) {
#pragma unroll
            for (int c = 0; c < CT; ++c) eval_located<Q>(HDc(c)[sp], IVc(c), k, t, reg, pv[m][s][c], dpv[m][s][c]);
          }
#pragma unroll
          for (int c = 0; c < CT; ++c) sm.px[m * CT + c][p] = pv[m][s][c];
        }
      }
    }
  }
  __syncthreads();

  // ---- phase 2: triplets (dF factored out)
  // With g = g(cos), g' = dg/dcos, every pair j needs, over the other pairs k
  // of its atom, S1 = sum g f_k, S2 = sum g' f_k cos, S3 = sum g' f_k d_k; its
  // three-body leg is then G_j = d_j (f'_j S1 - f_j S2 / r_j) + (f_j / r_j) S3
  // and the atom's density gains f_j S1 / 2 (each unordered triplet is seen
  // from both of its pairs).  All lanes of an atom walk its pairs k in the same
  // order, so partner records are LDS broadcasts and nothing is written until
  // the walk ends; cos and the g interval are shared by the CT candidates.
  if (MEAM) {
    const bool cut = !POINT && !args.pot.g_keep_all;
#pragma unroll
    for (int s = 0; s < PPL; ++s) {
      const int p = tid + s * BLOCK;
      if (ia[s] < 0) continue;
      const int kb = sm.poff[ia[s]], ke = sm.poff[ia[s] + 1];
      const int ti = args.types[ch.a0 + ia[s]];
      const int gfix = GP ? ti * (nt * (nt + 1) / 2) : ti;  // g of the lane's atom (and, with GP, first pair)
      const SplHead& Hg0 = GHc(0)[gfix];
      const int goff0 = Hg0.coef_off - args.pot.gco0;
      const double xj = ux[s], yj = uy[s], zj = uz[s];
      double S[CT][5], T[NJ][CT][5];
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int q = 0; q < 5; ++q) {
          S[c][q] = 0.0;
#pragma unroll
          for (int m = 0; m < NJ; ++m) T[m][c][q] = 0.0;
        }
      const double gx0 = Hg0.x0, ginv = Hg0.invstep;
      const int gkm = Hg0.n - 2;
      auto walk = [&](auto spn) {
      for (int k = kb; k < ke; ++k) {
        const double4 P = sm.rec[k];
        const double cs = xj * P.x + yj * P.y + zj * P.z;
        const int gi = GP ? gfix + pair_ij(tjs[s], sm.ptype[k], nt) : gfix;
        const SplHead& Hg = GP ? GHc(0)[gi] : Hg0;
        const int goff = GP ? Hg.coef_off - args.pot.gco0 : goff0;
        GLoc L;
        bool zero;
        if constexpr (decltype(spn)::value) {
          // the spline spans [-1, 1]: u > -1, truncation gives the clamped floor
          const double u = (cs - gx0) * ginv;
          const int kk = (int)u;
          L.off = kk < gkm ? kk : gkm;
          L.t = u - (double)L.off;
          L.side = 0;
          zero = k == p;
        } else {
          L = g_locate<POINT, Q>(Hg, cs);
          zero = k == p || (cut && !(cs < Hg.xcut));
        }
#pragma unroll
        for (int c = 0; c < CT; ++c) {
          double gv, gp;
          g_value<POINT, Q>(GHc(c)[gi], GTc(c) + goff, ngc, L, gv, gp);
          const double fk = zero ? 0.0 : (c == 0 ? P.w : sm.fx[c > 0 ? c - 1 : 0][k]);
          const double b = gp * fk;
          S[c][0] += gv * fk;
          S[c][1] += b * cs;
          S[c][2] += b * P.x;
          S[c][3] += b * P.y;
          S[c][4] += b * P.z;
        }
        // jmeam: q (q2) of the atom's type at the same cosine; the term is
        // absent for a knot-style cosine beyond the spline
#pragma unroll
        for (int m = 0; m < JM; ++m) {
          const int iq = args.pot.fn_off[6 + 2 * m] + ti;
          const SplHead& Hq = HDc(0)[iq];
          const GLoc Lq = g_locate<POINT, Q>(Hq, cs);
          const bool zq = k == p || (!POINT && !(cs < Hq.xcut));
          const int qoff = Hq.coef_off - args.pot.gco0;
#pragma unroll
          for (int c = 0; c < CT; ++c) {
            double qv, qp;
            g_value<POINT, Q>(HDc(c)[iq], GTc(c) + qoff, ngc, Lq, qv, qp);
            const double pk = zq ? 0.0 : sm.px[m * CT + c][k];
            const double b = qp * pk;
            T[m][c][0] += qv * pk;
            T[m][c][1] += b * cs;
            T[m][c][2] += b * P.x;
            T[m][c][3] += b * P.y;
            T[m][c][4] += b * P.z;
          }
        }
      }
      };
      if (!POINT && !GP && args.pot.g_span)
        walk(std::true_type{});
      else
        walk(std::false_type{});
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        sm.rp[c][p] += 0.5 * fv[s][c] * S[c][0];
        const double fr = fv[s][c] * invr[s];
        const double radial = dfv[s][c] * S[c][0] - fr * S[c][1];
        Gx[s][c] = xj * radial + fr * S[c][2];
        Gy[s][c] = yj * radial + fr * S[c][3];
        Gz[s][c] = zj * radial + fr * S[c][4];
        S1v[s][c] = S[c][0];
        // p p q terms enter the energy directly (each unordered triplet is
        // seen from both legs) and their legs bypass U'
#pragma unroll
        for (int m = 0; m < JM; ++m) {
          if (args.T1_out && valid[c]) args.T1_out[((size_t)cand[c] * args.npairs + ch.p0 + p) * 2 + m] = T[m][c][0];
          e_acc[c] += 0.5 * pv[m][s][c] * T[m][c][0];
          const double pr = pv[m][s][c] * invr[s];
          const double rq = dpv[m][s][c] * T[m][c][0] - pr * T[m][c][1];
          Hx[s][c] += xj * rq + pr * T[m][c][2];
          Hy[s][c] += yj * rq + pr * T[m][c][3];
          Hz[s][c] += zj * rq + pr * T[m][c][4];
        }
      }
    }
  }
  __syncthreads();
  // per-atom densities from the staged pair contributions (two candidates at
  // a time share one eight-lane pass); pair styles have none
#pragma unroll
  for (int c = 0; c < (args.pot.nodens ? 0 : CT); c += 2) {
    const int c2 = c + 1 < CT ? c + 1 : c;
    atom_sums3<BLOCK>(
        sm.poff, ch.na,
        [&](int q, double& x, double& y, double& z) {
          x = sm.rp[c][q];
          y = sm.rp[c2][q];
          z = 0.0;
        },
        [&](int a, double x, double y, double) {
          sm.rho[c][a] = x;
          sm.rho[c2][a] = y;
        });
  }
  __syncthreads();

  // ---- phase 3: embedding
  double c_acc[CT];
  const double wcell = args.weight[ch.cell];
#pragma unroll
  for (int c = 0; c < CT; ++c) c_acc[c] = 0.0;
  for (int a = tid; a < ch.na; a += BLOCK) {
    if (args.pot.nodens) {
#pragma unroll
      for (int c = 0; c < CT; ++c) sm.dF[c][a] = 0.0;
      continue;
    }
    const int su = args.pot.fn_off[2] + args.types[ch.a0 + a];
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      double rho = sm.rho[c][a];
      if (args.rho_out && valid[c]) args.rho_out[(size_t)cand[c] * args.natoms + ch.a0 + a] = rho;
      const SplHead& HU = HDc(c)[su];
      const double lo = HU.x0, hi = HU.xn;
      if (rho < lo) {
        c_acc[c] += wcell * kDummyWeight * 10. * (lo - rho) * (lo - rho);
        if (!args.pot.embed_extrap) rho = lo;
      } else if (rho > hi) {
        c_acc[c] += wcell * kDummyWeight * 10. * (rho - hi) * (rho - hi);
        if (!args.pot.embed_extrap) rho = hi;
      }
      double dU, d2U;
      e_acc[c] += eval_point<Q>(HU, IVc(c), rho, dU, d2U);
      sm.dF[c][a] = dU;
      if (args.dF_out && valid[c]) args.dF_out[(size_t)cand[c] * args.natoms + ch.a0 + a] = dU;
    }
  }
  __syncthreads();

  // ---- phase 4: pair forces, three-body forces, virial, chunk partials
  const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    double vals[8] = {e_acc[c], 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, c_acc[c]};
#pragma unroll
    for (int s = 0; s < PPL; ++s) {
      if (ia[s] < 0) continue;
      const int p = tid + s * BLOCK;
      const double dFi = sm.dF[c][ia[s]];
      const double sc = aphi[s][c] + brho[s][c] * dFi;
      double tx = ux[s] * sc + dFi * Gx[s][c];
      double ty = uy[s] * sc + dFi * Gy[s][c];
      double tz = uz[s] * sc + dFi * Gz[s][c];
      if (JM) {
        tx += Hx[s][c];
        ty += Hy[s][c];
        tz += Hz[s][c];
      }
      const size_t pg = (size_t)cand[c] * args.npairs + ch.p0 + p;
      if (valid[c]) {
        if (args.G_out)
          reinterpret_cast<double4*>(args.G_out)[pg] = make_double4(Gx[s][c], Gy[s][c], Gz[s][c], S1v[s][c]);
        // stored at the pair's slot in its neighbour's incoming run, which
        // k_resid then reads as one contiguous stream per atom
        double* tv = args.tvec + ((size_t)cand[c] * args.npairs + ipos[s]) * 3;
        tv[0] = tx;
        tv[1] = ty;
        tv[2] = tz;
      }
      sm.rec[p] = make_double4(tx, ty, tz, 0.0);
      const double rx = tx * r[s], ry = ty * r[s], rz = tz * r[s];
      vals[1] -= ux[s] * rx;
      vals[2] -= uy[s] * ry;
      vals[3] -= uz[s] * rz;
      vals[4] -= ux[s] * ry;
      vals[5] -= uy[s] * rz;
      vals[6] -= uz[s] * rx;
    }
    // chunk partials of energy, virial and penalty
    const double w8 = wave_sums8(vals);
    if ((lane & 7) == 0) sm.red8[c][((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1)][wid] = w8;
    __syncthreads();
    // own-force sums: F_a gains T_p over its own pairs
    atom_sums3<BLOCK>(
        sm.poff, ch.na,
        [&](int q, double& x, double& y, double& z) {
          const double4 T = sm.rec[q];
          x = T.x;
          y = T.y;
          z = T.z;
        },
        [&](int a, double x, double y, double z) {
          if (!valid[c]) return;
          double* fa = args.fown + ((size_t)cand[c] * args.natoms + ch.a0 + a) * 3;
          fa[0] = x;
          fa[1] = y;
          fa[2] = z;
        });
    __syncthreads();
  }
  if (tid < 8 * CT) {
    const int c = tid >> 3, q = tid & 7;
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) t += sm.red8[c][q][w];
    if (blockIdx.y * CT + c < args.ncand)
      args.chunk_acc[((size_t)(blockIdx.y * CT + c) * args.nchunks + blockIdx.x) * 8 + q] = t;
  }
}

// ------------------------------------------------------------------ k_resid

// Eight lanes per item; items are the cells, then the atoms (a cell's sum
// over its chunks is a serial chain of loads, up to a few hundred long, which
// then runs beside the atoms' streams instead of after them).  An atom's lanes
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
  const int item = (blockIdx.x * BLOCK + threadIdx.x) / kResidLanes - a.ncells;  // cells: -ncells .. -1
  const int l = threadIdx.x & (kResidLanes - 1);
  double sq = 0.0;
  double* ev = a.errvec ? a.errvec + (size_t)cand * a.nerr : nullptr;
  if (item >= 0 && item < a.natoms) {
    const int i = item;
    const double* tv = a.tvec + (size_t)cand * a.npairs * 3;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    const int b = a.in_off[i], e = a.in_off[i + 1];
    // three of the lane's pairs per iteration (nine loads in flight), each
    // lane's sums still taken in ascending pair order
    constexpr int U = 3;
    int q = b + l;
    for (; q + kResidLanes * (U - 1) < e; q += kResidLanes * U) {
      double x[U], y[U], z[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const double* t = tv + 3 * (size_t)(q + kResidLanes * u);
        x[u] = t[0];
        y[u] = t[1];
        z[u] = t[2];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        sx += x[u];
        sy += y[u];
        sz += z[u];
      }
    }
    for (; q < e; q += kResidLanes) {
      const double* t = tv + 3 * (size_t)q;
      sx += t[0];
      sy += t[1];
      sz += t[2];
    }
    static_assert(kResidLanes == 8, "incoming-pair sums use eight-lane groups");
    sx = group8_sum(sx);
    sy = group8_sum(sy);
    sz = group8_sum(sz);
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
  } else if (item < 0) {
    const int c = item + a.ncells;
    const double* cacc = a.chunk_acc + (size_t)cand * a.nchunks * 8;
    // four independent partial sums (four index-then-value load pairs in
    // flight), combined in a fixed order
    double a4[4] = {0.0, 0.0, 0.0, 0.0};
    int q = a.cell_chunk_off[c];
    const int qe = a.cell_chunk_off[c + 1];
    for (; q + 3 < qe; q += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) a4[u] += cacc[8 * (size_t)a.cell_chunks[q + u] + l];
    }
    for (; q < qe; ++q) a4[0] += cacc[8 * (size_t)a.cell_chunks[q] + l];
    const double acc = (a4[0] + a4[1]) + (a4[2] + a4[3]);
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
      if (ev && !a.nodens) ev[7 * (size_t)a.ncells + 3 * (size_t)a.natoms + c] = er;
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
  // eight independent partial sums per thread keep eight loads in flight
  // (a single chain is bound by load latency on the ~72k ternary partials),
  // combined in a fixed order
  const double* pa = a.partial + (size_t)cand * a.npartial;
  double s8[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int q = threadIdx.x;
  for (; q + 7 * 256 < a.npartial; q += 8 * 256) {
#pragma unroll
    for (int u = 0; u < 8; ++u) s8[u] += pa[q + u * 256];
  }
  for (; q < a.npartial; q += 256) s8[0] += pa[q];
  const double s = ((s8[0] + s8[1]) + (s8[2] + s8[3])) + ((s8[4] + s8[5]) + (s8[6] + s8[7]));
  const double tot = block_sum<256>(s, red);
  if (threadIdx.x != 0) return;
  double F = tot;
  double* ev = a.errvec ? a.errvec + (size_t)cand * a.nerr : nullptr;
  const size_t tail = 8 * (size_t)a.ncells + 3 * (size_t)a.natoms;
  const double* kv = a.knots + (size_t)cand * a.nknot;
  if (a.nodens) {
    // pair styles: no penalty term (src/pots/pot_pair/pot_pair.cpp:136-137)
  } else if (a.meam && a.fper) {
    // one term per f spline (src/pots/pot_gmeam_spline/pot_gmeam_spline.cpp:275-280)
    for (int q = 0; q < a.nf; ++q) {
      const int s2 = a.f0 + q;
      const double* y = kv + a.koff[s2];
      double m = 0.0;
      for (int k = 0; k < a.nknots[s2]; ++k)
        if (fabs(m) < fabs(y[k])) m = y[k];
      const double pe = kDummyWeight * 25. * (1.0 - fabs(m)) * (1.0 - fabs(m));
      F += pe * pe;
      if (ev) ev[tail + q] = pe;
    }
  } else if (a.meam) {
    // PotFns::get_max_y_mag over the f splines (src/pot_fns.cpp:100-111); the
    // jmeam styles add one such term over the p (and p2) splines
    // (src/pots/pot_jmeam2/pot_jmeam2.cpp:308-324)
    for (int g = 0; g < a.npgrp; ++g) {
      double mx = 0.0;
      for (int q = 0; q < a.nf; ++q) {
        const int s2 = a.pgrp0[g] + q;
        const double* y = kv + a.koff[s2];
        double m = 0.0;
        for (int k = 0; k < a.nknots[s2]; ++k)
          if (fabs(m) < fabs(y[k])) m = y[k];
        if (fabs(mx) < fabs(m)) mx = m;
      }
      const double m = fabs(mx);
      const double pe = kDummyWeight * 25. * (1.0 - m) * (1.0 - m);
      F += pe * pe;
      if (ev) ev[tail + g] = pe;
    }
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

template <int BLOCK, int PPL, int MAXA, int CT>
static void launch_atoms_t(hipStream_t st, const AtomArgs& a, int nchunks, int ncand) {
  const int style = a.pot.style;
  const bool meam = style == kMeamSpline || style == kMeam || style == kGMeamSpline || style == kGMeam;
  const bool point = style == kEam || style == kMeam || style == kGMeam || style == kPair;
  const bool q = a.pot.quintic != 0;
  const size_t dyn = atoms_dyn_lds(a.pot, CT);
  const bool tg = atoms_tables_global(a.pot, CT);  // always with gpair
  dim3 grid(nchunks, (ncand + CT - 1) / CT), block(BLOCK);
#define MZ_LAUNCH(M, P, Q_, TG_, GP_)                                                              \
  hipLaunchKernelGGL((k_atoms<BLOCK, PPL, MAXA, CT, M, P, Q_, TG_, GP_, 0>), grid, block, dyn, st, a)
  if (meam && a.pot.gpair) {
    if (point) { if (q) MZ_LAUNCH(true, true, true, true, true); else MZ_LAUNCH(true, true, false, true, true); }
    else { if (q) MZ_LAUNCH(true, false, true, true, true); else MZ_LAUNCH(true, false, false, true, true); }
  } else if (meam && tg) {
    if (point) { if (q) MZ_LAUNCH(true, true, true, true, false); else MZ_LAUNCH(true, true, false, true, false); }
    else { if (q) MZ_LAUNCH(true, false, true, true, false); else MZ_LAUNCH(true, false, false, true, false); }
  } else if (meam) {
    if (point) { if (q) MZ_LAUNCH(true, true, true, false, false); else MZ_LAUNCH(true, true, false, false, false); }
    else { if (q) MZ_LAUNCH(true, false, true, false, false); else MZ_LAUNCH(true, false, false, false, false); }
  } else {
    if (point) { if (q) MZ_LAUNCH(false, true, true, false, false); else MZ_LAUNCH(false, true, false, false, false); }
    else { if (q) MZ_LAUNCH(false, false, true, false, false); else MZ_LAUNCH(false, false, false, false, false); }
  }
#undef MZ_LAUNCH
}

// jmeam styles: one candidate per workgroup, spline tables in global memory
// (heads and the angular records staged), JM = 1 or 2 extra terms.
template <int BLOCK, int PPL, int MAXA>
static void launch_atoms_jm(hipStream_t st, const AtomArgs& a, int nchunks, int ncand) {
  const int style = a.pot.style;
  const bool point = style == kJMeam || style == kJMeam2;
  const bool q = a.pot.quintic != 0;
  const size_t dyn = atoms_dyn_lds(a.pot, 1);
  dim3 grid(nchunks, ncand), block(BLOCK);
#define MZ_LAUNCH(P, Q_, JM_) \
  hipLaunchKernelGGL((k_atoms<BLOCK, PPL, MAXA, 1, true, P, Q_, true, false, JM_>), grid, block, dyn, st, a)
  if (a.pot.njm == 1) {
    if (point) { if (q) MZ_LAUNCH(true, true, 1); else MZ_LAUNCH(true, false, 1); }
    else { if (q) MZ_LAUNCH(false, true, 1); else MZ_LAUNCH(false, false, 1); }
  } else {
    if (point) { if (q) MZ_LAUNCH(true, true, 2); else MZ_LAUNCH(true, false, 2); }
    else { if (q) MZ_LAUNCH(false, true, 2); else MZ_LAUNCH(false, false, 2); }
  }
#undef MZ_LAUNCH
}

// MEAM potentials whose staged tables would take more than 16 KB of LDS per
// workgroup read them from global memory instead (measured: the ternary
// potential's 17.5 KB tables cost a candidate per workgroup and a workgroup
// per CU; the one-element potential's 3.9 KB are faster staged).
bool atoms_tables_global(const PotDesc& pot, int ct) {
  const bool meam = pot.ngco > 0;
  return meam && (pot.gpair || pot.njm > 0 || (size_t)ct * (size_t)pot.table_bytes > 16 * 1024);
}

size_t atoms_dyn_lds(const PotDesc& pot, int ct) {
  const bool meam = pot.ngco > 0;
  const size_t g = meam ? (size_t)pot.ngco * 8 * g_record_doubles(pot.quintic != 0) : 0;
  if (atoms_tables_global(pot, ct)) return (size_t)ct * ((size_t)pot.nspl * sizeof(SplHead) + g);
  return (size_t)ct * ((size_t)pot.table_bytes + g);
}

size_t atoms_static_lds(int block, int ppl, int ct, const PotDesc& pot) {
  const size_t P = (size_t)block * ppl;
  return P * 32 + (size_t)(ct > 1 ? ct - 1 : 1) * P * 8 + (size_t)ct * P * 8 + (size_t)ct * 64 * 8 * 2 + 65 * 4 +
         (size_t)ct * 8 * (block / 64) * 8 + (pot.gpair ? P * 4 : 4) +
         (pot.njm > 0 ? (size_t)pot.njm * ct * P * 8 : 8);
}

int launch_atoms(hipStream_t st, const AtomArgs& a, int nchunks, int ncand, int block, int ppl, int ct) {
#define MZ_SHAPE(B, P_)                                                       \
  if (block == B && ppl == P_) {                                             \
    if (a.pot.njm > 0) {                                                     \
      if (ct != 1) return -1;                                                \
      launch_atoms_jm<B, P_, 64>(st, a, nchunks, ncand);                     \
    } else if (ct == 1) launch_atoms_t<B, P_, 64, 1>(st, a, nchunks, ncand); \
    else if (ct == 2) launch_atoms_t<B, P_, 64, 2>(st, a, nchunks, ncand);   \
    else if (ct == 4) launch_atoms_t<B, P_, 64, 4>(st, a, nchunks, ncand);   \
    else return -1;                                                          \
    return 0;                                                                \
  }
  MZ_SHAPE(256, 2)
  MZ_SHAPE(256, 1)
  MZ_SHAPE(128, 2)
  MZ_SHAPE(128, 4)
  MZ_SHAPE(512, 1)
  MZ_SHAPE(64, 4)
#undef MZ_SHAPE
  return -1;
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
