// Device helpers shared by the forward and adjoint kernels: reductions and
// spline evaluation from the per-interval power-basis tables.
#pragma once

#include <hip/hip_runtime.h>

#include "mz_common.h"

namespace mz {

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double* scratch) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) s += scratch[w];
  }
  return s;  // valid in thread 0
}

__device__ __forceinline__ int pair_ij(int ti, int tj, int nt) {
  return (ti <= tj) ? ti * nt + tj - ((ti * (ti + 1)) >> 1) : tj * nt + ti - ((tj * (tj + 1)) >> 1);
}

template <bool Q>
__device__ __forceinline__ void poly(const SplInterval& I, double t, double invstep, double& v, double& d) {
  if (Q) {
    v = I.c[0] + t * (I.c[1] + t * (I.c[2] + t * (I.c[3] + t * (I.c[4] + t * I.c[5]))));
    d = (I.c[1] + t * (2.0 * I.c[2] + t * (3.0 * I.c[3] + t * (4.0 * I.c[4] + t * 5.0 * I.c[5])))) * invstep;
  } else {
    v = I.c[0] + t * (I.c[1] + t * (I.c[2] + t * I.c[3]));
    d = (I.c[1] + t * (2.0 * I.c[2] + t * 3.0 * I.c[3])) * invstep;
  }
}

// Value and d/dt of an interval polynomial from its coefficients (c0, c1),
// (c2, c3) and, quintic, (c4, c5), by one Horner recurrence carrying the
// derivative beside the value (5 FMAs cubic, 9 quintic).
template <bool Q>
__device__ __forceinline__ void horner_d(const double2& c01, const double2& c23, const double2& c45, double t,
                                         double& v, double& d) {
  double p, q;
  if (Q) {
    p = fma(c45.y, t, c45.x);
    q = fma(c45.y, t, p);
    p = fma(p, t, c23.y);
    q = fma(q, t, p);
    p = fma(p, t, c23.x);
  } else {
    p = fma(c23.y, t, c23.x);
    q = fma(c23.y, t, p);
  }
  if (Q) q = fma(q, t, p);
  p = fma(p, t, c01.y);
  d = fma(q, t, p);
  v = fma(p, t, c01.x);
}

// Knot positions of one interval record, x_k and 1 / (x_(k+1) - x_k): what
// locating r needs of a spline table.  The knots sit at the same x for every
// candidate, so the kernels stage one copy of them in LDS.
struct KnotRec {
  double xk, invh;
};

// Knot-style location (Spline::get_knotshift src/spline.cpp:221-246): the
// interval k with x_k <= r < x_(k+1) (clamped to the first and last) and the
// shift t in it, from the spline's records (SplInterval or KnotRec).  Returns
// false when r lies beyond x_n (1 + 1e-8).
template <class Rec>
__device__ __forceinline__ bool locate_knot(const SplHead& H, const Rec* I, double r, int& k, double& t) {
  if (!(r < H.xcut)) return false;
  const int nm2 = H.n - 2;
  k = (int)floor((r - H.x0) * H.invstep);
  k = k < 0 ? 0 : (k > nm2 ? nm2 : k);
  while (k > 0 && I[k].xk > r) --k;
  while (k < nm2 && I[k + 1].xk <= r) ++k;
  t = (r - I[k].xk) * I[k].invh;
  return true;
}

// Knot-style evaluation (locate_knot, then splint_comb by knot).  Returns
// false when r lies beyond x_n (1 + 1e-8).
template <bool Q>
__device__ __forceinline__ bool eval_knot(const SplHead& H, const SplInterval* IV, double r, double& v, double& d) {
  int k;
  double t;
  if (!locate_knot(H, IV, r, k, t)) return false;
  poly<Q>(IV[k], t, H.invstep, v, d);
  return true;
}

// Point evaluation with end extrapolation (splint_comb(double) and
// splint_grad2(double), src/spline3eq.h:204-266, src/spline5eq.h:218-286).
template <bool Q>
__device__ __forceinline__ double eval_point(const SplHead& H, const SplInterval* IV, double r, double& d, double& d2) {
  if (r < H.x0) {
    const double e = r - H.x0;
    d = H.gL + H.g2L * e;
    d2 = H.g2L;
    return H.y0 + H.gL * e + 0.5 * H.g2L * e * e;
  }
  if (r > H.xn) {
    const double e = r - H.xn;
    d = H.gR + H.g2R * e;
    d2 = H.g2R;
    return H.yn + H.gR * e + 0.5 * H.g2R * e * e;
  }
  const double delta = r - H.x0;
  int k = (int)__dmul_rn(delta, H.invstep);
  if (k >= H.n - 1) k = H.n - 2;
  if (k < 0) k = 0;
  const double t = __dmul_rn(__dadd_rn(delta, -__dmul_rn((double)k, H.step)), H.invstep);
  const SplInterval& I = IV[H.coef_off + k];
  double v;
  poly<Q>(I, t, H.invstep, v, d);
  if (Q)
    d2 = (2.0 * I.c[2] + t * (6.0 * I.c[3] + t * (12.0 * I.c[4] + t * 20.0 * I.c[5]))) * H.invstep2;
  else
    d2 = (2.0 * I.c[2] + t * 6.0 * I.c[3]) * H.invstep2;
  return v;
}

// Evaluates one radial function of a pair: knot style (spline styles) or point
// style ("eam"/"meam": inside iff r <= x_n).  Returns false when outside.
template <bool POINT, bool Q>
__device__ __forceinline__ bool eval_radial(const SplHead& H, const SplInterval* IV, double r, double& v,
                                            double& d) {
  if (POINT) {
    if (!(r <= H.xn)) return false;
    double d2;
    v = eval_point<Q>(H, IV, r, d, d2);
    return true;
  }
  return eval_knot<Q>(H, IV + H.coef_off, r, v, d);
}

// Lane moves inside a row of sixteen lanes as DPP operand modes (VALU, no LDS
// traffic): quad_perm [1,0,3,2] and [2,3,0,1] exchange lanes i and i^1, i and
// i^2; row_half_mirror maps lane i of each group of eight to 7 - i; row_ror:8
// exchanges lanes i and i^8.
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppRor8 = 0x128;

template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

// Sum over each aligned group of eight lanes, the same in every lane of the
// group (pairs, then quads, then the two quads); all eight lanes must be active.
__device__ __forceinline__ double group8_sum(double x) {
  x += dpp_move<kDppXor1>(x);
  x += dpp_move<kDppXor2>(x);
  return x + dpp_move<kDppHalfMirror>(x);
}

// Sums three per-pair values (vals[q] for q in the atom's pair range) for
// every atom of the chunk, eight lanes per atom, fixed order; out(a, x, y, z)
// receives the sums.  Every lane of the workgroup must call it.
template <int BLOCK, class Get, class Out>
__device__ __forceinline__ void atom_sums3(const int* poff, int na, Get get, Out out) {
  const int l = threadIdx.x & 7;
  for (int base = 0; base < na; base += BLOCK / 8) {
    const int a = base + (int)(threadIdx.x >> 3);
    double x = 0.0, y = 0.0, z = 0.0;
    if (a < na) {
      for (int q = poff[a] + l; q < poff[a + 1]; q += 8) {
        double vx, vy, vz;
        get(q, vx, vy, vz);
        x += vx;
        y += vy;
        z += vz;
      }
    }
    x = group8_sum(x);
    y = group8_sum(y);
    z = group8_sum(z);
    if (a < na && l == 0) out(a, x, y, z);
  }
}

// Wave totals of eight per-lane values: three halving exchanges (the first
// two across rows by shuffles, the third by DPP) leave each lane with one
// value index summed over eight lanes; an eight-lane DPP sum finishes it.  Lane l with (l & 7) == 0 returns the total of
// value index ((l >> 5) & 1) * 4 + ((l >> 4) & 1) * 2 + ((l >> 3) & 1).
__device__ __forceinline__ double wave_sums8(const double (&v)[8]) {
  const int lane = threadIdx.x & 63;
  double a[4], b[2];
  const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double send = h5 ? v[i] : v[i + 4];
    const double keep = h5 ? v[i + 4] : v[i];
    a[i] = keep + __shfl_xor(send, 32, 64);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double send = h4 ? a[i] : a[i + 2];
    const double keep = h4 ? a[i + 2] : a[i];
    b[i] = keep + __shfl_xor(send, 16, 64);
  }
  const double c = (h3 ? b[1] : b[0]) + dpp_move<kDppRor8>(h3 ? b[0] : b[1]);
  return group8_sum(c);
}

}  // namespace mz
