// Definitions shared by the host code and the gfx950 kernels of libmeamz.
#pragma once

#include <cstdint>

namespace mz {

constexpr int kMaxKnots = 32;        // knots per spline
constexpr int kMaxTypes = 8;         // atom types
constexpr int kMaxSplines = 96;      // splines per potential
constexpr int kMaxImage = 127;       // |image index| per direction (packed as int8)
constexpr double kDummyWeight = 100.0;  // DUMMY_WEIGHT, src/pots/pot_pair/pot_pair.h:39

enum Style : int {
  kEamSpline = 0,
  kMeamSpline = 1,
  kEam = 2,
  kMeam = 3,
  kGMeamSpline = 4,
  kGMeam = 5,
  kPairSpline = 6,
  kPair = 7,
  kJMeamSpline = 8,
  kJMeam = 9,
  kJMeam2Spline = 10,
  kJMeam2 = 11
};

// Per-spline header in the device table (16 doubles = 128 bytes).
struct SplHead {
  double x0, xn, step, invstep;
  double xcut;          // xn * (1 + 1e-8): knot-style "inside" limit (src/spline.cpp:226)
  double y0, yn;        // end values for point extrapolation
  double gL, gR;        // boundary slopes (grad_)
  double g2L, g2R;      // boundary second derivatives (grad2_, zero for cubic)
  double invstep2;      // invstep^2
  int32_t n, basis;     // knots, 3 or 5
  int32_t coef_off;     // offset (in intervals) of this spline's interval records
  int32_t pad;
  double pad2[2];
};
static_assert(sizeof(SplHead) == 128, "SplHead layout");

// One interval of a spline in power basis of the shift t in [0,1]:
//   value(t) = c0 + c1 t + c2 t^2 + c3 t^3 + c4 t^4 + c5 t^5
// plus the knot position x_k and 1/(x_{k+1}-x_k) for the shift.
struct SplInterval {
  double c[6];
  double xk;
  double invh;
};
static_assert(sizeof(SplInterval) == 64, "SplInterval layout");

// Work unit of the atom kernel: a run of consecutive atoms of one cell.
struct Chunk {
  int32_t cell;
  int32_t a0;   // first atom (global index)
  int32_t na;   // atoms
  int32_t p0;   // first pair (global index)
  int32_t np;   // pairs
  int32_t pad[3];
};

// Static (per-context) description of the potential used by every kernel.
struct PotDesc {
  int32_t style;
  int32_t nt;
  int32_t nspl;
  int32_t embed_extrap;
  int32_t fn_off[9];    // first spline of phi, rho, U, f, g, p, q, p2, q2
  int32_t ncoef;        // total interval records in one table
  int32_t table_bytes;  // heads + intervals of one table
  int32_t nknot;        // knot-vector length
  int32_t quintic;      // any spline5eq present
  int32_t gco0;         // first interval record of the g splines (MEAM)
  int32_t ngco;         // interval records of the g splines
  int32_t g_keep_all;   // every g spline's x_n (1 + 1e-8) exceeds any cosine: no triplet is cut
  int32_t gpair;        // g indexed by (type i, types j k) (gmeam styles, more than one type)
  int32_t ng;           // number of g splines (nt, or nt * nt (nt + 1) / 2 with gpair)
  int32_t nodens;       // pair styles: phi only, no density, embedding or density penalty
  int32_t njm;          // jmeam styles: extra p_ij p_ik q(cos) energy terms (1, or 2 for jmeam2)
  int32_t g_span;       // knot style, every g spline spans [-1, 1] (x_0 <= -1, no cut): u = (cos - x_0) / h
                        // exceeds -1, so the interval is min(trunc(u), n - 2)
};

// Doubles per g-interval record in the compact LDS copies of the g tables
// (c0..c3 for cubic, c0..c5 for quintic splines).  The copies are stored as
// planes of coefficient pairs: plane m holds (c_2m, c_2m+1) of every record,
// 16 bytes apart, so that lanes reading different intervals with one 16-byte
// load touch different LDS banks (records 8 apart share a bank quad).
constexpr int g_record_doubles(bool quintic) { return quintic ? 6 : 4; }

// Packs an image triple into one 32-bit word (8 bits each, signed).
inline uint32_t pack_image(int ix, int iy, int iz) {
  return (uint32_t)(uint8_t)(int8_t)ix | ((uint32_t)(uint8_t)(int8_t)iy << 8) |
         ((uint32_t)(uint8_t)(int8_t)iz << 16);
}

}  // namespace mz
