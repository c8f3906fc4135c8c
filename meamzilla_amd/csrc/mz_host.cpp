// Host side of libmeamz: context creation (pair lists, work chunks, device
// upload) and the C-ABI entry points declared in include/meamz.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/meamz.h"
#include "mz_common.h"
#include "mz_kernels.h"
#include "mz_adjoint.h"
#include "mz_ctx.h"
#include "mz_list.h"

using mz::DevBuf;
using mz::Grid;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define MZ_HIP(call)                                                                         \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(MZ_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_));             \
  } while (0)

int pair_ij_h(int ti, int tj, int nt) {
  return (ti <= tj) ? ti * nt + tj - ((ti * (ti + 1)) / 2) : tj * nt + ti - ((tj * (tj + 1)) / 2);
}

}  // namespace

namespace {

int build_layout(mz_ctx* c, const mz_layout* L) {
  if (!L || L->style < 0 || L->style > 11) return fail(MZ_ERR_LAYOUT, "unknown style");
  if (L->ntypes < 1 || L->ntypes > mz::kMaxTypes) return fail(MZ_ERR_LAYOUT, "ntypes out of range");
  c->style = L->style;
  c->nt = L->ntypes;
  const bool gmeam = L->style == mz::kGMeamSpline || L->style == mz::kGMeam;
  // jmeam styles: MEAM plus p Pair_ij, q Triplet_i (jmeam2: also p2, q2),
  // src/pots/pot_jmeam/pot_jmeam.cpp:30-47, src/pots/pot_jmeam2/pot_jmeam2.cpp:30-46
  const int njm = (L->style == mz::kJMeamSpline || L->style == mz::kJMeam)     ? 1
                  : (L->style == mz::kJMeam2Spline || L->style == mz::kJMeam2) ? 2
                                                                                 : 0;
  c->meam = (L->style == mz::kMeamSpline || L->style == mz::kMeam || gmeam || njm > 0);
  c->point = (L->style == mz::kEam || L->style == mz::kMeam || L->style == mz::kGMeam || L->style == mz::kPair ||
              L->style == mz::kJMeam || L->style == mz::kJMeam2);
  // pair styles: phi Pair_ij alone (src/pots/pot_pair/pot_pair.cpp:24-37)
  c->pair = (L->style == mz::kPairSpline || L->style == mz::kPair);
  const int nt = c->nt, npij = nt * (nt + 1) / 2;
  // g: Triplet_i (meam), Triplet_i_jk (gmeam, src/pots/pot_gmeam/pot_gmeam.cpp:31-36)
  const int counts[9] = {npij, nt, nt, npij, gmeam ? nt * npij : nt, npij, nt, npij, nt};
  const int nfn = c->meam ? 5 + 2 * njm : (c->pair ? 1 : 3);
  int tot = 0;
  for (int f = 0; f < 9; ++f) c->pot.fn_off[f] = 0;
  for (int f = 0; f < nfn; ++f) {
    c->pot.fn_off[f] = tot;
    tot += counts[f];
  }
  if (L->nspl != tot || tot > mz::kMaxSplines)
    return fail(MZ_ERR_LAYOUT, "layout has " + std::to_string(L->nspl) + " splines, style needs " + std::to_string(tot));
  c->nspl = tot;
  c->basis.assign(L->basis, L->basis + tot);
  c->nknots.assign(L->nknots, L->nknots + tot);
  c->x0.assign(L->x0, L->x0 + tot);
  c->xn.assign(L->xn, L->xn + tot);
  c->bc.assign(L->bc, L->bc + 2 * tot);
  c->koff.resize(tot);
  c->coef_off.resize(tot);
  c->grids.resize(tot);
  int quintic = 0, koff = 0, coff = 0;
  for (int s = 0; s < tot; ++s) {
    if (c->basis[s] != 3 && c->basis[s] != 5) return fail(MZ_ERR_LAYOUT, "basis must be 3 or 5");
    if (c->nknots[s] < 3 || c->nknots[s] > mz::kMaxKnots)
      return fail(MZ_ERR_CAPACITY, "spline " + std::to_string(s) + ": knot count outside [3, 32]");
    if (!(c->xn[s] > c->x0[s])) return fail(MZ_ERR_LAYOUT, "spline " + std::to_string(s) + ": x_n <= x_0");
    for (int q = 0; q < 2; ++q)
      if (c->bc[2 * s + q] < -1 || c->bc[2 * s + q] > 1) return fail(MZ_ERR_LAYOUT, "bad boundary condition");
    if (c->basis[s] == 5) quintic = 1;
    c->koff[s] = koff;
    koff += c->nknots[s] + 4;
    c->grids[s].init(c->nknots[s], c->x0[s], c->xn[s]);
  }
  // interval records: radial and embedding splines first, then the angular
  // ones (g, and q, q2 of the jmeam styles), which the kernels stage and
  // accumulate as one contiguous range
  auto angular = [&](int s) {
    for (int f : {4, 6, 8})
      if (f < nfn && s >= c->pot.fn_off[f] && s < c->pot.fn_off[f] + counts[f]) return true;
    return false;
  };
  for (int pass = 0; pass < 2; ++pass)
    for (int s = 0; s < tot; ++s)
      if (angular(s) == (pass == 1)) {
        c->coef_off[s] = coff;
        coff += c->nknots[s] - 1;
      }
  c->nknot = koff;
  c->ncoef = coff;
  c->table_bytes = (int)(sizeof(mz::SplHead) * tot + sizeof(mz::SplInterval) * coff);
  if (c->table_bytes > 40 * 1024)
    return fail(MZ_ERR_CAPACITY, "spline tables exceed the 40 KB LDS budget of one workgroup");
  c->pot.style = c->style;
  c->pot.nt = nt;
  c->pot.nspl = tot;
  c->pot.embed_extrap = L->embed_extrap;
  c->pot.ncoef = coff;
  c->pot.table_bytes = c->table_bytes;
  c->pot.nknot = koff;
  c->pot.quintic = quintic;
  c->pot.gco0 = c->meam ? c->coef_off[c->pot.fn_off[4]] : coff;
  c->pot.ngco = c->meam ? coff - c->pot.gco0 : 0;
  c->pot.ng = c->meam ? counts[4] : 0;
  c->pot.gpair = gmeam && nt > 1;  // one type: gmeam is meam
  c->pot.nodens = c->pair ? 1 : 0;
  c->pot.njm = njm;
  c->nfpen = gmeam ? npij : (c->meam ? 1 + njm : 0);  // f-max, then one p-max per p function
  // |cos| <= 1 + a few ulps, so a knot-style g spline ending at or beyond 1
  // never cuts a triplet
  c->pot.g_keep_all = 0;
  c->pot.g_span = 0;
  if (c->meam && !c->point) {
    bool all = true;
    for (int t = 0; t < counts[4]; ++t) {
      const int s = c->pot.fn_off[4] + t;
      if (!(c->xn[s] * (1. + 1.e-8) > 1.0 + 1.e-12)) all = false;
    }
    c->pot.g_keep_all = all ? 1 : 0;
    bool span = all;
    for (int t = 0; t < counts[4]; ++t)
      if (!(c->x0[c->pot.fn_off[4] + t] <= -1.0)) span = false;
    c->pot.g_span = span ? 1 : 0;
  }
  return MZ_OK;
}

// Pair lists of one cell in the reference's order (Config::find_pairs
// src/config.cpp:162-229): for each atom i, each atom j of the cell and each
// image (ix, iy, iz) within +-ceil(rc |a_q x a_r| / V), keep r < rc.  The image
// ranges are pruned per (i, j) with the fractional separation, which cannot
// drop a pair: |d| >= |s_p + n_p| h_p for the height h_p of direction p.
struct CellPairs {
  std::vector<std::vector<uint2>> per_atom;
  std::string err;
};

void cell_pairs(const mz_ctx* c, int cc, const double* box, const double* pos, double max_rcut, double vol,
                CellPairs& out) {
  const int n = c->natoms_c[cc];
  const int base = c->abase[cc];
  const double* a0 = box;
  const double* a1 = box + 3;
  const double* a2 = box + 6;
  double c12[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2], a1[0] * a2[1] - a1[1] * a2[0]};
  double c20[3] = {a2[1] * a0[2] - a2[2] * a0[1], a2[2] * a0[0] - a2[0] * a0[2], a2[0] * a0[1] - a2[1] * a0[0]};
  double c01[3] = {a0[1] * a1[2] - a0[2] * a1[1], a0[2] * a1[0] - a0[0] * a1[2], a0[0] * a1[1] - a0[1] * a1[0]};
  const double ih[3] = {std::sqrt(c12[0] * c12[0] + c12[1] * c12[1] + c12[2] * c12[2]) / vol,
                        std::sqrt(c20[0] * c20[0] + c20[1] * c20[1] + c20[2] * c20[2]) / vol,
                        std::sqrt(c01[0] * c01[0] + c01[1] * c01[1] + c01[2] * c01[2]) / vol};
  const int sc[3] = {(int)std::ceil(max_rcut * ih[0]), (int)std::ceil(max_rcut * ih[1]),
                     (int)std::ceil(max_rcut * ih[2])};
  out.per_atom.assign(n, {});
  if (sc[0] <= 0 || sc[1] <= 0 || sc[2] <= 0) {
    out.err = "Cell #" + std::to_string(cc) + " is too small or potential is too big";
    return;
  }
  if (sc[0] > mz::kMaxImage || sc[1] > mz::kMaxImage || sc[2] > mz::kMaxImage) {
    out.err = "Cell #" + std::to_string(cc) + " needs more than 127 periodic images";
    return;
  }
  // fractional coordinates: s = x . (a_q x a_r) / V
  std::vector<double> fr(3 * (size_t)n);
  for (int a = 0; a < n; ++a) {
    const double* p = pos + 3 * (size_t)(base + a);
    fr[3 * a + 0] = (p[0] * c12[0] + p[1] * c12[1] + p[2] * c12[2]) / vol;
    fr[3 * a + 1] = (p[0] * c20[0] + p[1] * c20[1] + p[2] * c20[2]) / vol;
    fr[3 * a + 2] = (p[0] * c01[0] + p[1] * c01[1] + p[2] * c01[2]) / vol;
  }
  const double sgn = (a0[0] * c12[0] + a0[1] * c12[1] + a0[2] * c12[2]) < 0 ? -1.0 : 1.0;
  const double margin = 1e-7;
  const int nt = c->nt;
  for (int ia = 0; ia < n; ++ia) {
    const int gi = base + ia;
    const double* pi = pos + 3 * (size_t)gi;
    const int ti = c->types[gi];
    auto& lst = out.per_atom[ia];
    for (int j = 0; j < n; ++j) {
      const int gj = base + j;
      const double* pj = pos + 3 * (size_t)gj;
      int lo[3], hi[3];
      for (int q = 0; q < 3; ++q) {
        const double ds = sgn * (fr[3 * j + q] - fr[3 * ia + q]);
        const double lim = max_rcut * ih[q] + margin;
        lo[q] = std::max(-sc[q], (int)std::ceil(-lim - ds));
        hi[q] = std::min(sc[q], (int)std::floor(lim - ds));
      }
      const int tj = c->types[gj];
      const int pij = pair_ij_h(ti, tj, nt);
      for (int ix = lo[0]; ix <= hi[0]; ++ix)
        for (int iy = lo[1]; iy <= hi[1]; ++iy)
          for (int iz = lo[2]; iz <= hi[2]; ++iz) {
            if (ia == j && ix == 0 && iy == 0 && iz == 0) continue;
            double d[3];
            d[0] = pj[0] + ix * a0[0] + iy * a1[0] + iz * a2[0] - pi[0];
            d[1] = pj[1] + ix * a0[1] + iy * a1[1] + iz * a2[1] - pi[1];
            d[2] = pj[2] + ix * a0[2] + iy * a1[2] + iz * a2[2] - pi[2];
            const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            if (r >= max_rcut) continue;
            // check_pair: first-knot test and inside flags
            const int fns[5] = {c->pot.fn_off[0] + pij, c->pot.fn_off[1] + tj, c->pot.fn_off[3] + pij,
                                c->pot.fn_off[5] + pij, c->pot.fn_off[7] + pij};
            bool inside = false;
            for (int q = 0; q < (c->meam ? 3 + c->pot.njm : (c->pair ? 1 : 2)); ++q) {
              const Grid& g = c->grids[fns[q]];
              if (r < g.x[0]) {
                char buf[200];
                std::snprintf(buf, sizeof(buf),
                              "Pair distance too short for potential (r=%g) cell=%d atom=%d neigh=%d", r, cc, ia, j);
                out.err = buf;
                return;
              }
              inside |= c->point ? (r <= g.x[g.n - 1]) : (r < g.xcut());
            }
            if (!inside) continue;
            lst.push_back(make_uint2((unsigned)gj, mz::pack_image(ix, iy, iz)));
          }
    }
  }
}

// jmeam styles: a triplet is kept when any of its terms is present -- the
// cosine inside g or q (spline styles), both legs inside f or p (point
// styles); check_triplet src/pots/pot_jmeam/triplet_jmeam.cpp:48-88,
// src/pots/pot_jmeam_spline/triplet_jmeam_spline.cpp:49-88.
long long count_triplets_jmeam(const mz_ctx* c, int gi, const double* pos, const double* box) {
  const long long b = c->atom_poff[gi], e = c->atom_poff[gi + 1], n = e - b;
  const int ti = c->types[gi], nt = c->nt, nterm = 1 + c->pot.njm;
  std::vector<double> u(3 * (size_t)n), rr((size_t)n);
  std::vector<int> tj((size_t)n);
  for (long long p = b; p < e; ++p) {
    const int j = (int)c->pairs[p].x;
    const uint32_t im = c->pairs[p].y;
    const int ix = (int8_t)(im & 0xff), iy = (int8_t)((im >> 8) & 0xff), iz = (int8_t)((im >> 16) & 0xff);
    const double* pj = pos + 3 * (size_t)j;
    const double* pi = pos + 3 * (size_t)gi;
    double d[3];
    for (int q = 0; q < 3; ++q) d[q] = pj[q] + ix * box[q] + iy * box[3 + q] + iz * box[6 + q] - pi[q];
    const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int q = 0; q < 3; ++q) u[3 * (p - b) + q] = d[q] / r;
    rr[p - b] = r;
    tj[p - b] = c->types[j];
  }
  long long cnt = 0;
  for (long long p = 0; p < n - 1; ++p)
    for (long long q = p + 1; q < n; ++q) {
      bool keep = false;
      for (int m = 0; m < nterm && !keep; ++m) {
        if (c->point) {
          const int f = m == 0 ? 3 : 3 + 2 * m;
          const Grid& gj = c->grids[c->pot.fn_off[f] + pair_ij_h(ti, tj[p], nt)];
          const Grid& gk = c->grids[c->pot.fn_off[f] + pair_ij_h(ti, tj[q], nt)];
          keep = rr[p] <= gj.x[gj.n - 1] && rr[q] <= gk.x[gk.n - 1];
        } else {
          const double cs = u[3 * p] * u[3 * q] + u[3 * p + 1] * u[3 * q + 1] + u[3 * p + 2] * u[3 * q + 2];
          keep = cs < c->grids[c->pot.fn_off[4 + 2 * m] + ti].xcut();
        }
      }
      cnt += keep;
    }
  return cnt;
}

// Triplet count of one atom as the reference's find_triplets keeps them.
long long count_triplets(const mz_ctx* c, int gi, const double* pos, const double* box) {
  if (!c->meam) return 0;
  if (c->pot.njm) return count_triplets_jmeam(c, gi, pos, box);
  const long long b = c->atom_poff[gi], e = c->atom_poff[gi + 1];
  const int ti = c->types[gi];
  const int nt = c->nt;
  if (c->point) {
    long long m = 0;
    for (long long p = b; p < e; ++p) {
      const int j = (int)c->pairs[p].x;
      const Grid& g = c->grids[c->pot.fn_off[3] + pair_ij_h(ti, c->types[j], nt)];
      // r recomputed exactly as the pair builder did
      const uint32_t im = c->pairs[p].y;
      const int ix = (int8_t)(im & 0xff), iy = (int8_t)((im >> 8) & 0xff), iz = (int8_t)((im >> 16) & 0xff);
      const double* pj = pos + 3 * (size_t)j;
      const double* pi = pos + 3 * (size_t)gi;
      double d[3];
      for (int q = 0; q < 3; ++q) d[q] = pj[q] + ix * box[q] + iy * box[3 + q] + iz * box[6 + q] - pi[q];
      const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (r <= g.x[g.n - 1]) ++m;
    }
    // a point-style triplet is kept iff both legs lie inside f
    return m * (m - 1) / 2;
  }
  // g of the triplet: by the type of i (meam), or by i and the types of j, k
  // (gmeam, src/pot_fns.cpp:154-171)
  const int npij = nt * (nt + 1) / 2;
  const long long n = e - b;
  if (c->pot.g_keep_all) return n * (n - 1) / 2;  // |cos| <= 1: every triplet lies inside every g
  if (!c->pot.gpair && c->grids[c->pot.fn_off[4] + ti].xcut() > 1.0 + 1e-12) return n * (n - 1) / 2;
  std::vector<double> u(3 * (size_t)n);
  std::vector<int> tk((size_t)n);
  for (long long p = b; p < e; ++p) {
    const int j = (int)c->pairs[p].x;
    tk[p - b] = c->types[j];
    const uint32_t im = c->pairs[p].y;
    const int ix = (int8_t)(im & 0xff), iy = (int8_t)((im >> 8) & 0xff), iz = (int8_t)((im >> 16) & 0xff);
    const double* pj = pos + 3 * (size_t)j;
    const double* pi = pos + 3 * (size_t)gi;
    double d[3];
    for (int q = 0; q < 3; ++q) d[q] = pj[q] + ix * box[q] + iy * box[3 + q] + iz * box[6 + q] - pi[q];
    const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int q = 0; q < 3; ++q) u[3 * (p - b) + q] = d[q] / r;
  }
  long long cnt = 0;
  for (long long p = 0; p < n - 1; ++p)
    for (long long q = p + 1; q < n; ++q) {
      const int gs = c->pot.gpair ? ti * npij + pair_ij_h(tk[p], tk[q], nt) : ti;
      const double cs = u[3 * p] * u[3 * q] + u[3 * p + 1] * u[3 * q + 1] + u[3 * p + 2] * u[3 * q + 2];
      if (cs < c->grids[c->pot.fn_off[4] + gs].xcut()) ++cnt;
    }
  return cnt;
}

int make_chunks(mz_ctx* c) {
  c->chunks.clear();
  const int maxa = std::max(1, std::min(c->tile.max_atoms, 64));
  const int maxp = c->tile.max_pairs;
  for (int cc = 0; cc < c->ncells; ++cc) {
    int a = c->abase[cc];
    const int aend = a + c->natoms_c[cc];
    while (a < aend) {
      mz::Chunk ch{};
      ch.cell = cc;
      ch.a0 = a;
      ch.p0 = c->atom_poff[a];
      int na = 0;
      while (a < aend && na < maxa) {
        const int np_a = c->atom_poff[a + 1] - c->atom_poff[a];
        if (np_a > maxp) return fail(MZ_ERR_CAPACITY, "atom " + std::to_string(a) + " has " + std::to_string(np_a) +
                                                          " neighbours, more than a chunk holds");
        if (c->atom_poff[a + 1] - ch.p0 > maxp) break;
        ++a;
        ++na;
      }
      ch.na = na;
      ch.np = c->atom_poff[a] - ch.p0;
      c->chunks.push_back(ch);
    }
  }
  // longest work first, so the hardware dispatcher balances the tail
  std::vector<long long> work(c->chunks.size());
  for (size_t q = 0; q < c->chunks.size(); ++q) {
    long long w = c->chunks[q].np;
    for (int a = c->chunks[q].a0; a < c->chunks[q].a0 + c->chunks[q].na; ++a) w += 2 * c->ntrip_atom[a];
    work[q] = w;
  }
  std::vector<size_t> order(c->chunks.size());
  for (size_t q = 0; q < order.size(); ++q) order[q] = q;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work[x] > work[y]; });
  std::vector<mz::Chunk> sorted(order.size());
  for (size_t q = 0; q < order.size(); ++q) sorted[q] = c->chunks[order[q]];
  c->chunks.swap(sorted);
  c->cell_chunk_off.assign(c->ncells + 1, 0);
  for (const auto& ch : c->chunks) c->cell_chunk_off[ch.cell + 1]++;
  for (int k = 0; k < c->ncells; ++k) c->cell_chunk_off[k + 1] += c->cell_chunk_off[k];
  c->cell_chunks.assign(c->chunks.size(), 0);
  {
    std::vector<int> fill(c->cell_chunk_off.begin(), c->cell_chunk_off.end() - 1);
    for (size_t q = 0; q < c->chunks.size(); ++q) c->cell_chunks[fill[c->chunks[q].cell]++] = (int)q;
  }
  if (c->d_chunks.upload(c->chunks.data(), c->chunks.size()) ||
      c->d_cell_chunk_off.upload(c->cell_chunk_off.data(), c->cell_chunk_off.size()) ||
      c->d_cell_chunks.upload(c->cell_chunks.data(), c->cell_chunks.size()))
    return fail(MZ_ERR_NOMEM, "chunk upload failed");
  if (c->ncap > 0 && c->d_chunk_acc.alloc((size_t)c->ncap * c->chunks.size() * 8))
    return fail(MZ_ERR_NOMEM, "chunk accumulator allocation failed");
  return MZ_OK;
}

void drop_graphs(mz_ctx* c) {
  for (auto& g : c->graphs)
    if (g.ex) (void)hipGraphExecDestroy(g.ex);
  c->graphs.clear();
  for (auto& g : c->adjx.graphs)
    if (g.ex) (void)hipGraphExecDestroy(g.ex);
  c->adjx.graphs.clear();
}

int ensure_eval_buffers(mz_ctx* c, int ncand) {
  if (ncand <= c->ncap) return MZ_OK;
  drop_graphs(c);
  const size_t nc = (size_t)ncand;
  c->resid_blocks = mz::resid_blocks(c->natoms, c->ncells);
  if (c->d_knots.alloc(nc * c->nknot) || c->d_tables.alloc(nc * c->table_bytes) ||
      c->d_forces.alloc(nc * c->natoms * 3) || c->d_tvec.alloc(nc * c->npairs * 3) ||
      c->d_fown.alloc(nc * c->natoms * 3) || c->d_chunk_acc.alloc(nc * c->chunks.size() * 8) ||
      c->d_partial.alloc(nc * c->resid_blocks) || c->d_F.alloc(nc))
    return fail(MZ_ERR_NOMEM, "device allocation of evaluation buffers failed");
  c->ncap = ncand;
  const size_t need = nc * std::max<size_t>(c->nknot, 1);
  if (need > c->h_stage_n) {
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    c->h_stage = nullptr;
    if (hipHostMalloc((void**)&c->h_stage, need * sizeof(double), hipHostMallocDefault) != hipSuccess)
      return fail(MZ_ERR_NOMEM, "pinned staging allocation failed");
    c->h_stage_n = need;
  }
  if (c->h_F) (void)hipHostFree(c->h_F);
  c->h_F = nullptr;
  if (hipHostMalloc((void**)&c->h_F, nc * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return fail(MZ_ERR_NOMEM, "pinned result allocation failed");
  return MZ_OK;
}

mz::RefreshMeta refresh_meta(const mz_ctx* c) {
  mz::RefreshMeta m;
  m.nspl = c->nspl;
  m.nknots = c->d_nknots.p;
  m.basis = c->d_basis.p;
  m.koff = c->d_koff.p;
  m.x0 = c->d_x0.p;
  m.xn = c->d_xn.p;
  m.bc = c->d_bc.p;
  m.coef_off = c->d_coef_off.p;
  return m;
}

// The context's own per-evaluation buffers.
mz::EvalBufs ctx_bufs(const mz_ctx* c) {
  mz::EvalBufs b;
  b.knots = c->d_knots.p;
  b.tables = c->d_tables.p;
  b.forces = c->d_forces.p;
  b.tvec = c->d_tvec.p;
  b.fown = c->d_fown.p;
  b.chunk_acc = c->d_chunk_acc.p;
  b.partial = c->d_partial.p;
  b.F = c->d_F.p;
  b.nblocks = c->resid_blocks;
  return b;
}

mz::AtomArgs atom_args(const mz_ctx* c, const mz::EvalBufs& b) {
  mz::AtomArgs a{};
  a.pot = c->pot;
  a.chunks = c->d_chunks.p;
  a.atom_poff = c->d_atom_poff.p;
  a.pairs = c->d_pairs.p;
  a.pos = c->d_pos.p;
  a.box = c->d_box.p;
  a.types = c->d_types.p;
  a.weight = c->d_weight.p;
  a.tables = b.tables;
  a.table_bytes = c->table_bytes;
  a.natoms = c->natoms;
  a.ncells = c->ncells;
  a.npairs = c->npairs;
  a.nchunks = (int)c->chunks.size();
  a.in_pos = c->d_in_pos.p;
  a.tvec = b.tvec;
  a.fown = b.fown;
  a.chunk_acc = b.chunk_acc;
  return a;
}

mz::ResidArgs resid_args(const mz_ctx* c, const mz::EvalBufs& b) {
  mz::ResidArgs r{};
  r.natoms = c->natoms;
  r.ncells = c->ncells;
  r.nerr = c->nerr;
  r.cell_of = c->d_cell_of.p;
  r.natoms_c = c->d_natoms_c.p;
  r.abase = c->d_abase.p;
  r.weight = c->d_weight.p;
  r.eweight = c->d_eweight.p;
  r.sweight = c->d_sweight.p;
  r.vol = c->d_vol.p;
  r.energy0 = c->d_energy0.p;
  r.stress0 = c->d_stress0.p;
  r.force0 = c->d_force0.p;
  r.energy_weight = c->energy_weight;
  r.stress_weight = c->stress_weight;
  r.npairs = c->npairs;
  r.in_off = c->d_in_off.p;
  r.tvec = b.tvec;
  r.fown = b.fown;
  r.forces = b.forces;
  r.chunk_acc = b.chunk_acc;
  r.cell_chunk_off = c->d_cell_chunk_off.p;
  r.cell_chunks = c->d_cell_chunks.p;
  r.nchunks = (int)c->chunks.size();
  r.partial = b.partial;
  r.nodens = c->pot.nodens;
  return r;
}

mz::FinalArgs final_args(const mz_ctx* c, const mz::EvalBufs& b) {
  mz::FinalArgs f{};
  f.npartial = b.nblocks;
  f.partial = b.partial;
  f.meam = c->meam;
  f.quintic = c->pot.quintic;
  f.ncells = c->ncells;
  f.natoms = c->natoms;
  f.nerr = c->nerr;
  f.nt = c->nt;
  f.nspl = c->nspl;
  f.nknot = c->nknot;
  f.f0 = c->pot.fn_off[3];
  f.nf = c->nt * (c->nt + 1) / 2;
  f.fper = c->meam && c->nfpen == f.nf && (c->style == mz::kGMeamSpline || c->style == mz::kGMeam);
  f.u0 = c->pot.fn_off[2];
  f.nodens = c->pot.nodens;
  f.npgrp = c->meam && !f.fper ? 1 + c->pot.njm : 0;
  for (int g = 0; g < 3; ++g) f.pgrp0[g] = c->pot.fn_off[g == 0 ? 3 : 3 + 2 * g];
  f.koff = c->d_koff.p;
  f.nknots = c->d_nknots.p;
  f.knots = b.knots;
  f.tables = b.tables;
  f.table_bytes = c->table_bytes;
  f.F = b.F;
  return f;
}

}  // namespace

int mz_noalloc_setup(int dev);

// Launch sequence of one batch of ncand evaluations whose knot vectors are
// already in b.knots, on the per-evaluation buffers b.  Every buffer of b is
// written before it is read: nothing needs clearing between calls.
int mz_run_forward_on(mz_ctx* c, const mz::EvalBufs& b, int ncand, double* d_errvec, double* d_ce, double* d_cs,
                      double* d_rho, double* d_dF, double* d_G, double* d_lam, double* d_cadj, double* d_T1) {
  hipStream_t st = c->stream;
  if (c->timing) MZ_HIP(hipEventRecord(c->ev[0], st));
  mz::launch_refresh(st, b.knots, c->nknot, refresh_meta(c), b.tables, c->table_bytes, ncand);
  if (c->timing) MZ_HIP(hipEventRecord(c->ev[1], st));
  if (c->timing) MZ_HIP(hipEventRecord(c->ev[2], st));
  mz::AtomArgs aa = atom_args(c, b);
  aa.rho_out = d_rho;
  aa.dF_out = d_dF;
  aa.G_out = d_G;
  aa.T1_out = d_T1;
  aa.ncand = ncand;
  const int ppl = std::max(1, c->tile.max_pairs / c->tile.block);
  // candidates per workgroup: the tile's choice, at most what ncand fills and
  // what the LDS holds
  int ct = std::min(c->tile.cand_tile, ncand >= 4 ? 4 : (ncand >= 2 ? 2 : 1));
  if (c->pot.njm) ct = 1;  // the jmeam kernels take one candidate per workgroup
  while (ct > 1 && mz::atoms_static_lds(c->tile.block, ppl, ct, c->pot) + mz::atoms_dyn_lds(c->pot, ct) > 64 * 1024) ct >>= 1;
  if (mz::launch_atoms(st, aa, (int)c->chunks.size(), ncand, c->tile.block, ppl, ct))
    return fail(MZ_ERR_ARG, "unsupported tile (block, pairs per lane)");
  MZ_HIP(hipGetLastError());
  if (c->timing) MZ_HIP(hipEventRecord(c->ev[7], st));
  mz::ResidArgs ra = resid_args(c, b);
  ra.errvec = d_errvec;
  ra.cell_energy = d_ce;
  ra.cell_stress = d_cs;
  ra.lam = d_lam;
  ra.cell_adj = d_cadj;
  mz::launch_resid(st, ra, b.nblocks, ncand);
  mz::FinalArgs fa = final_args(c, b);
  fa.errvec = d_errvec;
  mz::launch_final(st, fa, ncand);
  MZ_HIP(hipGetLastError());
  if (c->timing) MZ_HIP(hipEventRecord(c->ev[4], st));
  return MZ_OK;
}

// The same sequence on the context's own buffers.  Shared by the plain,
// gradient and batched entry points.
int mz_run_forward(mz_ctx* c, int ncand, double* d_errvec, double* d_ce, double* d_cs, double* d_rho, double* d_dF,
                   double* d_G, double* d_lam, double* d_cadj, double* d_T1) {
  return mz_run_forward_on(c, ctx_bufs(c), ncand, d_errvec, d_ce, d_cs, d_rho, d_dF, d_G, d_lam, d_cadj, d_T1);
}

int mz_stage_knots(mz_ctx* c, int ncand, const double* knots) {
  const size_t n = (size_t)ncand * c->nknot;
  for (size_t q = 0; q < n; ++q)
    if (!std::isfinite(knots[q])) return fail(MZ_ERR_ARG, "knot vector holds a non-finite value");
  std::memcpy(c->h_stage, knots, n * sizeof(double));
  MZ_HIP(hipMemcpyAsync(c->d_knots.p, c->h_stage, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return MZ_OK;
}

// Adds the kernel times of the launch sequence that has just drained (events
// recorded with timing on) to acc; `adjoint`: the sequence ran the reverse
// sweep too (events 5-6 around k_adjoint, 3 at the end of the chain).
void mz_timing_add(mz_ctx* c, mz_timing* acc, bool adjoint) {
  if (!c->timing) return;
  float t01 = 0, t27 = 0, t74 = 0, tend = 0, t56 = 0;
  (void)hipEventElapsedTime(&t01, c->ev[0], c->ev[1]);
  (void)hipEventElapsedTime(&t27, c->ev[2], c->ev[7]);
  (void)hipEventElapsedTime(&t74, c->ev[7], c->ev[4]);
  (void)hipEventElapsedTime(&tend, c->ev[0], c->ev[adjoint ? 3 : 4]);
  if (adjoint) (void)hipEventElapsedTime(&t56, c->ev[5], c->ev[6]);
  acc->refresh_ms += t01;
  acc->atom_ms += t27;
  acc->resid_ms += t74;
  acc->total_ms += tend;
  acc->adjoint_ms += t56;
  acc->launches += 1;
}

// c->last from the one launch sequence that has just drained.
void mz_record_timing(mz_ctx* c) {
  if (!c->timing) return;
  c->last = mz_timing{};
  mz_timing_add(c, &c->last, false);
}

// Equal sub-batches of at most `fit` candidates covering ncand.
int mz_sub_batch(int ncand, size_t fit) {
  const int cap = (int)std::max<size_t>(1, std::min<size_t>((size_t)ncand, fit));
  const int parts = (ncand + cap - 1) / cap;
  return (ncand + parts - 1) / parts;
}

// Forward objective of ncand knot vectors through a captured graph:
// H2D copy of the staged knots, refresh, clears, atom kernel, residual and
// final reduction, D2H copy of F.  One graph per candidate count; graphs are
// dropped whenever buffers or chunks change.  With timing on, the same
// sequence is launched on the stream with its events (event timestamps are
// not kept by graph replays) and the kernel times are added to acc.
int mz_forward_graph(mz_ctx* c, int ncand, const double* knots, double* F, mz_timing* acc) {
  const size_t n = (size_t)ncand * c->nknot;
  for (size_t q = 0; q < n; ++q)
    if (!std::isfinite(knots[q])) return fail(MZ_ERR_ARG, "knot vector holds a non-finite value");
  std::memcpy(c->h_stage, knots, n * sizeof(double));
  if (c->timing) {
    MZ_HIP(hipMemcpyAsync(c->d_knots.p, c->h_stage, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    const int rc = mz_run_forward(c, ncand, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                  nullptr);
    if (rc) return rc;
    MZ_HIP(hipMemcpyAsync(c->h_F, c->d_F.p, ncand * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MZ_HIP(hipStreamSynchronize(c->stream));
    if (acc) mz_timing_add(c, acc, false);
    std::memcpy(F, c->h_F, ncand * sizeof(double));
    return MZ_OK;
  }
  hipGraphExec_t ex = nullptr;
  for (auto& g : c->graphs)
    if (g.ncand == ncand) ex = g.ex;
  if (!ex) {
    hipGraph_t graph = nullptr;
    MZ_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e1 = hipMemcpyAsync(c->d_knots.p, c->h_stage, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    int rc = mz_run_forward(c, ncand, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    hipError_t e2 = hipMemcpyAsync(c->h_F, c->d_F.p, ncand * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    hipError_t e3 = hipStreamEndCapture(c->stream, &graph);
    if (rc) return rc;
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
      return fail(MZ_ERR_HIP, "graph capture of the forward evaluation failed");
    hipError_t e4 = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e4 != hipSuccess) return fail(MZ_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e4));
    mz_ctx::Graph g;
    g.ex = ex;
    g.ncand = ncand;
    c->graphs.push_back(g);
  }
  MZ_HIP(hipGraphLaunch(ex, c->stream));
  MZ_HIP(hipStreamSynchronize(c->stream));
  std::memcpy(F, c->h_F, ncand * sizeof(double));
  return MZ_OK;
}

// Accessors used by the adjoint and batch translation units.
mz::AtomArgs mz_atom_args(const mz_ctx* c) { return atom_args(c, ctx_bufs(c)); }

extern "C" {

const char* mz_last_error(void) { return g_err.c_str(); }

static int create_impl(const mz_layout* layout, const mz_db* db, bool device_lists, mz_ctx** out) {
  if (!out || !layout || !db) return fail(MZ_ERR_ARG, "null argument");
  *out = nullptr;
  if (db->ncells < 1) return fail(MZ_ERR_ARG, "database has no cells");
  mz_ctx* c = new mz_ctx();
  int rc = build_layout(c, layout);
  if (rc) {
    delete c;
    return rc;
  }
  c->ncells = db->ncells;
  c->natoms_c.assign(db->natoms, db->natoms + db->ncells);
  c->abase.resize(c->ncells);
  int na = 0;
  for (int k = 0; k < c->ncells; ++k) {
    if (c->natoms_c[k] < 1) {
      delete c;
      return fail(MZ_ERR_ARG, "cell " + std::to_string(k) + " has no atoms");
    }
    c->abase[k] = na;
    na += c->natoms_c[k];
  }
  c->natoms = na;
  // pair styles add no constraint or penalty rows (their extra_error is empty,
  // src/pots/pot_pair/pot_pair.cpp:136-137)
  c->nerr = 7 * c->ncells + 3 * na + (c->pair ? 0 : c->ncells + (c->meam ? c->nfpen : c->nt));
  c->types.assign(db->types, db->types + na);
  c->cell_of.resize(na);
  for (int k = 0; k < c->ncells; ++k)
    for (int a = 0; a < c->natoms_c[k]; ++a) c->cell_of[c->abase[k] + a] = k;
  for (int a = 0; a < na; ++a)
    if (c->types[a] < 0 || c->types[a] >= c->nt) {
      delete c;
      return fail(MZ_ERR_ARG, "incorrect atom type specified - atom=" + std::to_string(a));
    }
  c->vol.resize(c->ncells);
  for (int k = 0; k < c->ncells; ++k) {
    const double* A = db->box + 9 * k;
    const double c12[3] = {A[4] * A[8] - A[5] * A[7], A[5] * A[6] - A[3] * A[8], A[3] * A[7] - A[4] * A[6]};
    c->vol[k] = std::abs(A[0] * c12[0] + A[1] * c12[1] + A[2] * c12[2]);  // Cell::compute_vol src/cell.cpp:176-179
    if (!(c->vol[k] > 0)) {
      delete c;
      return fail(MZ_ERR_GEOMETRY, "cell " + std::to_string(k) + " has zero volume");
    }
  }
  // max radial cutoff (src/potential.cpp:73-82)
  double max_rcut = 0.0;
  const int radial[5] = {0, 1, 3, 5, 7};
  const int nt = c->nt, npij = nt * (nt + 1) / 2;
  const int counts[9] = {npij, nt, nt, npij, nt, npij, nt, npij, nt};
  for (int q = 0; q < (c->meam ? 3 + c->pot.njm : (c->pair ? 1 : 2)); ++q)
    for (int s = c->pot.fn_off[radial[q]]; s < c->pot.fn_off[radial[q]] + counts[radial[q]]; ++s)
      max_rcut = std::max(max_rcut, c->xn[s]);

  // the HIP device and stream first: the device list builder runs on them
  if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return fail(MZ_ERR_HIP, "no HIP device available");
  }
  // device address of the allocation-free entry point's scratch array, once per device
  rc = mz_noalloc_setup(c->device);
  if (rc) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  const auto t_list = std::chrono::steady_clock::now();
  // the triplet counts read the list on the host unless every triplet of an
  // atom is kept (knot-style g spanning |cos| <= 1)
  const bool host_pairs = c->meam && (c->point || !c->pot.g_keep_all || c->pot.njm > 0);
  if (device_lists) {
    std::string e;
    rc = mz::build_pairs_device(c, db->pos, db->box, max_rcut, host_pairs, e);
    if (!rc) rc = mz::build_incoming_device(c, e);
    if (rc) {
      mz_destroy(c);
      return fail(rc, e);
    }
  } else {
    // pair lists, cells spread over host threads
    std::vector<CellPairs> cp(c->ncells);
    {
      unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
      std::atomic<int> next{0};
      std::vector<std::thread> th;
      for (unsigned t = 0; t < nth; ++t)
        th.emplace_back([&]() {
          for (int k; (k = next.fetch_add(1)) < c->ncells;)
            cell_pairs(c, k, db->box + 9 * k, db->pos, max_rcut, c->vol[k], cp[k]);
        });
      for (auto& t : th) t.join();
    }
    c->atom_poff.assign(na + 1, 0);
    long long np = 0;
    for (int k = 0; k < c->ncells; ++k) {
      if (!cp[k].err.empty()) {
        std::string e = cp[k].err;
        mz_destroy(c);
        return fail(MZ_ERR_GEOMETRY, e);
      }
      for (int a = 0; a < c->natoms_c[k]; ++a) {
        c->atom_poff[c->abase[k] + a] = (int)np;
        np += (long long)cp[k].per_atom[a].size();
      }
    }
    if (np > 2000000000LL) {
      mz_destroy(c);
      return fail(MZ_ERR_CAPACITY, "more than 2e9 pairs");
    }
    c->atom_poff[na] = (int)np;
    c->npairs = np;
    c->pairs.resize(np);
    for (int k = 0; k < c->ncells; ++k)
      for (int a = 0; a < c->natoms_c[k]; ++a) {
        auto& v = cp[k].per_atom[a];
        std::copy(v.begin(), v.end(), c->pairs.begin() + c->atom_poff[c->abase[k] + a]);
        std::vector<uint2>().swap(v);
      }
  }
  c->list_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_list).count();
  c->list_device = device_lists ? 1 : 0;
  const long long np = c->npairs;
  c->ntrip_atom.assign(na, 0);
  {
    unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nth; ++t)
      th.emplace_back([&]() {
        for (int a; (a = next.fetch_add(1)) < na;)
          c->ntrip_atom[a] = count_triplets(c, a, db->pos, db->box + 9 * c->cell_of[a]);
      });
    for (auto& t : th) t.join();
  }
  c->ntriplets = 0;
  for (int a = 0; a < na; ++a) c->ntriplets += c->ntrip_atom[a];
  // incoming pairs of every atom (counting sort of the pairs by neighbour):
  // the force on atom j from pair (i -> j) is -T of that pair; the device
  // builder made them already
  if (!device_lists) {
    c->in_off.assign(na + 1, 0);
    for (long long q = 0; q < np; ++q) c->in_off[c->pairs[q].x + 1]++;
    for (int a = 0; a < na; ++a) c->in_off[a + 1] += c->in_off[a];
    c->in_list.assign(np, 0);
    std::vector<int> fill(c->in_off.begin(), c->in_off.end() - 1);
    for (long long q = 0; q < np; ++q) c->in_list[fill[c->pairs[q].x]++] = (int)q;
    c->in_pos.assign(np, 0);
    for (long long q = 0; q < np; ++q) c->in_pos[c->in_list[q]] = (int)q;
  }

  // device upload (the device builder already placed the pair list and offsets)
  c->energy_weight = db->energy_weight;
  c->stress_weight = db->stress_weight;
  std::vector<double> s0(db->stress0, db->stress0 + 6 * (size_t)c->ncells);
  bool bad = (!device_lists && (c->d_atom_poff.upload(c->atom_poff.data(), c->atom_poff.size()) ||
                                 c->d_pairs.upload(c->pairs.data(), c->pairs.size()) ||
                                 c->d_in_off.upload(c->in_off.data(), c->in_off.size()) ||
                                 c->d_in_list.upload(c->in_list.data(), c->in_list.size()) ||
                                 c->d_in_pos.upload(c->in_pos.data(), c->in_pos.size()))) ||
             c->d_pos.upload(db->pos, 3 * (size_t)na) || c->d_box.upload(db->box, 9 * (size_t)c->ncells) ||
             c->d_types.upload(c->types.data(), na) || c->d_cell_of.upload(c->cell_of.data(), na) ||
             c->d_natoms_c.upload(c->natoms_c.data(), c->ncells) || c->d_abase.upload(c->abase.data(), c->ncells) ||
             c->d_weight.upload(db->weight, c->ncells) || c->d_eweight.upload(db->eweight, c->ncells) ||
             c->d_sweight.upload(db->sweight, c->ncells) || c->d_vol.upload(c->vol.data(), c->ncells) ||
             c->d_energy0.upload(db->energy0, c->ncells) || c->d_stress0.upload(s0.data(), s0.size()) ||
             c->d_force0.upload(db->force0, 3 * (size_t)na) ||
             c->d_nknots.upload(c->nknots.data(), c->nspl) || c->d_basis.upload(c->basis.data(), c->nspl) ||
             c->d_koff.upload(c->koff.data(), c->nspl) || c->d_bc.upload(c->bc.data(), 2 * c->nspl) ||
             c->d_coef_off.upload(c->coef_off.data(), c->nspl) || c->d_x0.upload(c->x0.data(), c->nspl) ||
             c->d_xn.upload(c->xn.data(), c->nspl);
  if (bad) {
    mz_destroy(c);
    return fail(MZ_ERR_NOMEM, "device upload failed");
  }
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) {
      mz_destroy(c);
      return fail(MZ_ERR_HIP, "event creation failed");
    }
  rc = make_chunks(c);
  if (!rc) rc = ensure_eval_buffers(c, 1);
  if (rc) {
    mz_destroy(c);
    return rc;
  }
  *out = c;
  return MZ_OK;
}

int mz_create(const mz_layout* layout, const mz_db* db, mz_ctx** out) { return create_impl(layout, db, true, out); }

int mz_create_host_lists(const mz_layout* layout, const mz_db* db, mz_ctx** out) {
  return create_impl(layout, db, false, out);
}

void mz_destroy(mz_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  mz_adjoint_drop(c);
  DevBuf<double>* dbufs[] = {&c->d_pos, &c->d_box, &c->d_weight, &c->d_eweight, &c->d_sweight, &c->d_vol,
                             &c->d_energy0, &c->d_stress0, &c->d_force0, &c->d_x0, &c->d_xn, &c->d_knots,
                             &c->d_forces, &c->d_tvec, &c->d_fown, &c->d_chunk_acc, &c->d_partial, &c->d_F, &c->d_errvec, &c->d_cell_e,
                             &c->d_cell_s, &c->d_rho, &c->d_dF, &c->d_G};
  for (auto* b : dbufs) b->release();
  DevBuf<int>* ibufs[] = {&c->d_in_off, &c->d_in_list, &c->d_in_pos, &c->d_cell_chunk_off, &c->d_cell_chunks, &c->d_atom_poff, &c->d_types, &c->d_cell_of, &c->d_natoms_c, &c->d_abase,
                          &c->d_nknots, &c->d_basis, &c->d_koff, &c->d_bc, &c->d_coef_off};
  for (auto* b : ibufs) b->release();
  c->d_pairs.release();
  c->d_chunks.release();
  c->d_tables.release();
  drop_graphs(c);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  if (c->h_F) (void)hipHostFree(c->h_F);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int mz_info(const mz_ctx* c, mz_info_t* out) {
  if (!c || !out) return fail(MZ_ERR_ARG, "null argument");
  out->ncells = c->ncells;
  out->natoms = c->natoms;
  out->nspl = c->nspl;
  out->nknot = c->nknot;
  out->nerr = c->nerr;
  out->npairs = c->npairs;
  out->ntriplets = c->ntriplets;
  out->nchunks = (int)c->chunks.size();
  out->device = c->device;
  out->list_ms = c->list_ms;
  out->list_device = c->list_device;
  return MZ_OK;
}

int mz_objective(mz_ctx* c, const double* knots, double* F, double* errvec, double* cell_energy, double* cell_stress,
                 double* forces) {
  if (!c || !knots) return fail(MZ_ERR_ARG, "null argument");
  int rc = ensure_eval_buffers(c, 1);
  if (rc) return rc;
  if (!errvec && !cell_energy && !cell_stress && !forces) {
    double Fh = 0.0;
    mz_timing acc{};
    rc = mz_forward_graph(c, 1, knots, &Fh, &acc);
    if (!rc && F) *F = Fh;
    if (!rc && c->timing) c->last = acc;
    return rc;
  }
  if (errvec && c->d_errvec.alloc(c->nerr)) return fail(MZ_ERR_NOMEM, "errvec allocation failed");
  if (cell_energy && c->d_cell_e.alloc(c->ncells)) return fail(MZ_ERR_NOMEM, "allocation failed");
  if (cell_stress && c->d_cell_s.alloc(6 * (size_t)c->ncells)) return fail(MZ_ERR_NOMEM, "allocation failed");
  rc = mz_stage_knots(c, 1, knots);
  if (rc) return rc;
  rc = mz_run_forward(c, 1, errvec ? c->d_errvec.p : nullptr, cell_energy ? c->d_cell_e.p : nullptr,
                      cell_stress ? c->d_cell_s.p : nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
  if (rc) return rc;
  hipStream_t st = c->stream;
  double Fh = 0.0;
  MZ_HIP(hipMemcpyAsync(&Fh, c->d_F.p, sizeof(double), hipMemcpyDeviceToHost, st));
  if (errvec) MZ_HIP(hipMemcpyAsync(errvec, c->d_errvec.p, sizeof(double) * c->nerr, hipMemcpyDeviceToHost, st));
  if (cell_energy)
    MZ_HIP(hipMemcpyAsync(cell_energy, c->d_cell_e.p, sizeof(double) * c->ncells, hipMemcpyDeviceToHost, st));
  if (cell_stress)
    MZ_HIP(hipMemcpyAsync(cell_stress, c->d_cell_s.p, sizeof(double) * 6 * c->ncells, hipMemcpyDeviceToHost, st));
  if (forces)
    MZ_HIP(hipMemcpyAsync(forces, c->d_forces.p, sizeof(double) * 3 * c->natoms, hipMemcpyDeviceToHost, st));
  MZ_HIP(hipStreamSynchronize(st));
  mz_record_timing(c);
  if (F) *F = Fh;
  return MZ_OK;
}

int mz_densities(mz_ctx* c, const double* knots, double* dens) {
  if (!c || !knots || !dens) return fail(MZ_ERR_ARG, "null argument");
  if (c->pair) return fail(MZ_ERR_LAYOUT, "pair styles have no densities");
  int rc = ensure_eval_buffers(c, 1);
  if (rc) return rc;
  if (c->d_rho.alloc(c->natoms)) return fail(MZ_ERR_NOMEM, "allocation failed");
  rc = mz_stage_knots(c, 1, knots);
  if (rc) return rc;
  rc = mz_run_forward(c, 1, nullptr, nullptr, nullptr, c->d_rho.p, nullptr, nullptr, nullptr, nullptr, nullptr);
  if (rc) return rc;
  MZ_HIP(hipMemcpyAsync(dens, c->d_rho.p, sizeof(double) * c->natoms, hipMemcpyDeviceToHost, c->stream));
  MZ_HIP(hipStreamSynchronize(c->stream));
  return MZ_OK;
}

int mz_set_tile(mz_ctx* c, const mz_tile* t) {
  if (!c || !t) return fail(MZ_ERR_ARG, "null argument");
  if (t->max_atoms < 1 || t->max_atoms > 64) return fail(MZ_ERR_ARG, "max_atoms must lie in [1, 64]");
  const int ppl = t->block > 0 ? t->max_pairs / t->block : 0;
  const bool ok = (t->block == 256 && (ppl == 1 || ppl == 2)) || (t->block == 128 && (ppl == 2 || ppl == 4)) ||
                  (t->block == 512 && ppl == 1) || (t->block == 64 && ppl == 4);
  if (!ok || ppl * t->block != t->max_pairs) return fail(MZ_ERR_ARG, "unsupported tile (block, max_pairs)");
  if (t->cand_tile != 0 && t->cand_tile != 1 && t->cand_tile != 2 && t->cand_tile != 4)
    return fail(MZ_ERR_ARG, "cand_tile must be 1, 2 or 4");
  if (mz::atoms_static_lds(t->block, ppl, 1, c->pot) + mz::atoms_dyn_lds(c->pot, 1) > 64 * 1024)
    return fail(MZ_ERR_CAPACITY, "tile and spline tables exceed 64 KB of LDS");
  mz_tile old = c->tile;
  drop_graphs(c);
  c->tile = *t;
  if (c->tile.cand_tile < 1) c->tile.cand_tile = 1;
  int rc = make_chunks(c);
  if (rc) {
    c->tile = old;
    (void)make_chunks(c);
  }
  return rc;
}

int mz_get_tile(const mz_ctx* c, mz_tile* t) {
  if (!c || !t) return fail(MZ_ERR_ARG, "null argument");
  *t = c->tile;
  return MZ_OK;
}

int mz_enable_timing(mz_ctx* c, int on) {
  if (!c) return fail(MZ_ERR_ARG, "null argument");
  c->timing = on != 0;
  return MZ_OK;
}

int mz_last_timing(const mz_ctx* c, mz_timing* out) {
  if (!c || !out) return fail(MZ_ERR_ARG, "null argument");
  *out = c->last;
  return MZ_OK;
}

}  // extern "C"

int mz_ensure_buffers(mz_ctx* c, int ncand) { return ensure_eval_buffers(c, ncand); }

// Device bytes that the per-candidate buffers of one launch may take: what is
// free now plus what the context's per-candidate buffers already hold (a
// larger batch reallocates them), less a reserve of 1/16 of the device (at
// least 2 GiB), and at most three quarters of the device.
size_t mz_candidate_budget(const mz_ctx* c) {
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) != hipSuccess || totalb == 0) return size_t(16) << 30;
  const mz::AdjointExtra& X = c->adjx;
  const size_t held = sizeof(double) * (c->d_knots.n + c->d_forces.n + c->d_tvec.n + c->d_fown.n + c->d_chunk_acc.n +
                                        c->d_partial.n + c->d_F.n + c->d_errvec.n + X.d_rho.n + X.d_dF.n + X.d_G.n +
                                        X.d_lam.n + X.d_cadj.n + X.d_slots.n + X.d_tabsum.n + X.d_grad.n + X.d_T1.n) +
                      c->d_tables.n;
  const size_t reserve = std::max<size_t>(size_t(2) << 30, totalb / 16);
  const size_t avail = freeb + held > reserve ? freeb + held - reserve : 0;
  return std::max<size_t>(size_t(1) << 30, std::min(avail, totalb / 4 * 3));
}

// Per-candidate device bytes of the forward launch sequence.
size_t mz_forward_bytes(const mz_ctx* c) {
  return (size_t)c->npairs * 24 + (size_t)c->natoms * 48 + c->chunks.size() * 64 + (size_t)c->nknot * 8 +
         (size_t)c->table_bytes + (size_t)mz::resid_blocks(c->natoms, c->ncells) * 8 + 64;
}

size_t mz_candidate_budget(const mz_ctx* c);
size_t mz_forward_bytes(const mz_ctx* c);
int mz_forward_graph(mz_ctx* c, int ncand, const double* knots, double* F, mz_timing* acc);

// Batched scoring of ncand knot vectors in one launch sequence (forward only
// when G is NULL; the gradient path lives in mz_adjoint.hip).

int mz_batch_impl(mz_ctx* c, int ncand, const double* knots, double* F, double* G) {
  if (!c || !knots || !F || ncand < 1) return fail(MZ_ERR_ARG, "null argument or ncand < 1");
  if (G) return mz_grad_impl(c, ncand, knots, F, G);
  // candidates go through in equal sub-batches whose per-candidate buffers
  // fit the device memory that is free
  const int sub = mz_sub_batch(ncand, mz_candidate_budget(c) / mz_forward_bytes(c));
  int rc = ensure_eval_buffers(c, sub);
  if (rc) return rc;
  mz_timing acc{};
  for (int c0 = 0; c0 < ncand; c0 += sub) {
    const int nb = std::min(sub, ncand - c0);
    rc = mz_forward_graph(c, nb, knots + (size_t)c0 * c->nknot, F + c0, &acc);
    if (rc) return rc;
  }
  if (c->timing) c->last = acc;
  return MZ_OK;
}

// Residual vectors of ncand knot vectors: the forward launch sequence with the
// residual output enabled for every candidate, in sub-batches that keep the
// per-candidate buffers and residual vectors within the free device memory.
int mz_residuals_batch_impl(mz_ctx* c, int ncand, const double* knots, double* F, double* errvec) {
  if (!c || !knots || !F || !errvec || ncand < 1) return fail(MZ_ERR_ARG, "null argument or ncand < 1");
  const int sub = mz_sub_batch(ncand, mz_candidate_budget(c) / (mz_forward_bytes(c) + (size_t)c->nerr * 8));
  int rc = ensure_eval_buffers(c, sub);
  if (rc) return rc;
  if (c->d_errvec.alloc((size_t)sub * c->nerr)) return fail(MZ_ERR_NOMEM, "residual buffer allocation failed");
  hipStream_t st = c->stream;
  mz_timing acc{};
  for (int c0 = 0; c0 < ncand; c0 += sub) {
    const int nb = std::min(sub, ncand - c0);
    rc = mz_stage_knots(c, nb, knots + (size_t)c0 * c->nknot);
    if (rc) return rc;
    rc = mz_run_forward(c, nb, c->d_errvec.p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    MZ_HIP(hipMemcpyAsync(F + c0, c->d_F.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st));
    MZ_HIP(hipMemcpyAsync(errvec + (size_t)c0 * c->nerr, c->d_errvec.p, sizeof(double) * nb * c->nerr,
                          hipMemcpyDeviceToHost, st));
    MZ_HIP(hipStreamSynchronize(st));
    mz_timing_add(c, &acc, false);
  }
  if (c->timing) c->last = acc;
  return MZ_OK;
}

int mz_fail(int code, const std::string& msg) { return fail(code, msg); }
