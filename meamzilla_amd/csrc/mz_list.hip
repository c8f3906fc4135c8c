// Device-side pair-list build.
//
// The same lists as the host builder (cell_pairs in mz_host.cpp), built on the
// GPU in two passes over one thread per atom: a count pass, a host scan of the
// counts into the per-atom offsets, then a fill pass writing (neighbour,
// packed image) records straight into the resident pair array.  Both follow
// Config::find_pairs src/config.cpp:162-229: for atom i, every atom j of its
// cell and every periodic image (ix, iy, iz) in ascending order, the distance
// vector pj + ix a0 + iy a1 + iz a2 - pi evaluated left to right, kept if
// r < max_rcut and r lies inside one of the radial functions
// (check_pair: src/pots/pot_meam_spline/pair_meam_spline.cpp:50-99 and its
// siblings; r below a function's first knot is an error).  Every rounding is
// explicit, so the device and host lists agree pair for pair.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "mz_ctx.h"
#include "mz_list.h"

namespace mz {

namespace {

struct ListCell {
  double ih[3];  // |a_q x a_r| / V per direction
  double sgn;    // sign of the triple product
  int sc[3];     // image range per direction
  int n;         // atoms
  int abase;     // first atom
  int pad;
};

struct ListArgs {
  const double* pos;    // [natoms][3]
  const double* fr;     // [natoms][3] fractional coordinates
  const double* box;    // [ncells][9]
  const ListCell* cells;
  const int* types;
  const double* x0;     // per spline: first knot
  const double* xn;     // per spline: last knot
  int fn_off[9];
  int nrad;             // radial functions checked per pair (phi, rho, f, p, p2 prefix)
  int point;            // point-evaluated styles: inside means r <= x_n
  int nt;
  double max_rcut;
  const int* poff;      // fill pass: first pair of every atom
  uint2* pairs;         // fill pass: output
  int* counts;          // count pass: pairs per atom
  unsigned long long* err_key;  // smallest (atom << 32 | neighbour) with r below a first knot
  int cell0;            // first cell of this launch (grid.y is at most 65535)
};

__device__ __forceinline__ int pair_ij_d(int ti, int tj, int nt) {
  return (ti <= tj) ? ti * nt + tj - ((ti * (ti + 1)) / 2) : tj * nt + ti - ((tj * (tj + 1)) / 2);
}

template <bool FILL>
__global__ void __launch_bounds__(128) k_list(const ListArgs a) {
  const int cell = a.cell0 + (int)blockIdx.y;
  const ListCell C = a.cells[cell];
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= C.n) return;
  const int gi = C.abase + ia;
  const double* box = a.box + 9 * (size_t)cell;
  const double pi[3] = {a.pos[3 * (size_t)gi], a.pos[3 * (size_t)gi + 1], a.pos[3 * (size_t)gi + 2]};
  const double fi[3] = {a.fr[3 * (size_t)gi], a.fr[3 * (size_t)gi + 1], a.fr[3 * (size_t)gi + 2]};
  const int ti = a.types[gi];
  const double margin = 1e-7;
  const double xcut_scale = 1. + 1.e-8;
  int cnt = 0;
  long long out = FILL ? a.poff[gi] : 0;
  const long long end = FILL ? a.poff[gi + 1] : 0;  // the count pass found exactly this many
  for (int j = 0; j < C.n; ++j) {
    const int gj = C.abase + j;
    int lo[3], hi[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double ds = C.sgn * (a.fr[3 * (size_t)gj + q] - fi[q]);
      const double lim = a.max_rcut * C.ih[q] + margin;
      lo[q] = max(-C.sc[q], (int)ceil(-lim - ds));
      hi[q] = min(C.sc[q], (int)floor(lim - ds));
    }
    const double pj[3] = {a.pos[3 * (size_t)gj], a.pos[3 * (size_t)gj + 1], a.pos[3 * (size_t)gj + 2]};
    const int tj = a.types[gj];
    const int pij = pair_ij_d(ti, tj, a.nt);
    const int fns[5] = {a.fn_off[0] + pij, a.fn_off[1] + tj, a.fn_off[3] + pij, a.fn_off[5] + pij,
                        a.fn_off[7] + pij};
    for (int ix = lo[0]; ix <= hi[0]; ++ix)
      for (int iy = lo[1]; iy <= hi[1]; ++iy)
        for (int iz = lo[2]; iz <= hi[2]; ++iz) {
          if (ia == j && ix == 0 && iy == 0 && iz == 0) continue;
          double d[3];
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            double t = __dadd_rn(pj[m], __dmul_rn((double)ix, box[m]));
            t = __dadd_rn(t, __dmul_rn((double)iy, box[3 + m]));
            t = __dadd_rn(t, __dmul_rn((double)iz, box[6 + m]));
            d[m] = __dadd_rn(t, -pi[m]);
          }
          const double r =
              __dsqrt_rn(__dadd_rn(__dadd_rn(__dmul_rn(d[0], d[0]), __dmul_rn(d[1], d[1])), __dmul_rn(d[2], d[2])));
          if (!(r < a.max_rcut)) continue;
          bool inside = false, bad = false;
          for (int q = 0; q < a.nrad; ++q) {
            const int s = fns[q];
            bad |= r < a.x0[s];
            inside |= a.point ? (r <= a.xn[s]) : (r < __dmul_rn(a.xn[s], xcut_scale));
          }
          if (bad) {
            atomicMin(a.err_key, ((unsigned long long)gi << 32) | (unsigned)j);
            continue;
          }
          if (!inside) continue;
          if (FILL) {
            const unsigned im = (unsigned)(unsigned char)(signed char)ix | ((unsigned)(unsigned char)(signed char)iy << 8) |
                                ((unsigned)(unsigned char)(signed char)iz << 16);
            if (out < end) a.pairs[out] = make_uint2((unsigned)gj, im);
            ++out;
          } else {
            ++cnt;
          }
        }
  }
  if (!FILL) a.counts[gi] = cnt;
}

// Incoming-pair index (the pairs whose neighbour is each atom, ascending pair
// index within an atom, as the host's counting sort leaves them): count per
// neighbour, fill in any order, then sort each atom's short run.
__global__ void __launch_bounds__(256) k_in_count(const uint2* pairs, long long np, int* cnt) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < np) atomicAdd(cnt + pairs[q].x, 1);
}

__global__ void __launch_bounds__(256) k_in_fill(const uint2* pairs, long long np, const int* off, int* fill,
                                                 int* list) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  const int j = (int)pairs[q].x;
  list[off[j] + atomicAdd(fill + j, 1)] = (int)q;
}

// Inverse of the incoming-pair index: the slot of every pair in it.
__global__ void __launch_bounds__(256) k_in_pos(const int* list, long long np, int* pos) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < np) pos[list[q]] = (int)q;
}

__global__ void __launch_bounds__(256) k_in_sort(const int* off, int* list, int na) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= na) return;
  const int b = off[a], e = off[a + 1];
  for (int i = b + 1; i < e; ++i) {
    const int v = list[i];
    int k = i - 1;
    while (k >= b && list[k] > v) {
      list[k + 1] = list[k];
      --k;
    }
    list[k + 1] = v;
  }
}

#define MZ_LCHECK(x)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) {                                                             \
      err = std::string("pair-list build: ") + #x + ": " + hipGetErrorString(e_);       \
      return MZ_ERR_HIP;                                                                \
    }                                                                                   \
  } while (0)

}  // namespace

int build_incoming_device(mz_ctx* c, std::string& err) {
  const int na = c->natoms;
  const long long np = c->npairs;
  hipStream_t st = c->stream;
  struct Scratch {
    DevBuf<int> cnt;
    ~Scratch() { cnt.release(); }
  } tmp;
  if (tmp.cnt.alloc(na) || c->d_in_off.alloc((size_t)na + 1) || c->d_in_list.alloc((size_t)std::max(np, 1LL)) ||
      c->d_in_pos.alloc((size_t)std::max(np, 1LL))) {
    err = "incoming-pair index: device allocation failed";
    return MZ_ERR_NOMEM;
  }
  MZ_LCHECK(hipMemsetAsync(tmp.cnt.p, 0, sizeof(int) * (size_t)na, st));
  const unsigned nbq = (unsigned)((np + 255) / 256);
  if (np) hipLaunchKernelGGL(k_in_count, dim3(nbq), dim3(256), 0, st, c->d_pairs.p, np, tmp.cnt.p);
  MZ_LCHECK(hipGetLastError());
  std::vector<int> cnt(na);
  MZ_LCHECK(hipMemcpyAsync(cnt.data(), tmp.cnt.p, sizeof(int) * (size_t)na, hipMemcpyDeviceToHost, st));
  MZ_LCHECK(hipStreamSynchronize(st));
  c->in_off.assign((size_t)na + 1, 0);
  for (int a = 0; a < na; ++a) c->in_off[a + 1] = c->in_off[a] + cnt[a];
  MZ_LCHECK(hipMemcpyAsync(c->d_in_off.p, c->in_off.data(), sizeof(int) * ((size_t)na + 1), hipMemcpyHostToDevice, st));
  MZ_LCHECK(hipMemsetAsync(tmp.cnt.p, 0, sizeof(int) * (size_t)na, st));
  if (np) hipLaunchKernelGGL(k_in_fill, dim3(nbq), dim3(256), 0, st, c->d_pairs.p, np, c->d_in_off.p, tmp.cnt.p,
                             c->d_in_list.p);
  hipLaunchKernelGGL(k_in_sort, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, st, c->d_in_off.p, c->d_in_list.p,
                     na);
  if (np) hipLaunchKernelGGL(k_in_pos, dim3(nbq), dim3(256), 0, st, c->d_in_list.p, np, c->d_in_pos.p);
  MZ_LCHECK(hipGetLastError());
  MZ_LCHECK(hipStreamSynchronize(st));
  return MZ_OK;
}

int build_pairs_device(mz_ctx* c, const double* pos, const double* box, double max_rcut, bool host_copy,
                       std::string& err) {
  const int na = c->natoms, nc = c->ncells;
  hipStream_t st = c->stream;
  // per-cell image ranges and fractional coordinates, as the host builder forms them
  std::vector<ListCell> cells(nc);
  std::vector<double> fr(3 * (size_t)na);
  int maxn = 0;
  for (int k = 0; k < nc; ++k) {
    const double* a0 = box + 9 * (size_t)k;
    const double* a1 = a0 + 3;
    const double* a2 = a0 + 6;
    const double c12[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2], a1[0] * a2[1] - a1[1] * a2[0]};
    const double c20[3] = {a2[1] * a0[2] - a2[2] * a0[1], a2[2] * a0[0] - a2[0] * a0[2], a2[0] * a0[1] - a2[1] * a0[0]};
    const double c01[3] = {a0[1] * a1[2] - a0[2] * a1[1], a0[2] * a1[0] - a0[0] * a1[2], a0[0] * a1[1] - a0[1] * a1[0]};
    const double vol = c->vol[k];
    ListCell& L = cells[k];
    L.ih[0] = std::sqrt(c12[0] * c12[0] + c12[1] * c12[1] + c12[2] * c12[2]) / vol;
    L.ih[1] = std::sqrt(c20[0] * c20[0] + c20[1] * c20[1] + c20[2] * c20[2]) / vol;
    L.ih[2] = std::sqrt(c01[0] * c01[0] + c01[1] * c01[1] + c01[2] * c01[2]) / vol;
    for (int q = 0; q < 3; ++q) L.sc[q] = (int)std::ceil(max_rcut * L.ih[q]);
    if (L.sc[0] <= 0 || L.sc[1] <= 0 || L.sc[2] <= 0) {
      err = "Cell #" + std::to_string(k) + " is too small or potential is too big";
      return MZ_ERR_GEOMETRY;
    }
    if (L.sc[0] > kMaxImage || L.sc[1] > kMaxImage || L.sc[2] > kMaxImage) {
      err = "Cell #" + std::to_string(k) + " needs more than 127 periodic images";
      return MZ_ERR_GEOMETRY;
    }
    L.sgn = (a0[0] * c12[0] + a0[1] * c12[1] + a0[2] * c12[2]) < 0 ? -1.0 : 1.0;
    L.n = c->natoms_c[k];
    L.abase = c->abase[k];
    L.pad = 0;
    maxn = std::max(maxn, L.n);
    for (int a = 0; a < L.n; ++a) {
      const double* p = pos + 3 * (size_t)(L.abase + a);
      double* f = fr.data() + 3 * (size_t)(L.abase + a);
      f[0] = (p[0] * c12[0] + p[1] * c12[1] + p[2] * c12[2]) / vol;
      f[1] = (p[0] * c20[0] + p[1] * c20[1] + p[2] * c20[2]) / vol;
      f[2] = (p[0] * c01[0] + p[1] * c01[1] + p[2] * c01[2]) / vol;
    }
  }
  struct Scratch {  // released on every return
    DevBuf<ListCell> cells;
    DevBuf<double> fr;
    DevBuf<int> counts;
    DevBuf<unsigned long long> err;
    ~Scratch() {
      cells.release();
      fr.release();
      counts.release();
      err.release();
    }
  } tmp;
  DevBuf<ListCell>& d_cells = tmp.cells;
  DevBuf<double>& d_fr = tmp.fr;
  DevBuf<int>& d_counts = tmp.counts;
  DevBuf<unsigned long long>& d_err = tmp.err;
  if (d_cells.upload(cells.data(), nc) || d_fr.upload(fr.data(), fr.size()) || d_counts.alloc(na) ||
      d_err.alloc(1) || c->d_pos.upload(pos, 3 * (size_t)na) || c->d_box.upload(box, 9 * (size_t)nc) ||
      c->d_types.upload(c->types.data(), na) || c->d_x0.upload(c->x0.data(), c->nspl) ||
      c->d_xn.upload(c->xn.data(), c->nspl)) {
    err = "pair-list build: device allocation failed";
    return MZ_ERR_NOMEM;
  }
  const unsigned long long none = ~0ull;
  MZ_LCHECK(hipMemcpyAsync(d_err.p, &none, sizeof(none), hipMemcpyHostToDevice, st));
  ListArgs a{};
  a.pos = c->d_pos.p;
  a.fr = d_fr.p;
  a.box = c->d_box.p;
  a.cells = d_cells.p;
  a.types = c->d_types.p;
  a.x0 = c->d_x0.p;
  a.xn = c->d_xn.p;
  for (int f = 0; f < 9; ++f) a.fn_off[f] = c->pot.fn_off[f];
  a.nrad = c->meam ? 3 + c->pot.njm : (c->pair ? 1 : 2);
  a.point = c->point ? 1 : 0;
  a.nt = c->nt;
  a.max_rcut = max_rcut;
  a.counts = d_counts.p;
  a.err_key = d_err.p;
  const int ybatch = 65535;
  auto launch = [&](bool fill) {
    for (int c0 = 0; c0 < nc; c0 += ybatch) {
      a.cell0 = c0;
      const dim3 grid((unsigned)((maxn + 127) / 128), (unsigned)std::min(ybatch, nc - c0)), block(128);
      if (fill) hipLaunchKernelGGL(k_list<true>, grid, block, 0, st, a);
      else hipLaunchKernelGGL(k_list<false>, grid, block, 0, st, a);
    }
  };
  launch(false);
  MZ_LCHECK(hipGetLastError());
  std::vector<int> counts(na);
  unsigned long long ekey = none;
  MZ_LCHECK(hipMemcpyAsync(counts.data(), d_counts.p, sizeof(int) * (size_t)na, hipMemcpyDeviceToHost, st));
  MZ_LCHECK(hipMemcpyAsync(&ekey, d_err.p, sizeof(ekey), hipMemcpyDeviceToHost, st));
  MZ_LCHECK(hipStreamSynchronize(st));
  if (ekey != none) {
    // the first offending (atom, neighbour) in the host builder's order; its
    // first image below a first knot names the distance
    const int gi = (int)(ekey >> 32), j = (int)(ekey & 0xffffffffu);
    const int k = c->cell_of[gi], ia = gi - c->abase[k];
    const ListCell& L = cells[k];
    const double* b = box + 9 * (size_t)k;
    const double* pi = pos + 3 * (size_t)gi;
    const double* pj = pos + 3 * (size_t)(L.abase + j);
    double rbad = 0.0;
    for (int ix = -L.sc[0]; ix <= L.sc[0] && rbad == 0.0; ++ix)
      for (int iy = -L.sc[1]; iy <= L.sc[1] && rbad == 0.0; ++iy)
        for (int iz = -L.sc[2]; iz <= L.sc[2] && rbad == 0.0; ++iz) {
          if (ia == j && ix == 0 && iy == 0 && iz == 0) continue;
          double d[3];
          for (int m = 0; m < 3; ++m) d[m] = pj[m] + ix * b[m] + iy * b[3 + m] + iz * b[6 + m] - pi[m];
          const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          if (!(r < max_rcut)) continue;
          const int ti = c->types[gi], tj = c->types[L.abase + j];
          const int pij = (ti <= tj) ? ti * c->nt + tj - ((ti * (ti + 1)) / 2) : tj * c->nt + ti - ((tj * (tj + 1)) / 2);
          const int fns[5] = {a.fn_off[0] + pij, a.fn_off[1] + tj, a.fn_off[3] + pij, a.fn_off[5] + pij,
                              a.fn_off[7] + pij};
          for (int q = 0; q < a.nrad; ++q)
            if (r < c->x0[fns[q]]) rbad = r;
        }
    char buf[200];
    std::snprintf(buf, sizeof(buf), "Pair distance too short for potential (r=%g) cell=%d atom=%d neigh=%d", rbad, k,
                  ia, j);
    err = buf;
    return MZ_ERR_GEOMETRY;
  }
  c->atom_poff.assign(na + 1, 0);
  long long np = 0;
  for (int q = 0; q < na; ++q) {
    c->atom_poff[q] = (int)np;
    np += counts[q];
    if (np > 2000000000LL) {
      err = "more than 2e9 pairs";
      return MZ_ERR_CAPACITY;
    }
  }
  c->atom_poff[na] = (int)np;
  c->npairs = np;
  if (c->d_atom_poff.upload(c->atom_poff.data(), c->atom_poff.size()) || c->d_pairs.alloc((size_t)std::max(np, 1LL))) {
    err = "pair-list build: device allocation failed";
    return MZ_ERR_NOMEM;
  }
  a.poff = c->d_atom_poff.p;
  a.pairs = c->d_pairs.p;
  launch(true);
  MZ_LCHECK(hipGetLastError());
  // a host copy only where the triplet counts need the geometry
  if (host_copy) {
    c->pairs.resize(np);
    if (np)
      MZ_LCHECK(hipMemcpyAsync(c->pairs.data(), c->d_pairs.p, sizeof(uint2) * (size_t)np, hipMemcpyDeviceToHost, st));
  }
  MZ_LCHECK(hipStreamSynchronize(st));
  return MZ_OK;
}

}  // namespace mz
