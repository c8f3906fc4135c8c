// C++ host layer over the libmeamz C-ABI (see mzhost.h).
#include "mzhost.h"

#include <algorithm>
#include <atomic>
#include <thread>
#include <chrono>
#include <cmath>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <sstream>

namespace mzh {

unsigned long long Potential::ncalls_ = 0;

namespace {

std::vector<std::string> content_lines(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw Error("Unable to open file to read: " + path);
  std::vector<std::string> out;
  std::string line;
  while (std::getline(in, line)) {
    line.erase(std::remove(line.begin(), line.end(), '\r'), line.end());
    if (line.find_first_not_of(" \t\n\v\f") != std::string::npos) out.push_back(line);
  }
  return out;
}

int style_code(const std::string& s) {
  if (s == "eam/spline") return MZ_STYLE_EAM_SPLINE;
  if (s == "meam/spline") return MZ_STYLE_MEAM_SPLINE;
  if (s == "eam") return MZ_STYLE_EAM;
  if (s == "meam") return MZ_STYLE_MEAM;
  if (s == "gmeam/spline") return MZ_STYLE_GMEAM_SPLINE;
  if (s == "gmeam") return MZ_STYLE_GMEAM;
  if (s == "pair/spline") return MZ_STYLE_PAIR_SPLINE;
  if (s == "pair") return MZ_STYLE_PAIR;
  if (s == "jmeam/spline") return MZ_STYLE_JMEAM_SPLINE;
  if (s == "jmeam") return MZ_STYLE_JMEAM;
  if (s == "jmeam2/spline") return MZ_STYLE_JMEAM2_SPLINE;
  if (s == "jmeam2") return MZ_STYLE_JMEAM2;
  throw Error("invalid potential style \"" + s + "\"");
}

void check(int rc) {
  if (rc != MZ_OK) throw Error(std::string("libmeamz: ") + mz_last_error());
}

}  // namespace

// ------------------------------------------------------------------ formats

int SplineRec::ncoeff() const {
  int c = 0;
  for (int k = 0; k < n; ++k) c += fixed[k] == 0;
  for (int s = 0; s < 2; ++s) c += bc[s] == 0;
  return c;
}

int splines_for_style(const std::string& style, int nt) {
  const int npij = nt * (nt + 1) / 2;
  const int eam = npij + nt + nt;
  style_code(style);
  if (style.compare(0, 4, "pair") == 0) return npij;  // phi Pair_ij only
  if (style.compare(0, 5, "gmeam") == 0) return eam + npij + nt * npij;  // g Triplet_i_jk
  // jmeam: p Pair_ij, q Triplet_i after g; jmeam2: also p2, q2
  if (style.compare(0, 6, "jmeam2") == 0) return eam + npij + nt + 2 * (npij + nt);
  if (style.compare(0, 5, "jmeam") == 0) return eam + npij + nt + (npij + nt);
  return (style.compare(0, 4, "meam") == 0) ? eam + npij + nt : eam;
}

PotFile read_potential_file(const std::string& path) {
  auto L = content_lines(path);
  PotFile pf;
  if (L.size() < 3) throw Error("potential file too short: " + path);
  pf.comment = L[0];
  {
    std::istringstream is(L[1]);
    std::string lab;
    is >> lab >> pf.style >> pf.ntypes;
    if (lab != "#T" || is.fail()) throw Error("potential type line (#T ...) missing or incorrect");
  }
  const int nspl = splines_for_style(pf.style, pf.ntypes);
  size_t ln = 2;
  while (ln < L.size()) {
    if (L[ln].compare(0, 2, "#P") != 0) throw Error("potential line (#P ...) missing");
    ++ln;
    std::vector<SplineRec> pot;
    for (int s = 0; s < nspl; ++s) {
      if (ln >= L.size() || L[ln].compare(0, 2, "#B") != 0) throw Error("potential line (#B ...) missing");
      std::istringstream hb(L[ln++]);
      std::string lab, basis;
      hb >> lab >> basis;
      SplineRec r;
      if (basis == "spline3eq") r.basis = 3;
      else if (basis == "spline5eq") r.basis = 5;
      else throw Error("invalid basis type \"" + basis + "\"");
      std::istringstream a(L[ln++]);
      a >> r.x0 >> r.xn >> r.n;
      if (a.fail() || r.n < 3) throw Error("incorrect spline definition (<left x> <right x> <nknots>)");
      std::istringstream b(L[ln++]);
      b >> r.bc[0] >> r.bc[1];
      if (b.fail()) throw Error("incorrect boundary condition constraint line");
      std::istringstream g(L[ln++]);
      g >> r.grad[0] >> r.grad[1];
      if (g.fail()) throw Error("incorrect boundary condition line");
      if (r.basis == 5) {
        std::istringstream g2(L[ln++]);
        g2 >> r.grad2[0] >> r.grad2[1];
        if (g2.fail()) throw Error("incorrect boundary condition line (grad2)");
      }
      r.y.resize(r.n);
      r.fixed.resize(r.n);
      for (int k = 0; k < r.n; ++k) {
        if (ln >= L.size()) throw Error("unexpected end of potential file");
        std::istringstream kk(L[ln++]);
        kk >> r.fixed[k] >> r.y[k];
        if (kk.fail()) throw Error("incorrect spline knot line (<is y fixed?> <y-value>)");
      }
      pot.push_back(r);
    }
    pf.pots.push_back(pot);
  }
  return pf;
}

void write_potential_file(const std::string& path, const PotFile& pf) {
  std::ofstream out(path);
  out << pf.comment << "\n#T " << pf.style << " " << pf.ntypes << "\n";
  out << std::scientific << std::setprecision(15);
  for (const auto& pot : pf.pots) {
    out << "#P\n";
    for (const auto& s : pot) {
      out << "#B " << (s.basis == 3 ? "spline3eq" : "spline5eq") << "\n";
      out << s.x0 << " " << s.xn << " " << s.n << "\n";
      out << s.bc[0] << " " << s.bc[1] << "\n";
      out << s.grad[0] << " " << s.grad[1] << "\n";
      if (s.basis == 5) out << s.grad2[0] << " " << s.grad2[1] << "\n";
      for (int k = 0; k < s.n; ++k) out << s.fixed[k] << " " << s.y[k] << "\n";
      out << "\n";
    }
  }
}

ConfigDb read_config_file(const std::string& path) {
  auto L = content_lines(path);
  ConfigDb db;
  size_t ln = 0;
  while (ln < L.size()) {
    std::istringstream hn(L[ln]);
    std::string lab;
    int n = 0, useforce = 1;
    hn >> lab >> n >> useforce;
    if (lab != "#N" || hn.fail()) throw Error("number line (#N ...) is missing or incorrect");
    double box[9] = {0}, e0 = 0, s0[6] = {0}, w = 1, ew = 1, sw = 1;
    ++ln;
    while (ln < L.size() && L[ln].compare(0, 2, "#F") != 0) {
      std::istringstream h(L[ln]);
      h >> lab;
      if (lab == "#X") h >> box[0] >> box[1] >> box[2];
      else if (lab == "#Y") h >> box[3] >> box[4] >> box[5];
      else if (lab == "#Z") h >> box[6] >> box[7] >> box[8];
      else if (lab == "#E") h >> e0;
      else if (lab == "#S") h >> s0[0] >> s0[1] >> s0[2] >> s0[3] >> s0[4] >> s0[5];
      else if (lab == "#W") h >> w;
      else if (lab == "#EW") h >> ew;
      else if (lab == "#SW") h >> sw;
      if (h.fail()) throw Error("header line incorrect: " + L[ln]);
      ++ln;
    }
    if (ln + n >= L.size() + 1) throw Error("unexpected end of config file");
    ++ln;
    for (int a = 0; a < n; ++a, ++ln) {
      if (ln >= L.size() || L[ln][0] == '#') throw Error("too few atoms");
      std::istringstream al(L[ln]);
      int t;
      double p[3], f[3];
      al >> t >> p[0] >> p[1] >> p[2] >> f[0] >> f[1] >> f[2];
      if (al.fail()) throw Error("atom line incorrect - atom=" + std::to_string(a));
      db.types.push_back(t);
      db.pos.insert(db.pos.end(), p, p + 3);
      db.force0.insert(db.force0.end(), f, f + 3);
    }
    db.natoms.push_back(n);
    db.box.insert(db.box.end(), box, box + 9);
    db.energy0.push_back(e0);
    db.stress0.insert(db.stress0.end(), s0, s0 + 6);
    db.weight.push_back(w);
    db.eweight.push_back(ew);
    db.sweight.push_back(sw);
    ++db.ncells;
  }
  return db;
}

Params Params::read(const std::string& path) {
  Params p;
  for (const auto& line : content_lines(path)) {
    std::istringstream is(line);
    std::string key, v;
    is >> key;
    if (key.empty() || p.kv.count(key)) continue;
    std::vector<std::string> vals;
    while (is >> v) vals.push_back(v);
    p.kv[key] = vals;
  }
  return p;
}

std::string Params::str(const std::string& k, const std::string& def) const {
  auto it = kv.find(k);
  return (it == kv.end() || it->second.empty()) ? def : it->second[0];
}

double Params::num(const std::string& k, double def) const {
  auto it = kv.find(k);
  return (it == kv.end() || it->second.empty()) ? def : std::stod(it->second[0]);
}

std::vector<std::string> Params::list(const std::string& k) const {
  auto it = kv.find(k);
  return it == kv.end() ? std::vector<std::string>() : it->second;
}

// ---------------------------------------------------------------- evaluator

Evaluator::Evaluator(const std::string& style, int nt, const std::vector<SplineRec>& spl, int embed_extrap,
                     const ConfigDb& db) {
  std::vector<int> basis, nk, bc;
  std::vector<double> x0, xn;
  for (const auto& s : spl) {
    basis.push_back(s.basis);
    nk.push_back(s.n);
    x0.push_back(s.x0);
    xn.push_back(s.xn);
    bc.push_back(s.bc[0]);
    bc.push_back(s.bc[1]);
  }
  mz_layout lay{style_code(style), nt, (int)spl.size(), basis.data(), nk.data(), x0.data(), xn.data(), bc.data(),
                embed_extrap};
  mz_db d{db.ncells, db.natoms.data(), db.box.data(), db.types.data(), db.pos.data(), db.force0.data(),
          db.energy0.data(), db.stress0.data(), db.weight.data(), db.eweight.data(), db.sweight.data(),
          db.energy_weight, db.stress_weight};
  check(mz_create(&lay, &d, &ctx_));
  check(mz_info(ctx_, &info_));
}

Evaluator::~Evaluator() { mz_destroy(ctx_); }

// ---------------------------------------------------------------- potential

Potential::Potential(std::shared_ptr<Evaluator> ev, const std::string& style, int nt, std::vector<SplineRec> spl,
                     int embed_extrap)
    : ev_(std::move(ev)), style_(style), ntypes_(nt), embed_extrap_(embed_extrap), spl_(std::move(spl)) {
  int off = 0;
  std::vector<int> walk;
  for (const auto& s : spl_) {
    koff_.push_back(off);
    for (int k = 0; k < s.n; ++k)
      if (!s.fixed[k]) walk.push_back(off + k);
    for (int side = 0; side < 2; ++side)
      if (s.bc[side] == 0) {
        walk.push_back(off + s.n + side);
        if (s.basis == 5) walk.push_back(off + s.n + 2 + side);
      }
    ncoeff_ += s.ncoeff();
    off += s.n + 4;
  }
  walk.resize(std::min<size_t>(walk.size(), ncoeff_));
  slots_ = walk;
}

double* Potential::slot_ptr(int slot) {
  for (size_t s = 0; s < spl_.size(); ++s) {
    const int n = spl_[s].n;
    if (slot >= koff_[s] && slot < koff_[s] + n + 4) {
      const int q = slot - koff_[s];
      if (q < n) return &spl_[s].y[q];
      if (q < n + 2) return &spl_[s].grad[q - n];
      return &spl_[s].grad2[q - n - 2];
    }
  }
  throw Error("unable to access coefficient for optimization");
}

double& Potential::coeff(int idx) {
  if (idx < 0 || idx >= ncoeff_) throw Error("unable to access coefficient for optimization");
  return *slot_ptr(slots_[idx]);
}

void Potential::modify_coeff(int idx, double height, double width) {
  int sum = 0;
  for (auto& s : spl_) {
    const int nc = s.ncoeff();
    sum += nc;
    if (idx < sum) {
      int local = idx + nc - sum;
      int kfound = -1;
      for (int k = 0; k < s.n; ++k)
        if (!s.fixed[k] && local-- == 0) {
          kfound = k;
          break;
        }
      if (kfound < 0) {
        for (int side = 0; side < 2; ++side)
          if (s.bc[side] == 0) {
            if (local-- == 0) {
              s.grad[side] += height;
              return;
            }
            if (s.basis == 5 && local-- == 0) {
              s.grad2[side] += height;
              return;
            }
          }
        return;
      }
      const double inv_sqrt_2pi = 0.3989422804014326779;
      const int span = (int)(4 * width);
      for (int i = -span; i <= span; ++i) {
        const int k = kfound + i;
        if (k >= 0 && k < s.n && !s.fixed[k]) s.y[k] += inv_sqrt_2pi * std::exp(-i * i * 0.5 / (width * width)) * height;
      }
      return;
    }
  }
}

std::vector<double> Potential::knot_vector() const {
  std::vector<double> kv;
  for (const auto& s : spl_) {
    kv.insert(kv.end(), s.y.begin(), s.y.end());
    kv.push_back(s.grad[0]);
    kv.push_back(s.grad[1]);
    kv.push_back(s.grad2[0]);
    kv.push_back(s.grad2[1]);
  }
  return kv;
}

void Potential::set_knot_vector(const std::vector<double>& kv) {
  for (size_t s = 0; s < spl_.size(); ++s) {
    auto& r = spl_[s];
    const int o = koff_[s];
    for (int k = 0; k < r.n; ++k) r.y[k] = kv[o + k];
    r.grad[0] = kv[o + r.n];
    r.grad[1] = kv[o + r.n + 1];
    r.grad2[0] = kv[o + r.n + 2];
    r.grad2[1] = kv[o + r.n + 3];
  }
}

double Potential::compute(ErrorVec* ev) {
  std::vector<double> kv = knot_vector();
  double F = 0.0;
  if (ev) {
    ev->assign(ev_->info().nerr, 0.0);
    check(mz_objective(ev_->ctx(), kv.data(), &F, ev->data(), nullptr, nullptr, nullptr));
  } else {
    check(mz_objective(ev_->ctx(), kv.data(), &F, nullptr, nullptr, nullptr, nullptr));
  }
  error_sum_ = F;
  ++ncalls_;
  return F;
}

double Potential::compute_outputs(ErrorVec* ev, std::vector<double>* cell_energy, std::vector<double>* cell_stress,
                                  std::vector<double>* forces) {
  std::vector<double> kv = knot_vector();
  const mz_info_t& in = ev_->info();
  ev->assign(in.nerr, 0.0);
  cell_energy->assign(in.ncells, 0.0);
  cell_stress->assign(6 * (size_t)in.ncells, 0.0);
  forces->assign(3 * (size_t)in.natoms, 0.0);
  double F = 0.0;
  check(mz_objective(ev_->ctx(), kv.data(), &F, ev->data(), cell_energy->data(), cell_stress->data(),
                     forces->data()));
  error_sum_ = F;
  ++ncalls_;
  return F;
}

void Potential::compute_densities(std::vector<double>* dens) {
  std::vector<double> kv = knot_vector();
  dens->assign(ev_->info().natoms, 0.0);
  check(mz_densities(ev_->ctx(), kv.data(), dens->data()));
}

double Potential::compute_grad(std::vector<double>* grad) {
  std::vector<double> kv = knot_vector();
  std::vector<double> g(kv.size());
  double F = 0.0;
  check(mz_objective_grad(ev_->ctx(), kv.data(), &F, g.data()));
  grad->assign(ncoeff_, 0.0);
  for (int i = 0; i < ncoeff_; ++i) (*grad)[i] = g[slots_[i]];
  error_sum_ = F;
  ++ncalls_;
  return F;
}

void Potential::set_dir(const std::vector<double>& dir, double x) {
  for (int i = 0; i < ncoeff_; ++i) coeff(i) += x * dir[i];
}

double Potential::compute_dir(ErrorVec* ev, const std::vector<double>& dir, double x) {
  Potential tmp(*this);
  tmp.set_dir(dir, x);
  return tmp.compute(ev);
}

// ---------------------------------------------------------------- line minimisation

namespace {
constexpr double CGOLD = 0.381966012;      // src/line_min.h:45
constexpr unsigned MAX_BRACKET = 100;      // src/line_min.h:46
constexpr unsigned ITMAX = 100;            // src/line_min_brent.h:29
constexpr double TOL = 1.0e-1;             // src/line_min_brent.h:30
constexpr double ZEPS = 1.0e-9;            // src/line_min_brent.h:31
}  // namespace

void LineMinBrent::bracket() {
  double xl = 0.0, xm, xr = 0.1;
  double fl = eval_(xl), fm, fr = eval_(xr);
  if (fr >= fl) {
    xm = xl;
    fm = fl;
    xl = -(xr - xm) / CGOLD + xr;
    fl = eval_(xl);
  } else {
    xm = xr;
    fm = fr;
    xr = (xm - xl) / CGOLD + xl;
    fr = eval_(xr);
  }
  for (unsigned count = 0; count < MAX_BRACKET; ++count) {
    if (fm < fl) {
      if (fm < fr) {
        xa_ = xl; xb_ = xm; xc_ = xr;
        fa_ = fl; fb_ = fm; fc_ = fr;
        return;
      } else if (fm > fr) {
        xl = xm; fl = fm;
        xm = xr; fm = fr;
        xr = (xm - xl) / CGOLD + xl;
        fr = eval_(xr);
      } else {
        xr = (xr - xl) * CGOLD + xr;
        fr = eval_(xr);
      }
    } else if (fm > fl) {
      xr = xm; fr = fm;
      xm = xl; fm = fl;
      xl = -(xr - xm) / CGOLD + xr;
      fl = eval_(xl);
    } else {
      xl = -(xr - xl) * CGOLD + xl;
      fl = eval_(xl);
    }
  }
  throw Error("problems with bracketing of minimum");
}

double LineMinBrent::find_minima() {
  double a = std::min(xa_, xc_), b = std::max(xa_, xc_);
  double v = xb_, w = v, x = v, d = 0.0, e = 0.0;
  double fx = fb_, fv = fx, fw = fx;
  for (unsigned it = 0; it < ITMAX; ++it) {
    double u = 0.0;
    const double xm = 0.5 * (a + b);
    const double tol1 = TOL * std::abs(x) + ZEPS, tol2 = 2. * tol1;
    if (std::abs(x - xm) <= (tol2 - 0.5 * (b - a))) {
      xmin1_ = x; xmin2_ = w; fmin1_ = fx; fmin2_ = fw;
      return fmin1_;
    }
    if (std::abs(e) > tol1) {
      double r = (x - w) * (fx - fv);
      double q = (x - v) * (fx - fw);
      double p = (x - v) * q - (x - w) * r;
      q = 2. * (q - r);
      if (q > 0.) p = -p; else q = -q;
      const double etemp = e;
      e = d;
      if (std::abs(p) >= std::abs(0.5 * q * etemp) || p <= q * (a - x) || p >= q * (b - x)) {
        e = (x < xm) ? b - x : -(x - a);
        d = CGOLD * e;
      } else {
        d = p / q;
        u = x + d;
        if ((u - a) < tol2 || (b - u) < tol2) d = (x < xm) ? tol1 : -tol1;
      }
    } else {
      e = (x < xm) ? b - x : -(x - a);
      d = CGOLD * e;
    }
    u = (std::abs(d) >= tol1) ? x + d : x + ((d > 0) ? tol1 : -tol1);
    const double fu = eval_(u);
    if (fu <= fx) {
      if (u >= x) a = x; else b = x;
      v = w; fv = fw; w = x; fw = fx; x = u; fx = fu;
    } else {
      if (u < x) a = u; else b = u;
      if (fu <= fw || w == x) { v = w; fv = fw; w = u; fw = fu; }
      else if (fu <= fv || v == x || v == w) { v = u; fv = fu; }
    }
  }
  throw Error("too many iterations in Brent minimization");
}

// ---------------------------------------------------------------- Powell

namespace {

constexpr double P_EPS = 1.E-4;       // src/opt_powell.h:43-47
constexpr double P_PRECISION = 1.E-7;
constexpr double P_NOTHING = 1.E-12;
constexpr int P_INNERLOOPS = 801;
constexpr double P_TOOBIG = 1.E+4;

// Solves the symmetric system A q = p by LU with partial pivoting and one
// step of iterative refinement; returns forward and backward error estimates
// in the sense of dsysvx (componentwise backward error, relative correction).
bool sym_solve(int n, const std::vector<double>& A, const std::vector<double>& p, std::vector<double>& q,
               double& ferr, double& berr) {
  std::vector<double> LU = A;
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int pk = k;
    for (int i = k + 1; i < n; ++i)
      if (std::abs(LU[i * n + k]) > std::abs(LU[pk * n + k])) pk = i;
    piv[k] = pk;
    if (LU[pk * n + k] == 0.0) return false;
    if (pk != k)
      for (int c = 0; c < n; ++c) std::swap(LU[k * n + c], LU[pk * n + c]);
    for (int i = k + 1; i < n; ++i) {
      LU[i * n + k] /= LU[k * n + k];
      for (int c = k + 1; c < n; ++c) LU[i * n + c] -= LU[i * n + k] * LU[k * n + c];
    }
  }
  auto solve = [&](std::vector<double> b) {
    for (int k = 0; k < n; ++k) {
      std::swap(b[k], b[piv[k]]);
      for (int i = k + 1; i < n; ++i) b[i] -= LU[i * n + k] * b[k];
    }
    for (int k = n - 1; k >= 0; --k) {
      for (int c = k + 1; c < n; ++c) b[k] -= LU[k * n + c] * b[c];
      b[k] /= LU[k * n + k];
    }
    return b;
  };
  q = solve(p);
  std::vector<double> r(n);
  berr = 0.0;
  for (int i = 0; i < n; ++i) {
    double s = p[i], den = std::abs(p[i]);
    for (int j = 0; j < n; ++j) {
      s -= A[i * n + j] * q[j];
      den += std::abs(A[i * n + j] * q[j]);
    }
    r[i] = s;
    if (den > 0) berr = std::max(berr, std::abs(s) / den);
  }
  std::vector<double> dq = solve(r);
  double nq = 0, nd = 0;
  for (int i = 0; i < n; ++i) {
    q[i] += dq[i];
    nq = std::max(nq, std::abs(q[i]));
    nd = std::max(nd, std::abs(dq[i]));
  }
  ferr = nq > 0 ? nd / nq : 0.0;
  return true;
}

}  // namespace

Potential* powell(const Potential& start, const OptOptions& opt) {
  std::ostream* out = opt.out;
  Potential* pot = start.clone();
  ErrorVec ev;
  double F = pot->compute(&ev);
  const int ncoeff = pot->get_ncoeff();
  const int esize = (int)ev.size();
  if (opt.max_steps == 0 || ncoeff == 0) return pot;
  if (F < P_NOTHING) {
    if (out) *out << "Error already too small to optimize, aborting..." << std::endl;
    return pot;
  }
  std::vector<double> gamma((size_t)esize * ncoeff), d((size_t)ncoeff * ncoeff), les((size_t)ncoeff * ncoeff),
      p(ncoeff), q(ncoeff), delta(ncoeff);
  auto G = [&](int j, int i) -> double& { return gamma[(size_t)i * esize + j]; };
  auto D = [&](int i, int j) -> double& { return d[(size_t)i * ncoeff + j]; };
  // Forward-difference Jacobian of the residual vector, one column per
  // coefficient (src/opt_powell.cpp:348-389); the ncoeff perturbed residual
  // vectors come from batched calls of at most kJacBatch knot vectors.
  constexpr int kJacBatch = 64;
  auto gamma_init = [&](const Potential* base) -> int {
    for (int i = 0; i < ncoeff; ++i)
      for (int j = 0; j < ncoeff; ++j) D(i, j) = (i == j) ? 1. : 0.;
    const std::vector<double> kv0 = base->knot_vector();
    const std::vector<int>& slots = base->coeff_slots();
    const size_t nk = kv0.size();
    mz_ctx* ctx = base->evaluator()->ctx();
    std::vector<double> kb, Fb, Eb;
    for (int i0 = 0; i0 < ncoeff; i0 += kJacBatch) {
      const int nb = std::min(kJacBatch, ncoeff - i0);
      kb.assign((size_t)nb * nk, 0.0);
      for (int b = 0; b < nb; ++b) {
        std::copy(kv0.begin(), kv0.end(), kb.begin() + (size_t)b * nk);
        kb[(size_t)b * nk + slots[i0 + b]] += P_EPS;
      }
      Fb.resize(nb);
      Eb.resize((size_t)nb * esize);
      if (mz_residuals_batch(ctx, nb, kb.data(), Fb.data(), Eb.data()) != MZ_OK)
        throw Error(std::string("mz_residuals_batch: ") + mz_last_error());
      Potential::add_ncalls(nb);
      for (int b = 0; b < nb; ++b) {
        const int i = i0 + b;
        const double* tv = Eb.data() + (size_t)b * esize;
        double sum = 0.0;
        for (int j = 0; j < esize; ++j) {
          const double t = (tv[j] - ev[j]) / P_EPS;
          G(j, i) = t;
          sum += t * t;
        }
        const double mag = std::sqrt(sum);
        if (mag < P_NOTHING) return i + 1;
        for (int j = 0; j < esize; ++j) G(j, i) /= mag;
        D(i, i) /= mag;
      }
    }
    return 0;
  };
  // Normal equations of the Gauss-Newton step: p = -Gamma^T e and
  // les = Gamma^T Gamma (all of it, or one row and column).  The full product
  // is O(ncoeff^2 esize) and is split over host threads by column.
  auto lineqsys = [&](int only_col) {
    auto column = [&](int i) {
      p[i] = 0.;
      for (int j = 0; j < esize; ++j) p[i] -= G(j, i) * ev[j];
      if (only_col >= 0 && i != only_col) return;
      for (int j = (only_col >= 0 ? 0 : i); j < ncoeff; ++j) {
        double s = 0.0;
        for (int k = 0; k < esize; ++k) s += G(k, i) * G(k, j);
        les[(size_t)i * ncoeff + j] = s;
        les[(size_t)j * ncoeff + i] = s;
      }
    };
    const int nth = only_col >= 0 ? 1 : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (nth == 1) {
      for (int i = 0; i < ncoeff; ++i) column(i);
      return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nth; ++t)
      th.emplace_back([&]() {
        for (int i = next++; i < ncoeff; i = next++) column(i);
      });
    for (auto& x : th) x.join();
  };
  if (out) *out << std::setw(10) << " " << std::setw(3) << "n" << std::setw(8) << "ncalls" << "   F" << std::endl;
  double F_older = F;
  int n = 0;
  do {
    const auto calls0 = Potential::get_ncalls();
    if (int i = gamma_init(pot)) {
      if (out) *out << "F does not depend on coefficient " << i - 1 << ", fit impossible!" << std::endl;
      break;
    }
    lineqsys(-1);
    F_older = F;
    int m = 0;
    double dF = 0.0, delta_mag = 0.0;
    do {
      const double F_old = F;
      double ferr = 0, berr = 0;
      if (!sym_solve(ncoeff, les, p, q, ferr, berr)) {
        if (out) *out << "Linear equation system singular after step " << m << std::endl;
        break;
      }
      delta_mag = 0.0;
      for (int i = 0; i < ncoeff; ++i) {
        delta[i] = 0.;
        for (int j = 0; j < ncoeff; ++j) delta[i] += D(i, j) * q[j];
        delta_mag += delta[i] * delta[i];
      }
      LineMinBrent lm;
      lm.set_func([&](double x) { return pot->compute_dir(nullptr, delta, x); });
      F = lm.find_minima();
      pot->set_dir(delta, lm.xmin1());
      pot->compute(&ev);
      if (ferr + berr > 1. && m > 5) break;
      int max_col = 0;
      double max_change = 0.;
      for (int i = 0; i < ncoeff; ++i)
        if (std::abs(p[i] * q[i]) > max_change) {
          max_change = std::abs(p[i] * q[i]);
          max_col = i;
        }
      // gamma_update (src/opt_powell.cpp:397-434)
      {
        const double l1 = lm.xmin1(), l2 = lm.xmin2();
        const double Fm = lm.fmin1();
        ErrorVec f2;
        pot->compute_dir(&f2, delta, l2 - l1);
        double mu = 0.0;
        for (int j = 0; j < esize; ++j) {
          const double t = (ev[j] - f2[j]) / (l1 - l2);
          G(j, max_col) = t;
          mu += t * ev[j];
        }
        mu /= Fm;
        double sum = 0.0;
        for (int j = 0; j < esize; ++j) {
          const double t = G(j, max_col) - mu * ev[j];
          G(j, max_col) = t;
          sum += t * t;
        }
        const double mag = std::sqrt(sum);
        if (mag < P_NOTHING) {
          if (out) *out << "Matrix gamma singular after step " << m << ", restarting inner loop" << std::endl;
          break;
        }
        for (int j = 0; j < esize; ++j) G(j, max_col) /= mag;
        for (int i = 0; i < ncoeff; ++i) delta[i] /= mag;
      }
      for (int i = 0; i < ncoeff; ++i) D(i, max_col) = delta[i];
      lineqsys(max_col);
      dF = F_old - F;
      ++m;
    } while ((m < ncoeff || (m <= P_INNERLOOPS && dF > P_PRECISION)) && dF < P_TOOBIG && delta_mag < P_TOOBIG);
    if (out)
      *out << std::setw(13) << m << std::setw(8) << Potential::get_ncalls() - calls0 << "   " << std::scientific
           << std::setprecision(8) << F << std::endl;
    ++n;
  } while (((F_older - F > P_PRECISION / 10.) || (F_older - F < 0)) && (F_older - F > opt.d_eps) &&
           n < opt.max_steps);
  return pot;
}

// ---------------------------------------------------------------- L-BFGS

// Two-loop recursion (Nocedal & Wright, Numerical Optimization, alg. 7.4)
// with a backtracking line search on the sufficient-decrease (Armijo)
// condition; gradients come from the adjoint sweep.
Potential* lbfgs(const Potential& start, const OptOptions& opt) {
  std::ostream* out = opt.out;
  Potential* pot = start.clone();
  const int n = pot->get_ncoeff();
  std::vector<double> x(n), g(n), gnew(n), dir(n);
  for (int i = 0; i < n; ++i) x[i] = pot->coeff(i);
  double F = pot->compute_grad(&g);
  const int M = std::max(1, opt.lbfgs_mem);
  std::vector<std::vector<double>> S, Y;
  std::vector<double> rho;
  if (out) *out << "  iter        F              |g|            step    seconds" << std::endl;
  const auto t0 = std::chrono::steady_clock::now();
  for (int it = 0; it < opt.max_steps; ++it) {
    // two-loop recursion
    std::vector<double> qv = g, alpha(S.size());
    for (int k = (int)S.size() - 1; k >= 0; --k) {
      double a = 0;
      for (int i = 0; i < n; ++i) a += S[k][i] * qv[i];
      a *= rho[k];
      alpha[k] = a;
      for (int i = 0; i < n; ++i) qv[i] -= a * Y[k][i];
    }
    double gamma = 1.0;
    if (!S.empty()) {
      double sy = 0, yy = 0;
      for (int i = 0; i < n; ++i) {
        sy += S.back()[i] * Y.back()[i];
        yy += Y.back()[i] * Y.back()[i];
      }
      gamma = sy / yy;
    } else {
      double gg = 0;
      for (int i = 0; i < n; ++i) gg += g[i] * g[i];
      gamma = gg > 0 ? 1e-3 / std::sqrt(gg) * std::max(1.0, F) / std::max(1.0, F) : 1.0;
    }
    for (int i = 0; i < n; ++i) qv[i] *= gamma;
    for (size_t k = 0; k < S.size(); ++k) {
      double b = 0;
      for (int i = 0; i < n; ++i) b += Y[k][i] * qv[i];
      b *= rho[k];
      for (int i = 0; i < n; ++i) qv[i] += S[k][i] * (alpha[k] - b);
    }
    double slope = 0;
    for (int i = 0; i < n; ++i) {
      dir[i] = -qv[i];
      slope += dir[i] * g[i];
    }
    if (!(slope < 0)) {  // not a descent direction: restart from steepest descent
      S.clear(); Y.clear(); rho.clear();
      double gg = 0;
      for (int i = 0; i < n; ++i) gg += g[i] * g[i];
      const double sc = 1e-3 / std::sqrt(std::max(gg, 1e-300));
      slope = 0;
      for (int i = 0; i < n; ++i) {
        dir[i] = -g[i] * sc;
        slope += dir[i] * g[i];
      }
    }
    double step = 1.0, Fnew = F;
    Potential* trial = nullptr;
    for (int ls = 0; ls < 40; ++ls) {
      delete trial;
      trial = pot->clone();
      trial->set_dir(dir, step);
      Fnew = trial->compute_grad(&gnew);
      if (std::isfinite(Fnew) && Fnew <= F + 1e-4 * step * slope) break;
      step *= 0.5;
    }
    if (!(Fnew < F)) {
      delete trial;
      if (out) *out << "L-BFGS: no decrease along the search direction, stopping" << std::endl;
      break;
    }
    std::vector<double> s(n), y(n);
    double sy = 0;
    for (int i = 0; i < n; ++i) {
      s[i] = trial->coeff(i) - pot->coeff(i);
      y[i] = gnew[i] - g[i];
      sy += s[i] * y[i];
    }
    delete pot;
    pot = trial;
    const double Fold = F;
    F = Fnew;
    g = gnew;
    if (sy > 1e-300) {
      S.push_back(s);
      Y.push_back(y);
      rho.push_back(1.0 / sy);
      if ((int)S.size() > M) {
        S.erase(S.begin());
        Y.erase(Y.begin());
        rho.erase(rho.begin());
      }
    }
    double gn = 0;
    for (int i = 0; i < n; ++i) gn += g[i] * g[i];
    if (out)
      *out << std::setw(6) << it + 1 << "  " << std::scientific << std::setprecision(8) << F << "  "
           << std::sqrt(gn) << "  " << step << "  " << std::fixed << std::setprecision(4)
           << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() << std::endl;
    if (Fold - F <= opt.d_eps * std::max(1.0, std::abs(F)) && opt.d_eps > 0) break;
  }
  return pot;
}

// ---------------------------------------------------------------- simulated annealing

Potential* simann(const Potential& start, const OptOptions& opt, std::mt19937& rng) {
  constexpr double EPS = 0.1;      // src/opt_simann.h:41-46
  constexpr int NEPS = 4;
  constexpr int NSTEP = 20;
  constexpr double STEPVAR = 2.0;
  constexpr double TEMPVAR = 0.85;
  std::normal_distribution<double> norm(0.0, 1.0);
  std::uniform_real_distribution<double> unif(0.0, 1.0);
  Potential* pot = start.clone();
  double F = pot->compute();
  const int ncoeff = pot->get_ncoeff();
  if (opt.max_steps == 0 || ncoeff == 0) return pot;
  Potential* best = pot->clone();
  double Fopt = F;
  std::vector<double> v(ncoeff, 1.0);
  std::vector<int> naccept(ncoeff, 0);
  std::vector<double> Fvar(opt.max_steps + 5 + NEPS, F);
  int k = 0;
  double T = opt.anneal_temp;
  bool loopagain = true;
  do {
    for (int m = 0; m < 3 * ncoeff; ++m) {
      for (int j = 0; j < NSTEP; ++j)
        for (int c = 0; c < ncoeff; ++c) {
          Potential trial(*pot);
          const double height = norm(rng) * v[c];
          const double width = std::abs(norm(rng));
          trial.modify_coeff(c, height, width);
          const double F2 = trial.compute();
          bool acc = false;
          if (F2 <= F) {
            acc = true;
            if (F2 <= Fopt) {
              delete best;
              best = trial.clone();
              Fopt = F2;
            }
          } else if (unif(rng) < std::exp((F - F2) / T)) {
            acc = true;
          }
          if (acc) {
            delete pot;
            pot = trial.clone();
            F = F2;
            ++naccept[c];
          }
        }
      for (int c = 0; c < ncoeff; ++c) {
        if (naccept[c] > 0.6 * NSTEP) v[c] *= (1 + STEPVAR * ((double)naccept[c] / NSTEP - 0.6) / 0.4);
        else if (naccept[c] < 0.4 * NSTEP) v[c] /= (1 + STEPVAR * (0.4 - (double)naccept[c] / NSTEP) / 0.4);
        naccept[c] = 0;
      }
      if (opt.out)
        *opt.out << std::setw(3) << k << "  T=" << std::fixed << T << "  m=" << m + 1 << "  F=" << std::scientific
                 << F << "  Fopt=" << Fopt << std::endl;
    }
    T *= TEMPVAR;
    ++k;
    loopagain = false;
    Fvar[k + NEPS] = F;
    for (int i = 1; i <= NEPS; ++i)
      if (std::abs(F - Fvar[k + NEPS - i]) > (EPS * F * 0.01)) loopagain = true;
    if (!loopagain && ((F - Fopt) > (EPS * F * 0.01))) {
      delete pot;
      pot = best->clone();
      F = Fopt;
      loopagain = true;
    }
  } while (k < opt.max_steps && loopagain);
  delete pot;
  return best;
}

// ---------------------------------------------------------------- genetic algorithm

std::vector<Potential*> genalg(const std::vector<Potential*>& pop_in, const OptOptions& opt, std::mt19937& rng) {
  std::normal_distribution<double> norm(0.0, 1.0);
  std::uniform_real_distribution<double> unif(0.0, 1.0);
  std::vector<Potential*> pop;
  for (auto* p : pop_in) pop.push_back(p->clone());
  const int ncoeff = pop.front()->get_ncoeff();
  while ((int)pop.size() < opt.pop_size) {
    Potential* p = pop.front()->clone();
    for (int c = 0; c < ncoeff; ++c) {
      const double height = norm(rng) * p->coeff(c) * opt.init_scale;
      const double width = std::abs(norm(rng));
      p->modify_coeff(c, height, width);
    }
    pop.push_back(p);
  }
  auto evaluator = pop.front()->evaluator();
  const int nknot = evaluator->info().nknot;
  // one batched call scores the whole population
  auto score = [&](std::vector<Potential*>& ps) {
    std::vector<double> K((size_t)ps.size() * nknot), F(ps.size());
    for (size_t i = 0; i < ps.size(); ++i) {
      auto kv = ps[i]->knot_vector();
      std::copy(kv.begin(), kv.end(), K.begin() + i * nknot);
    }
    check(mz_objective_batch(evaluator->ctx(), (int)ps.size(), K.data(), F.data(), nullptr));
    Potential::add_ncalls(ps.size());
    for (size_t i = 0; i < ps.size(); ++i) ps[i]->set_error(F[i]);
  };
  score(pop);
  std::stable_sort(pop.begin(), pop.end(), [](Potential* a, Potential* b) { return *a < *b; });
  const int P = (int)pop.size();
  for (int gen = 0; gen < opt.ngens; ++gen) {
    std::vector<Potential*> kids;
    int parent = 0;
    std::vector<int> idx;
    for (int i = 0; i < P - 2; ++i) idx.push_back(i + 1);
    for (int i = 0; i < P; ++i) {
      Potential* kid = pop[i]->clone();
      if (i < 2) {
        kids.push_back(kid);
        continue;
      }
      if (i > 2 && unif(rng) < opt.fit_frac) ++parent;
      int p2 = i - 1;
      if (!idx.empty()) {
        int t;
        do {
          t = (int)(unif(rng) * idx.size());
          p2 = idx[t];
          if (idx.size() == 1 && p2 == parent) ++parent;
        } while (p2 == parent);
        idx.erase(idx.begin() + t);
      }
      int sw = (int)(unif(rng) * 2), success = 0;
      for (int c = 0; c < ncoeff; ++c) {
        if (unif(rng) < opt.cross_rate) {
          sw = (sw + 1) % 2;
          success = 1;
        }
        if (c == ncoeff - 1 && !success) sw = (sw + 1) % 2;
        kid->coeff(c) = pop[(sw == 0) ? std::min(parent, P - 1) : p2]->coeff(c);
        if (unif(rng) < opt.mut_rate) {
          int r;
          do {
            r = (int)(unif(rng) * P);
          } while (r == parent || r == p2);
          kid->coeff(c) = pop[r]->coeff(c);
        }
      }
      kids.push_back(kid);
    }
    for (auto* p : pop) delete p;
    pop = kids;
    score(pop);
    std::stable_sort(pop.begin(), pop.end(), [](Potential* a, Potential* b) { return *a < *b; });
    if (opt.out) *opt.out << "generation " << gen + 1 << "  best F = " << std::scientific << pop[0]->get_error()
                          << std::endl;
  }
  return pop;
}

// ---------------------------------------------------------------- output files

namespace {

// dy/dx of a refreshed spline in interval k at shift t (a = 1 - t): the
// derivative of the cubic (and quintic) interpolation formula of
// src/spline3eq.h:96-105 and src/spline5eq.h:95-107.
double spline_slope(const SplineRec& s, const std::vector<double>& ypp, const std::vector<double>& yp4, int k,
                    double t) {
  const double step = (s.xn - s.x0) / (s.n - 1), invstep = 1. / step, a = 1.0 - t;
  double d = (s.y[k + 1] - s.y[k]) * invstep + ((3 * t * t - 1) * ypp[k + 1] - (3 * a * a - 1) * ypp[k]) * step / 6.0;
  if (s.basis == 5)
    d += ((15 * t * t * t * t - 30 * t * t + 7) * yp4[k + 1] - (15 * a * a * a * a - 30 * a * a + 7) * yp4[k]) *
         (step * step * step) / 360.0;
  return d;
}

std::string numbered(std::string base, const char* suffix, int num) {
  base += suffix;
  if (num >= 0) base += "." + std::to_string(num);
  return base;
}

std::ofstream open_out(const std::string& path) {
  std::ofstream out(path);
  if (!out) throw Error("Unable to open file to write: " + path);
  return out;
}

}  // namespace

void write_lammps_file(const std::string& path, const PotFile& pf) {
  for (size_t p = 0; p < pf.pots.size(); ++p) {
    std::ofstream out = open_out(pf.pots.size() > 1 ? path + "." + std::to_string(p) : path);
    out << pf.comment << "\n" << pf.style << "\n" << std::scientific << std::setprecision(15);
    for (const SplineRec& s : pf.pots[p]) {
      const int n = s.n;
      std::vector<double> rec(s.y);
      rec.insert(rec.end(), {s.grad[0], s.grad[1], s.grad2[0], s.grad2[1]});
      std::vector<double> ypp(n), yp4(n);
      check(mz_spline_refresh(s.basis, n, s.x0, s.xn, s.bc[0] == -1, s.bc[1] == -1, rec.data(), ypp.data(),
                              yp4.data()));
      out << (s.basis == 3 ? "spline3eq" : "spline5eq") << "\n" << n << "\n";
      out << spline_slope(s, ypp, yp4, 0, 0.0) << " " << spline_slope(s, ypp, yp4, n - 2, 1.0) << "\n";
      const double step = (s.xn - s.x0) / (n - 1);
      for (int k = 0; k < n; ++k) {
        const double x = k == 0 ? s.x0 : (k == n - 1 ? s.xn : s.x0 + step * k);
        out << x << " " << s.y[k] << " " << ypp[k];
        if (s.basis == 5) out << " " << yp4[k];
        out << "\n";
      }
    }
    out << "\n";
  }
}

void write_data_files(const std::string& base, Potential& pot, const ConfigDb& db, int num) {
  ErrorVec ev;
  std::vector<double> ce, cs, fr, dens;
  pot.compute_outputs(&ev, &ce, &cs, &fr);
  // pair styles write neither densities nor constraints: Potential's
  // write_extras and write_punish do nothing (src/potential.cpp:713-725)
  const bool pair = pot.style().compare(0, 4, "pair") == 0;
  if (!pair) pot.compute_densities(&dens);
  const int nc = db.ncells;
  int natoms = 0;
  for (int c = 0; c < nc; ++c) natoms += db.natoms[c];
  const bool gmeam = pot.style().compare(0, 5, "gmeam") == 0;
  const bool jmeam = pot.style().compare(0, 5, "jmeam") == 0;
  const int njm = jmeam ? (pot.style().compare(0, 6, "jmeam2") == 0 ? 2 : 1) : 0;
  const bool meam = gmeam || jmeam || pot.style().compare(0, 4, "meam") == 0;
  using std::setw;
  using std::setprecision;
  if (!pair) {  // densities (src/pots/pot_eam/pot_eam.cpp:506-543)
    std::ofstream o = open_out(numbered(base, ".rho", num));
    o << setw(6) << std::right << "conf:" << setw(5) << "atom" << setw(8) << "type" << setw(15) << "rho" << "\n";
    for (int c = 0, a = 0; c < nc; ++c)
      for (int i = 0; i < db.natoms[c]; ++i, ++a)
        o << setw(5) << c << ":" << setw(5) << i << setw(8) << db.types[a] << setw(15) << std::fixed << dens[a]
          << "\n";
  }
  {  // energies (src/potential.cpp:497-542)
    std::ofstream o = open_out(numbered(base, ".energy", num));
    o << "Global energy weight is " << std::fixed << db.energy_weight << "\n";
    o << std::right << setw(3) << "#" << setw(11) << "conf_w" << setw(11) << "conf_ew" << setw(9) << "natoms"
      << setw(15) << "energy error" << setw(18) << "e" << setw(18) << "e0" << setw(15) << "|e-e0|" << setw(15)
      << "e-e0" << setw(15) << "de/e0" << "\n";
    for (int c = 0, idx = 0; c < nc; ++c) {
      const double e = ce[c], e0 = db.energy0[c];
      o << setw(3) << c << std::fixed << setw(11) << setprecision(4) << db.weight[c] << setw(11) << db.eweight[c]
        << setw(9) << db.natoms[c] << setw(15) << setprecision(6) << ev[idx] * ev[idx] << setw(18)
        << setprecision(10) << e << setw(18) << e0 << setw(15) << setprecision(6) << std::abs(e - e0) << setw(15)
        << e - e0;
      if (e0 == 0) o << setw(15) << "inf";
      else o << setw(15) << (e0 - e) / e0;
      o << "\n";
      idx += 7 + 3 * db.natoms[c];
    }
  }
  {  // stresses (src/potential.cpp:548-591)
    std::ofstream o = open_out(numbered(base, ".stress", num));
    o << "Global stress weight is " << std::fixed << db.stress_weight << "\n";
    o << std::right << setw(3) << "#" << setw(11) << "conf_w" << setw(11) << "conf_sw" << setw(15) << "stress error"
      << setw(18) << "s" << setw(18) << "s0" << setw(15) << "ds/s0" << "\n";
    for (int c = 0, idx = 0; c < nc; ++c) {
      ++idx;
      for (int d = 0; d < 6; ++d, ++idx) {
        const double sv = cs[6 * c + d], s0 = db.stress0[6 * c + d];
        o << setw(3) << c << std::fixed << setw(11) << setprecision(4) << db.weight[c] << setw(11) << db.sweight[c]
          << setw(15) << setprecision(6) << ev[idx] * ev[idx] << setw(18) << setprecision(10) << sv << setw(18)
          << s0;
        if (s0 == 0) o << setw(15) << "inf";
        else o << setw(15) << setprecision(6) << (s0 - sv) / s0;
        o << "\n";
      }
      idx += 3 * db.natoms[c];
    }
  }
  {  // forces (src/potential.cpp:597-641)
    std::ofstream o = open_out(numbered(base, ".force", num));
    o << std::right << setw(6) << "conf:" << setw(5) << "atom" << setw(8) << "type" << setw(15) << "force error"
      << setw(15) << "f" << setw(15) << "f0" << setw(15) << "df/f0" << "\n";
    for (int c = 0, idx = 0, a = 0; c < nc; ++c) {
      idx += 7;
      for (int i = 0; i < db.natoms[c]; ++i, ++a)
        for (int d = 0; d < 3; ++d, ++idx) {
          const double f = fr[3 * a + d], f0 = db.force0[3 * a + d];
          o << setw(5) << c << ":" << setw(5) << i << setw(8) << db.types[a] << setw(15) << std::fixed
            << ev[idx] * ev[idx] << setw(15) << f << setw(15) << f0;
          if (f0 == 0) o << setw(15) << "inf";
          else o << setw(15) << (f0 - f) / f0;
          o << "\n";
        }
    }
  }
  {  // error sums (src/potential.cpp:647-707)
    double se = 0, ss = 0, sf = 0, sp = 0;
    size_t idx = 0;
    for (int c = 0; c < nc; ++c) {
      se += ev[idx] * ev[idx];
      ++idx;
      for (int d = 0; d < 6; ++d, ++idx) ss += ev[idx] * ev[idx];
      for (int i = 0; i < 3 * db.natoms[c]; ++i, ++idx) sf += ev[idx] * ev[idx];
    }
    for (; idx < ev.size(); ++idx) sp += ev[idx] * ev[idx];
    const long npun = (long)ev.size() - 3L * natoms - 7L * nc;
    std::ofstream o = open_out(numbered(base, ".error", num));
    o << "Total error sum " << std::fixed << se + ss + sf + sp << " count " << ev.size() << " (" << 3 * natoms
      << " forces, " << 6 * nc << " stresses, " << nc << " energies";
    if (npun == 0) o << ")\n";
    else o << ", " << npun << " punishments)\n";
    o << "Sum of force errors  = " << sf << "\n";
    o << "Sum of energy errors = " << se << "\n";
    o << "Sum of stress errors = " << ss << "\n";
    if (npun) o << "Sum of punish errors = " << sp << "\n";
  }
  if (!pair) {  // constraints (src/pots/pot_eam/pot_eam.cpp:439-488, src/pots/pot_meam/pot_meam.cpp:444-490)
    std::ofstream o = open_out(numbered(base, ".punish", num));
    o << "Limiting constraints (atomic density lies inside F boundaries)\n";
    o << std::right << setw(3) << "#" << setw(11) << "conf_w" << setw(15) << "punish error" << setw(15)
      << "punishment" << "\n";
    size_t idx = 7 * (size_t)nc + 3 * (size_t)natoms;
    for (int c = 0; c < nc; ++c, ++idx)
      o << setw(3) << c << std::fixed << setw(11) << setprecision(4) << db.weight[c] << setw(15) << setprecision(6)
        << ev[idx] * ev[idx] << setw(15) << ev[idx] / db.weight[c] << "\n";
    const int nt = pot.ntypes();
    if (gmeam) {
      // one line per f spline (src/pots/pot_gmeam/pot_gmeam.cpp:246-292)
      const int f0 = nt * (nt + 1) / 2 + 2 * nt, nf = nt * (nt + 1) / 2;
      o << "\nDummy constraints (|f_max(r)| = 1)\n";
      o << setw(15) << "punish error" << setw(15) << "|f_max(r)|" << "\n";
      for (int q = f0; q < f0 + nf; ++q, ++idx) {
        double m = 0.0;
        for (double y : pot.splines()[q].y)
          if (std::abs(m) < std::abs(y)) m = y;
        o << std::fixed << setw(15) << setprecision(6) << ev[idx] * ev[idx] << setw(15) << std::abs(m) << "\n";
      }
    } else if (meam) {
      // signed knot value of largest magnitude over the f splines (src/pot_fns.cpp:100-111)
      const int f0 = nt * (nt + 1) / 2 + 2 * nt, nf = nt * (nt + 1) / 2;
      double fmax = 0.0;
      for (int q = f0; q < f0 + nf; ++q) {
        double m = 0.0;
        for (double y : pot.splines()[q].y)
          if (std::abs(m) < std::abs(y)) m = y;
        if (std::abs(fmax) < std::abs(m)) fmax = m;
      }
      o << "\nDummy constraints (|f_max(r)| = 1)\n";
      o << setw(15) << "punish error" << setw(15) << "|f_max(r)|" << "\n";
      o << std::fixed << setw(15) << setprecision(6) << ev[idx] * ev[idx] << setw(15) << std::abs(fmax) << "\n";
      // jmeam: the same for the p (and p2) splines (src/pots/pot_jmeam/pot_jmeam.cpp:357-368,
      // src/pots/pot_jmeam2/pot_jmeam2.cpp:381-405)
      for (int m = 0; m < njm; ++m) {
        ++idx;
        const int p0 = f0 + nf + nt + m * (nf + nt);
        double pmax = 0.0;
        for (int q = p0; q < p0 + nf; ++q) {
          double v = 0.0;
          for (double y : pot.splines()[q].y)
            if (std::abs(v) < std::abs(y)) v = y;
          if (std::abs(pmax) < std::abs(v)) pmax = v;
        }
        const std::string nm = m == 0 ? "p" : "p2";
        o << "\nDummy constraints (|" << nm << "_max(r)| = 1)\n";
        o << setw(15) << "punish error" << setw(15) << ("|" + nm + "_max(r)|") << "\n";
        o << std::fixed << setw(15) << setprecision(6) << ev[idx] * ev[idx] << setw(15) << std::abs(pmax) << "\n";
      }
    } else {
      o << "\nDummy constraints (F'(n_avg) = 0)\n";
      o << setw(7) << "element" << setw(15) << "punish error" << setw(15) << "F'(n_avg)" << "\n";
      for (int t = 0; t < nt; ++t, ++idx)
        o << setw(7) << t << std::fixed << setw(15) << setprecision(6) << ev[idx] * ev[idx] << setw(15)
          << ev[idx] / 100.0 << "\n";  // DUMMY_WEIGHT, src/pots/pot_pair/pot_pair.h:39
    }
  }
}

// ---------------------------------------------------------------- driver

int run_param_file(const std::string& param_path, std::ostream& out) {
  Params prm = Params::read(param_path);
  std::string dir;
  const auto slash = param_path.find_last_of('/');
  if (slash != std::string::npos) dir = param_path.substr(0, slash + 1);
  auto path = [&](const std::string& f) { return (!f.empty() && f[0] == '/') ? f : dir + f; };
  if (!prm.has("config") || !prm.has("startpot")) throw Error("Parameter not found in file: config / startpot");
  PotFile pf = read_potential_file(path(prm.str("startpot")));
  ConfigDb db = read_config_file(path(prm.str("config")));
  db.energy_weight = prm.num("energy_weight", 0.0);  // src/config.cpp:27-28 defaults
  db.stress_weight = prm.num("stress_weight", 0.0);
  const int extrap = (int)prm.num("embed_extrap", 0);
  auto ev = std::make_shared<Evaluator>(pf.style, pf.ntypes, pf.pots[0], extrap, db);
  out << "Total # atoms = " << ev->info().natoms << " inside a total # cells = " << ev->info().ncells << std::endl;
  std::vector<Potential*> pots;
  for (auto& spl : pf.pots) pots.push_back(new Potential(ev, pf.style, pf.ntypes, spl, extrap));
  for (size_t i = 0; i < pots.size(); ++i)
    out << "P#" << i << " - Initial total sum error = " << std::fixed << pots[i]->compute() << std::endl;
  OptOptions opt;
  opt.d_eps = prm.num("d_eps", 0.0);
  opt.max_steps = (int)prm.num("max_steps", 1000000);
  opt.anneal_temp = prm.num("anneal_temp", 1.0);
  opt.pop_size = (int)prm.num("pop_size", 10);
  opt.ngens = (int)prm.num("max_steps", 10);
  opt.cross_rate = prm.num("cross_rate", 0.1);
  opt.mut_rate = prm.num("mut_rate", 0.0);
  opt.init_scale = prm.num("init_scale", 10.0);
  opt.fit_frac = prm.num("fit_rate", 0.2);
  opt.lbfgs_mem = (int)prm.num("lbfgs_mem", 10);
  opt.seed = (unsigned)prm.num("seed", 0);
  opt.out = &out;
  std::mt19937 rng(opt.seed);
  for (const auto& st : prm.list("optstyle")) {
    rng.seed(opt.seed);
    const auto calls0 = Potential::get_ncalls();
    const auto t0 = std::chrono::steady_clock::now();
    auto summary = [&]() {
      double best = pots[0]->get_error();
      for (auto* p : pots) best = std::min(best, p->get_error());
      out << "optimizer " << st << ": " << Potential::get_ncalls() - calls0 << " objective calls, " << std::fixed
          << std::setprecision(3) << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count()
          << " s, best F = " << std::scientific << std::setprecision(10) << best << std::endl;
    };
    if (st == "genalg") {
      auto res = genalg(pots, opt, rng);
      for (auto* p : pots) delete p;
      pots = res;
      summary();
      continue;
    }
    for (auto*& p : pots) {
      Potential* np = nullptr;
      if (st == "powell") np = powell(*p, opt);
      else if (st == "lbfgs") np = lbfgs(*p, opt);
      else if (st == "simann") np = simann(*p, opt, rng);
      else throw Error("Invalid optimizer style - " + st);
      delete p;
      p = np;
    }
    summary();
  }
  PotFile endf = pf;
  endf.pots.clear();
  for (auto* p : pots) endf.pots.push_back(p->splines());
  if (prm.has("endpot")) write_potential_file(path(prm.str("endpot")), endf);
  if (prm.has("lammpsfile")) write_lammps_file(path(prm.str("lammpsfile")), endf);
  if (prm.has("datafile"))
    for (size_t i = 0; i < pots.size(); ++i)
      write_data_files(path(prm.str("datafile")), *pots[i], db, pots.size() > 1 ? (int)i : -1);
  for (size_t i = 0; i < pots.size(); ++i)
    out << "P#" << i << " - Final total sum error = " << std::fixed << pots[i]->compute() << std::endl;
  for (auto* p : pots) delete p;
  return 0;
}

}  // namespace mzh
