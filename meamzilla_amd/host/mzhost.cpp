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

// ---------------------------------------------------------------- least squares

namespace {

constexpr int kBatch = 64;  // knot vectors per batched call of the optimizers

// Knot vectors of several potentials, row-major, for one batched call.
std::vector<double> knot_rows(const std::vector<const Potential*>& ps) {
  std::vector<double> K;
  for (const Potential* p : ps) {
    const std::vector<double> kv = p->knot_vector();
    K.insert(K.end(), kv.begin(), kv.end());
  }
  return K;
}

// Objectives of several potentials through mz_objective_batch; each
// potential keeps its own (get_error).
std::vector<double> batch_objective(mz_ctx* ctx, const std::vector<Potential*>& ps) {
  std::vector<double> F(ps.size());
  if (ps.empty()) return F;
  std::vector<double> K = knot_rows(std::vector<const Potential*>(ps.begin(), ps.end()));
  check(mz_objective_batch(ctx, (int)ps.size(), K.data(), F.data(), nullptr));
  Potential::add_ncalls(ps.size());
  for (size_t i = 0; i < ps.size(); ++i) ps[i]->set_error(F[i]);
  return F;
}

// In-place Cholesky factor (lower triangle, row-major) of a symmetric positive
// definite matrix; false when a pivot is not positive.
bool cholesky(int n, std::vector<double>& M) {
  for (int j = 0; j < n; ++j) {
    double d = M[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= M[(size_t)j * n + k] * M[(size_t)j * n + k];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    M[(size_t)j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      double t = M[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) t -= M[(size_t)i * n + k] * M[(size_t)j * n + k];
      M[(size_t)i * n + j] = t / d;
    }
  }
  return true;
}

// Solves L L^T x = b for the factor of cholesky().
std::vector<double> cholesky_solve(int n, const std::vector<double>& L, std::vector<double> b) {
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < i; ++k) b[i] -= L[(size_t)i * n + k] * b[k];
    b[i] /= L[(size_t)i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    for (int k = i + 1; k < n; ++k) b[i] -= L[(size_t)k * n + i] * b[k];
    b[i] /= L[(size_t)i * n + i];
  }
  return b;
}

// Runs body(i) for i in [0, n) on up to 16 host threads.
template <class Body>
void parallel_for(int n, Body body) {
  const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (nth == 1 || n < 8) {
    for (int i = 0; i < n; ++i) body(i);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < nth; ++t)
    th.emplace_back([&]() {
      for (int i = next++; i < n; i = next++) body(i);
    });
  for (auto& x : th) x.join();
}

}  // namespace

// Levenberg-Marquardt on the residual vector (Levenberg, Quart. Appl. Math. 2
// (1944) 164-168; Marquardt, J. SIAM 11 (1963) 431-441, with Marquardt's
// diagonal scaling).  Every iteration forms the Jacobian of the residuals by
// forward differences -- all its columns from batched residual calls
// (mz_residuals_batch, the perturbed knot vectors as candidates) -- and the
// normal equations A = J^T J, g = J^T e; it then solves
// (A + lambda diag(A)) delta = -g for four damping values lambda * {1/10, 1, 10,
// 100}, scores the four trial potentials in one batched call and keeps the
// lowest.  The damping follows the ratio of the actual to the predicted
// decrease (F - |e + J delta|^2); a step that decreases nothing raises it a
// hundredfold.  The damping also keeps the system regular along the
// objective's scale gauges (rho -> a rho with U rescaled, f -> b f with g / b^2).
Potential* least_squares(const Potential& start, const OptOptions& opt) {
  std::ostream* out = opt.out;
  Potential* pot = start.clone();
  const int n = pot->get_ncoeff();
  ErrorVec e;
  double F = pot->compute(&e);
  if (opt.max_steps == 0 || n == 0) return pot;
  const size_t m = e.size();
  mz_ctx* ctx = pot->evaluator()->ctx();
  const int nk = pot->evaluator()->info().nknot;
  std::vector<double> J(m * (size_t)n), A((size_t)n * n), g(n);  // J column-major: J[i*m + r] = de_r / dx_i
  const double kFactors[4] = {0.1, 1.0, 10.0, 100.0};
  double lambda = -1.0;
  int fails = 0;
  const auto t0 = std::chrono::steady_clock::now();
  if (out) *out << "Levenberg-Marquardt least squares, " << n << " coefficients, " << m << " residuals" << std::endl
                << "  iter        F             lambda        calls   seconds" << std::endl;
  for (int it = 0; it < opt.max_steps; ++it) {
    if (F < 1e-12) {
      if (out) *out << "objective below 1e-12: nothing left to fit" << std::endl;
      break;
    }
    // Jacobian columns: x_i + h_i, h_i = 1e-7 max(|x_i|, 1)
    const std::vector<double> kv = pot->knot_vector();
    const std::vector<int>& slots = pot->coeff_slots();
    std::vector<double> kb, Fb, Eb;
    for (int i0 = 0; i0 < n; i0 += kBatch) {
      const int nb = std::min(kBatch, n - i0);
      kb.resize((size_t)nb * nk);
      std::vector<double> h(nb);
      for (int b = 0; b < nb; ++b) {
        std::copy(kv.begin(), kv.end(), kb.begin() + (size_t)b * nk);
        const int sl = slots[i0 + b];
        h[b] = 1e-7 * std::max(std::abs(kv[sl]), 1.0);
        kb[(size_t)b * nk + sl] += h[b];
      }
      Fb.resize(nb);
      Eb.resize((size_t)nb * m);
      if (mz_residuals_batch(ctx, nb, kb.data(), Fb.data(), Eb.data()) != MZ_OK)
        throw Error(std::string("mz_residuals_batch: ") + mz_last_error());
      Potential::add_ncalls(nb);
      for (int b = 0; b < nb; ++b) {
        double* col = J.data() + (size_t)(i0 + b) * m;
        const double* eb = Eb.data() + (size_t)b * m;
        for (size_t r = 0; r < m; ++r) col[r] = (eb[r] - e[r]) / h[b];
      }
    }
    // normal equations, one column of A per task
    parallel_for(n, [&](int i) {
      const double* ci = J.data() + (size_t)i * m;
      double s = 0.0;
      for (size_t r = 0; r < m; ++r) s += ci[r] * e[r];
      g[i] = s;
      for (int j = 0; j <= i; ++j) {
        const double* cj = J.data() + (size_t)j * m;
        double t = 0.0;
        for (size_t r = 0; r < m; ++r) t += ci[r] * cj[r];
        A[(size_t)i * n + j] = t;
        A[(size_t)j * n + i] = t;
      }
    });
    double dmax = 0.0;
    for (int i = 0; i < n; ++i) dmax = std::max(dmax, A[(size_t)i * n + i]);
    if (!(dmax > 0.0)) {
      if (out) *out << "the residuals do not depend on the coefficients: nothing to fit" << std::endl;
      break;
    }
    if (lambda < 0) lambda = 1e-3;
    // trial steps for the four damping values
    std::vector<std::vector<double>> steps;
    std::vector<double> lams, preds;
    std::vector<Potential*> trials;
    for (double fct : kFactors) {
      const double lam = lambda * fct;
      std::vector<double> M = A;
      for (int i = 0; i < n; ++i) M[(size_t)i * n + i] += lam * std::max(A[(size_t)i * n + i], 1e-12 * dmax);
      if (!cholesky(n, M)) continue;
      std::vector<double> mg(n);
      for (int i = 0; i < n; ++i) mg[i] = -g[i];
      std::vector<double> dlt = cholesky_solve(n, M, mg);
      // predicted decrease: -(2 g.delta + delta.A.delta)
      double gd = 0.0, dad = 0.0;
      for (int i = 0; i < n; ++i) {
        gd += g[i] * dlt[i];
        double t = 0.0;
        for (int j = 0; j < n; ++j) t += A[(size_t)i * n + j] * dlt[j];
        dad += dlt[i] * t;
      }
      Potential* tp = pot->clone();
      tp->set_dir(dlt, 1.0);
      trials.push_back(tp);
      steps.push_back(dlt);
      lams.push_back(lam);
      preds.push_back(-(2.0 * gd + dad));
    }
    const std::vector<double> Ft = batch_objective(ctx, trials);
    int best = -1;
    for (size_t k = 0; k < Ft.size(); ++k)
      if (std::isfinite(Ft[k]) && Ft[k] < F && (best < 0 || Ft[k] < Ft[best])) best = (int)k;
    const double Fold = F;
    if (best < 0) {
      for (auto* t : trials) delete t;
      lambda *= 100.0;
      if (++fails > 8 || lambda > 1e16) {
        if (out) *out << "no damping value decreases the objective: stopping" << std::endl;
        break;
      }
      continue;
    }
    fails = 0;
    delete pot;
    pot = trials[best];
    for (size_t k = 0; k < trials.size(); ++k)
      if ((int)k != best) delete trials[k];
    F = pot->compute(&e);
    const double ratio = preds[best] > 0 ? (Fold - F) / preds[best] : 0.0;
    lambda = lams[best] * (ratio > 0.75 ? 1.0 / 3.0 : (ratio < 0.25 ? 2.0 : 1.0));
    if (out)
      *out << std::setw(6) << it + 1 << "  " << std::scientific << std::setprecision(8) << F << "  " << lams[best]
           << "  " << std::setw(7) << Potential::get_ncalls() << "  " << std::fixed << std::setprecision(4)
           << std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() << std::endl;
    if (Fold - F < opt.d_eps) break;
    if (Fold - F <= 1e-12 * Fold) break;
  }
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

// Simulated annealing with the per-coordinate step control of Corana,
// Marchesi, Martini and Ridella (ACM Trans. Math. Softw. 13 (1987) 262-280):
// a move at coefficient c is a Gaussian bump of height N(0,1) v_c and width
// |N(0,1)| (Potential::modify_coeff), accepted by the Metropolis rule at
// temperature T.  After kCycles sweeps over the coefficients every v_c is
// scaled toward an acceptance fraction between 0.4 and 0.6 (factor
// 1 + 2 (a - 0.6) / 0.4 above, division by 1 + 2 (0.4 - a) / 0.4 below);
// after 2 n such adjustments T is multiplied by 0.85.  The run ends after
// max_steps temperatures, or once the objective at the end of the last four
// temperatures and the best one agree to 1e-3 relative; if they agree with one
// another but not with the best, the walk restarts from the best.
//
// The moves of a sweep are scored speculatively: the next moves (for the
// following coefficients) are all built from the current state and scored in
// one batched call; they are then tested in order, and the first accepted one
// becomes the new state, discarding the rest, which are rebuilt from it.  Each
// move is thus tested against the state current when it is tested, as in the
// sequential walk; only the random draws of the discarded moves are extra.  The
// number of moves per call follows the running acceptance fraction (about two
// expected acceptances per call, at most 32).
Potential* simann(const Potential& start, const OptOptions& opt, std::mt19937& rng) {
  constexpr int kCycles = 20;
  constexpr double kCool = 0.85;
  constexpr int kSettle = 4;
  constexpr double kSettleTol = 1e-3;
  constexpr int kMaxWidth = 32;
  std::normal_distribution<double> gauss(0.0, 1.0);
  std::uniform_real_distribution<double> unif(0.0, 1.0);
  Potential* cur = start.clone();
  double Fcur = cur->compute();
  const int n = cur->get_ncoeff();
  if (opt.max_steps == 0 || n == 0) return cur;
  mz_ctx* ctx = cur->evaluator()->ctx();
  Potential* best = cur->clone();
  double Fbest = Fcur;
  std::vector<double> v(n, 1.0);
  std::vector<int> acc(n, 0);
  double T = opt.anneal_temp;
  double rate = 0.5;  // running acceptance fraction
  std::vector<double> ends;  // objective at the end of every temperature
  const int adjustments = 2 * n;
  for (int temp = 0; temp < opt.max_steps; ++temp) {
    for (int adj = 0; adj < adjustments; ++adj) {
      for (int cyc = 0; cyc < kCycles; ++cyc) {
        int c = 0;
        while (c < n) {
          const int width = std::max(1, std::min({kMaxWidth, n - c, (int)std::ceil(2.0 / std::max(rate, 1.0 / 64))}));
          std::vector<Potential*> moves;
          std::vector<double> u(width);
          for (int b = 0; b < width; ++b) {
            Potential* mv = cur->clone();
            const double height = gauss(rng) * v[c + b];
            const double w = std::abs(gauss(rng));
            mv->modify_coeff(c + b, height, w);
            moves.push_back(mv);
            u[b] = unif(rng);
          }
          const std::vector<double> Fm = batch_objective(ctx, moves);
          int taken = -1;
          for (int b = 0; b < width && taken < 0; ++b) {
            const bool ok = Fm[b] <= Fcur || u[b] < std::exp((Fcur - Fm[b]) / T);
            rate = 0.98 * rate + 0.02 * (ok ? 1.0 : 0.0);
            if (ok) taken = b;
          }
          if (taken >= 0) {
            ++acc[c + taken];
            delete cur;
            cur = moves[taken];
            moves[taken] = nullptr;
            Fcur = Fm[taken];
            if (Fcur < Fbest) {
              delete best;
              best = cur->clone();
              Fbest = Fcur;
            }
            c += taken + 1;
          } else {
            c += width;
          }
          for (auto* mv : moves) delete mv;
        }
      }
      for (int k = 0; k < n; ++k) {
        const double a = (double)acc[k] / kCycles;
        if (a > 0.6) v[k] *= 1.0 + 2.0 * (a - 0.6) / 0.4;
        else if (a < 0.4) v[k] /= 1.0 + 2.0 * (0.4 - a) / 0.4;
        acc[k] = 0;
      }
    }
    if (opt.out)
      *opt.out << "temperature " << temp + 1 << "  T=" << std::scientific << std::setprecision(6) << T
               << "  F=" << Fcur << "  best F=" << Fbest << std::endl;
    T *= kCool;
    ends.push_back(Fcur);
    if ((int)ends.size() >= kSettle) {
      bool settled = true;
      for (int q = 1; q <= kSettle; ++q)
        settled = settled && std::abs(ends[ends.size() - q] - Fcur) <= kSettleTol * std::abs(Fcur);
      if (settled) {
        if (std::abs(Fcur - Fbest) <= kSettleTol * std::abs(Fcur)) break;
        delete cur;
        cur = best->clone();
        Fcur = Fbest;
      }
    }
  }
  delete cur;
  return best;
}

// ---------------------------------------------------------------- genetic algorithm

// Generational genetic algorithm over a population of knot vectors, scored
// with one batched call per generation.  The population is filled up to
// pop_size with copies of the first member carrying a Gaussian bump of height
// N(0,1) |x_c| init_scale at every coefficient; each generation keeps the best
// max(1, fit_rate P) members unchanged and breeds the rest: two parents chosen
// by binary tournaments, uniform crossover (each coefficient from the second
// parent with probability cross_rate), and with probability mut_rate per
// coefficient a bump of height N(0,1) |x_c| / 10.
std::vector<Potential*> genalg(const std::vector<Potential*>& pop_in, const OptOptions& opt, std::mt19937& rng) {
  std::normal_distribution<double> gauss(0.0, 1.0);
  std::uniform_real_distribution<double> unif(0.0, 1.0);
  std::vector<Potential*> pop;
  for (auto* p : pop_in) pop.push_back(p->clone());
  const int n = pop.front()->get_ncoeff();
  const int P = std::max<int>({opt.pop_size, (int)pop.size(), 2});
  while ((int)pop.size() < P) {
    Potential* p = pop.front()->clone();
    for (int c = 0; c < n; ++c) p->modify_coeff(c, gauss(rng) * std::abs(p->coeff(c)) * opt.init_scale,
                                               std::abs(gauss(rng)));
    pop.push_back(p);
  }
  mz_ctx* ctx = pop.front()->evaluator()->ctx();
  auto score_and_sort = [&]() {
    for (size_t i0 = 0; i0 < pop.size(); i0 += kBatch) {
      const size_t nb = std::min<size_t>(kBatch, pop.size() - i0);
      std::vector<Potential*> part(pop.begin() + i0, pop.begin() + i0 + nb);
      const std::vector<double> F = batch_objective(ctx, part);
      for (size_t b = 0; b < nb; ++b) pop[i0 + b]->set_error(std::isfinite(F[b]) ? F[b] : HUGE_VAL);
    }
    std::stable_sort(pop.begin(), pop.end(), [](Potential* a, Potential* b) { return *a < *b; });
  };
  auto tournament = [&](int avoid) {
    int a = (int)(unif(rng) * P), b = (int)(unif(rng) * P);
    a = std::min(a, P - 1);
    b = std::min(b, P - 1);
    int w = pop[a]->get_error() <= pop[b]->get_error() ? a : b;
    if (w == avoid) w = (w == a) ? b : a;
    return w;
  };
  score_and_sort();
  const int elite = std::max(1, std::min(P - 1, (int)std::lround(opt.fit_frac * P)));
  for (int gen = 0; gen < opt.ngens; ++gen) {
    std::vector<Potential*> next;
    for (int i = 0; i < elite; ++i) next.push_back(pop[i]->clone());
    while ((int)next.size() < P) {
      const int pa = tournament(-1), pb = tournament(pa);
      Potential* kid = pop[pa]->clone();
      for (int c = 0; c < n; ++c) {
        if (unif(rng) < opt.cross_rate) kid->coeff(c) = pop[pb]->coeff(c);
        if (unif(rng) < opt.mut_rate)
          kid->modify_coeff(c, gauss(rng) * std::abs(kid->coeff(c)) * 0.1, std::abs(gauss(rng)));
      }
      next.push_back(kid);
    }
    for (auto* p : pop) delete p;
    pop = next;
    score_and_sort();
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
      if (st == "powell") np = least_squares(*p, opt);
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
