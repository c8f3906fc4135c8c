// C++ host layer over the libmeamz C-ABI: the potential / optimizer contract of
// the fitting program (Potential::compute, coeff, compute_dir, clone;
// Powell, simulated annealing, genetic algorithm) with every objective
// evaluation going through mz_objective / mz_objective_grad /
// mz_objective_batch, plus the L-BFGS optimizer that uses the adjoint
// gradient.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <ostream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/meamz.h"

namespace mzh {

using ErrorVec = std::vector<double>;

struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

// ---------------------------------------------------------------- file formats

// One "#B" block (src/spline3eq.cpp:54-106, src/spline5eq.cpp:57-116).
struct SplineRec {
  int basis = 3;
  int n = 0;
  double x0 = 0, xn = 0;
  int bc[2] = {-1, -1};      // -1 natural, 0 free, 1 fixed
  double grad[2] = {0, 0};
  double grad2[2] = {0, 0};
  std::vector<double> y;
  std::vector<int> fixed;
  int ncoeff() const;
};

struct PotFile {
  std::string comment;
  std::string style;
  int ntypes = 0;
  std::vector<std::vector<SplineRec>> pots;  // one entry per "#P" block
};

PotFile read_potential_file(const std::string& path);
void write_potential_file(const std::string& path, const PotFile& pf);
int splines_for_style(const std::string& style, int ntypes);

// Configuration database in the layout of mz_db.
struct ConfigDb {
  int ncells = 0;
  std::vector<int> natoms, types;
  std::vector<double> box, pos, force0, energy0, stress0, weight, eweight, sweight;
  double energy_weight = 0.0, stress_weight = 0.0;
};
ConfigDb read_config_file(const std::string& path);

// Parameter file: "key value..." lines, first match wins (src/input.h:47-77).
struct Params {
  std::map<std::string, std::vector<std::string>> kv;
  static Params read(const std::string& path);
  bool has(const std::string& k) const { return kv.count(k) != 0; }
  std::string str(const std::string& k, const std::string& def = "") const;
  double num(const std::string& k, double def) const;
  std::vector<std::string> list(const std::string& k) const;
};

// ---------------------------------------------------------------- evaluator

// One device context shared by every potential of a run (the reference keeps
// configurations and lists as global singletons, src/potential.h note on
// ownership).
class Evaluator {
 public:
  Evaluator(const std::string& style, int ntypes, const std::vector<SplineRec>& layout, int embed_extrap,
            const ConfigDb& db);
  ~Evaluator();
  Evaluator(const Evaluator&) = delete;
  Evaluator& operator=(const Evaluator&) = delete;
  mz_ctx* ctx() const { return ctx_; }
  const mz_info_t& info() const { return info_; }

 private:
  mz_ctx* ctx_ = nullptr;
  mz_info_t info_{};
};

// ---------------------------------------------------------------- potential

class Potential {
 public:
  Potential(std::shared_ptr<Evaluator> ev, const std::string& style, int ntypes, std::vector<SplineRec> spl,
            int embed_extrap);
  Potential* clone() const { return new Potential(*this); }

  int get_ncoeff() const { return ncoeff_; }
  double& coeff(int idx);                                    // src/potential.cpp:192-206
  void modify_coeff(int idx, double height, double width);   // src/potential.cpp:212-226, src/spline.cpp:267-296
  double compute(ErrorVec* error_vec = nullptr);             // src/potential.h:67
  // compute plus per-cell energies (per atom), stresses (per volume) and forces
  double compute_outputs(ErrorVec* ev, std::vector<double>* cell_energy, std::vector<double>* cell_stress,
                         std::vector<double>* forces);
  void compute_densities(std::vector<double>* dens);        // src/pots/pot_eam/pot_eam.cpp:395-433
  double compute_grad(std::vector<double>* grad);            // objective + d/dcoeff (adjoint)
  void set_dir(const std::vector<double>& dir, double x);    // src/potential.cpp:268-274
  double compute_dir(ErrorVec* ev, const std::vector<double>& dir, double x);  // :281-292
  double get_error() const { return error_sum_; }
  void set_error(double F) { error_sum_ = F; }
  static unsigned long long get_ncalls() { return ncalls_; }
  static void add_ncalls(unsigned long long n) { ncalls_ += n; }
  static void reset_ncalls() { ncalls_ = 0; }

  std::vector<double> knot_vector() const;
  void set_knot_vector(const std::vector<double>& kv);
  const std::vector<int>& coeff_slots() const { return slots_; }
  const std::vector<SplineRec>& splines() const { return spl_; }
  const std::string& style() const { return style_; }
  int ntypes() const { return ntypes_; }
  std::shared_ptr<Evaluator> evaluator() const { return ev_; }
  bool operator<(const Potential& o) const { return error_sum_ < o.error_sum_; }

 private:
  std::shared_ptr<Evaluator> ev_;
  std::string style_;
  int ntypes_ = 0, embed_extrap_ = 0;
  std::vector<SplineRec> spl_;
  std::vector<int> koff_;     // knot-vector offset of each spline
  std::vector<int> slots_;    // knot-vector slot of each coefficient
  int ncoeff_ = 0;
  double error_sum_ = 0.0;
  static unsigned long long ncalls_;
  double* slot_ptr(int slot);
};

// ---------------------------------------------------------------- optimizers

struct OptOptions {
  double d_eps = 0.0;          // least squares / L-BFGS: stop when a step improves less
  int max_steps = 1000000;     // least-squares and L-BFGS iterations, annealing temperatures
  double anneal_temp = 1.0;    // simulated annealing start temperature
  int pop_size = 10;           // genetic algorithm
  int ngens = 10;
  double cross_rate = 0.1, mut_rate = 0.0, init_scale = 10.0, fit_frac = 0.2;
  int lbfgs_mem = 10;
  unsigned seed = 0;
  std::ostream* out = nullptr;
};

// The optimizers of the driver.  They take the optstyle names of the
// reference's parameter files (powell, simann, genalg) and its parameter keys,
// but are methods of their own built on the batched entry points; the
// reference's own optimizers run on this objective through the binding of
// integration/.
//
// Least squares on the residual vector (optstyle powell): Levenberg-Marquardt
// with batched forward-difference Jacobians and four damping values scored per
// batched call.
Potential* least_squares(const Potential& start, const OptOptions& opt);
// Limited-memory BFGS on the adjoint gradient (no reference counterpart).
Potential* lbfgs(const Potential& start, const OptOptions& opt);
// Simulated annealing with Corana et al.'s step control, moves scored in
// batched calls.
Potential* simann(const Potential& start, const OptOptions& opt, std::mt19937& rng);
// Genetic algorithm (elitism, tournaments, uniform crossover, bump mutations)
// over a population scored with batched calls.
std::vector<Potential*> genalg(const std::vector<Potential*>& pop, const OptOptions& opt, std::mt19937& rng);

// LAMMPS export of every potential of a list (src/pot_list.cpp:333-361,
// src/pot_fns.cpp:360-368): comment line, style, then per spline its basis
// name, knot count, end slopes and knot lines x y y'' [y''''].
void write_lammps_file(const std::string& path, const PotFile& pf);
// Fitting data files <base>.rho/.energy/.stress/.force/.error/.punish[.num]
// (src/potential.cpp:325-342,497-707, src/pots/pot_eam/pot_eam.cpp:439-543,
// src/pots/pot_meam/pot_meam.cpp:444-490); num < 0 for a single potential.
void write_data_files(const std::string& base, Potential& pot, const ConfigDb& db, int num);

// Runs a parameter file the way `meamz -p <file>` does: initial objective,
// optimizer list, end potential, LAMMPS and data files, final objective.
// Returns 0 on success.
int run_param_file(const std::string& param_path, std::ostream& out);

}  // namespace mzh
