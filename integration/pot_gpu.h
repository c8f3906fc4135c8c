// GPU-evaluated spline potentials for the reference fitting program.
//
// This is the binding a maintainer adds to the reference tree to evaluate its
// objective with libmeamz (as src/pots/pot_meam_gpu/pot_meam_gpu.h, registered
// with PotentialStyle(meam/gpu, PotMEAMGpu) in src/style_pot.h:12-23 so that
// the factory in src/pot_list.cpp:115-126 creates it).  PotGpu<Base, Style>
// derives from one of the reference's spline potentials and replaces the two
// evaluation calls of the Potential interface:
//   compute           virtual double Potential::compute(const Comm&, ErrorVec*)  src/potential.h:67
//   compute_densities virtual void PotEAM::compute_densities(const Comm&, Vector<double>*)
//                                                                    src/pots/pot_eam/pot_eam.h:45
// with one call each into the C-ABI of include/meamz.h.  Everything else
// (reading and writing potentials, the coefficient walk coeff/modify_coeff,
// rescale, clone semantics, the optimizers) stays the reference's own.
//
// The configuration database is a singleton of the program (mmz->config), so
// one device context serves every potential of the list and all their clones;
// it is created on the first evaluation from the database and the layout of
// the first potential evaluated.  compute leaves behind what the reference's
// leaves (src/potential.cpp:399-405,455-478): per-cell energy per atom,
// stress per volume and per-atom forces in mmz->config, the error vector on
// the root rank, error_sum_ and the call count.  One MPI rank drives one GPU.
#ifndef MEAMZ_POT_GPU_H
#define MEAMZ_POT_GPU_H

#include <memory>
#include <string>
#include <vector>

#include "atom.h"
#include "cell.h"
#include "config.h"
#include "error.h"
#include "meamz.h"
#include "meamzilla.h"
#include "pot_list.h"
#include "spline3eq.h"
#include "spline5eq.h"

namespace MEAMZ_NS {

namespace gpu_detail {

// Read access to the protected spline records (knot values, boundary slopes
// and flags) through pointers to members formed in derived classes.
struct SplineRecs : Spline3Eq {
  static std::vector<double> Spline::*y() { return &SplineRecs::y_; }
  static std::vector<double> Spline::*x() { return &SplineRecs::x_; }
  static int (Spline::*nat())[2] { return &SplineRecs::is_nat_; }
  static double (Spline3Eq::*grad())[2] { return &SplineRecs::grad_; }
  static int (Spline3Eq::*grad_fixed())[2] { return &SplineRecs::is_grad_fixed_; }
};
struct Spline5Recs : Spline5Eq {
  static double (Spline5Eq::*grad2())[2] { return &Spline5Recs::grad2_; }
};

struct CtxDeleter {
  void operator()(mz_ctx* c) const { mz_destroy(c); }
};

// The program's one device context (the database is a program singleton).
inline std::shared_ptr<mz_ctx>& shared_ctx() {
  static std::shared_ptr<mz_ctx> ctx;
  return ctx;
}

inline void check(int rc, const char* what) {
  if (rc != MZ_OK) throw Error(std::string("ERROR: libmeamz ") + what + ": " + mz_last_error() + " - ");
}

}  // namespace gpu_detail

template <class Base, int Style>
class PotGpu : public Base {
 public:
  PotGpu(class Meamzilla* mmz, int ntypes) : Base(mmz, ntypes) {}
  explicit PotGpu(const Base& b) : Base(b) {}
  virtual ~PotGpu() {}

  virtual PotGpu* clone() const { return new PotGpu(*this); }

  virtual double compute(const Comm& comm, ErrorVec* error_vec = nullptr) {
    require_one_rank(comm);
    mz_ctx* ctx = context();
    mz_info_t info;
    gpu_detail::check(mz_info(ctx, &info), "mz_info");
    const std::vector<double> kv = knot_vector();
    std::vector<double> ev(info.nerr), ce(info.ncells), cs(6 * (size_t)info.ncells), fr(3 * (size_t)info.natoms);
    double F = 0.0;
    gpu_detail::check(mz_objective(ctx, kv.data(), &F, ev.data(), ce.data(), cs.data(), fr.data()), "mz_objective");
    // what the reference's compute leaves in the database (src/potential.cpp:455-478)
    size_t a = 0;
    int c = 0;
    for (Cell* cell : this->mmz->config->cells) {
      cell->energy = ce[c];
      for (int j = 0; j < 6; ++j) cell->stress[j] = cs[6 * (size_t)c + j];
      for (Atom* atom : cell->atoms) {
        atom->force.x = fr[3 * a];
        atom->force.y = fr[3 * a + 1];
        atom->force.z = fr[3 * a + 2];
        ++a;
      }
      ++c;
    }
    if (error_vec && comm.is_root()) error_vec->swap(ev);
    this->error_sum_ = F;
    ++Potential::ncalls_;
    return F;
  }

 protected:
  virtual void compute_densities(const Comm& comm, Vector<double>* density_ptr) {
    require_one_rank(comm);
    mz_ctx* ctx = context();
    mz_info_t info;
    gpu_detail::check(mz_info(ctx, &info), "mz_info");
    const std::vector<double> kv = knot_vector();
    std::vector<double> dens(info.natoms);
    gpu_detail::check(mz_densities(ctx, kv.data(), dens.data()), "mz_densities");
    if (comm.is_root()) density_ptr->swap(dens);
  }

 private:
  static void require_one_rank(const Comm& comm) {
    if (comm.get_size() != 1) throw Error("ERROR: a GPU potential runs on one MPI rank per GPU - ");
  }

  // The splines in file order (phi, rho, U, f, g; alloy members in index order).
  std::vector<Spline*> splines() const {
    std::vector<Spline*> out;
    for (const PotFns& pf : this->pot_fns_)
      for (Basis* b : pf.fns) {
        Spline* s = dynamic_cast<Spline*>(b);
        if (!s) throw Error("ERROR: GPU potentials evaluate spline3eq / spline5eq bases only - ");
        out.push_back(s);
      }
    return out;
  }

  // Knot vector of the C-ABI: per spline y_0..y_{n-1}, grad_L, grad_R, grad2_L, grad2_R.
  std::vector<double> knot_vector() const {
    using R = gpu_detail::SplineRecs;
    std::vector<double> kv;
    for (Spline* s : splines()) {
      const std::vector<double>& y = s->*R::y();
      kv.insert(kv.end(), y.begin(), y.end());
      Spline3Eq* s3 = dynamic_cast<Spline3Eq*>(s);
      if (!s3) throw Error("ERROR: GPU potentials evaluate spline3eq / spline5eq bases only - ");
      const double(&g)[2] = s3->*R::grad();
      kv.push_back(g[0]);
      kv.push_back(g[1]);
      Spline5Eq* s5 = dynamic_cast<Spline5Eq*>(s);
      if (s5) {
        const double(&g2)[2] = s5->*gpu_detail::Spline5Recs::grad2();
        kv.push_back(g2[0]);
        kv.push_back(g2[1]);
      } else {
        kv.push_back(0.0);
        kv.push_back(0.0);
      }
    }
    return kv;
  }

  mz_ctx* context() const {
    std::shared_ptr<mz_ctx>& ctx = gpu_detail::shared_ctx();
    if (ctx) return ctx.get();
    using R = gpu_detail::SplineRecs;
    // layout of this potential (what read_basis parsed from the #B blocks)
    std::vector<int> basis, nk, bc;
    std::vector<double> x0, xn;
    for (Spline* s : splines()) {
      const std::vector<double>& x = s->*R::x();
      basis.push_back(dynamic_cast<Spline5Eq*>(s) ? 5 : 3);
      nk.push_back((int)x.size());
      x0.push_back(x.front());
      xn.push_back(x.back());
      Spline3Eq* s3 = dynamic_cast<Spline3Eq*>(s);
      const int(&nat)[2] = s->*R::nat();
      const int(&fixed)[2] = s3->*R::grad_fixed();
      for (int side = 0; side < 2; ++side) bc.push_back(nat[side] ? -1 : (fixed[side] ? 1 : 0));
    }
    mz_layout L;
    L.style = Style;
    L.ntypes = this->mmz->potlist->get_ntypes();  // Potential's own ntypes_ is left unset by its constructor
    L.nspl = (int)basis.size();
    L.basis = basis.data();
    L.nknots = nk.data();
    L.x0 = x0.data();
    L.xn = xn.data();
    L.bc = bc.data();
    L.embed_extrap = this->embed_extrap_;  // PotEAM reads it from the parameter file
    // the configuration database (src/cell.cpp:39-136 read it into mmz->config)
    const Config& cfg = *this->mmz->config;
    std::vector<int> natoms, types;
    std::vector<double> box, pos, f0, e0, s0, w, ew, sw;
    for (Cell* cell : cfg.cells) {
      natoms.push_back(cell->natoms);
      for (int k = 0; k < 3; ++k) {
        box.push_back(cell->a[k].x);
        box.push_back(cell->a[k].y);
        box.push_back(cell->a[k].z);
      }
      e0.push_back(cell->energy0);
      for (int j = 0; j < 6; ++j) s0.push_back(cell->stress0[j]);
      w.push_back(cell->weight);
      ew.push_back(cell->eweight);
      sw.push_back(cell->sweight);
      for (Atom* atom : cell->atoms) {
        types.push_back(atom->typ);
        pos.push_back(atom->pos.x);
        pos.push_back(atom->pos.y);
        pos.push_back(atom->pos.z);
        f0.push_back(atom->force0.x);
        f0.push_back(atom->force0.y);
        f0.push_back(atom->force0.z);
      }
    }
    mz_db db;
    db.ncells = (int)natoms.size();
    db.natoms = natoms.data();
    db.box = box.data();
    db.types = types.data();
    db.pos = pos.data();
    db.force0 = f0.data();
    db.energy0 = e0.data();
    db.stress0 = s0.data();
    db.weight = w.data();
    db.eweight = ew.data();
    db.sweight = sw.data();
    db.energy_weight = cfg.get_energy_weight();
    db.stress_weight = cfg.get_stress_weight();
    mz_ctx* c = nullptr;
    gpu_detail::check(mz_create(&L, &db, &c), "mz_create");
    ctx.reset(c, gpu_detail::CtxDeleter());
    return c;
  }
};

}  // namespace MEAMZ_NS

#endif  // MEAMZ_POT_GPU_H
