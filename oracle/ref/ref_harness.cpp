// Test-infrastructure harness around the reference program (built from the
// reference's own sources by oracle/ref/Makefile into oracle/_ref/).  It is a
// checker only: it reads a reference parameter file, evaluates every
// potential of the list through the reference's virtual
// Potential::compute(const Comm&, ErrorVec*) (src/potential.h:67) and prints
// the results with full double precision as JSON on rank 0.
//
//   meamz_ref -p <param file> [--fd <h>] [--time <n>] [--dens] [--write]
//
//   --fd h    also print the central-difference gradient of the objective over
//             every adjustable coefficient, step h*max(|c|,1)
//   --time n  time n objective evaluations (all ranks) and print evals/s
//   --dens    also print per-atom densities (PotEAM::compute_densities)
//   --write   afterwards write the LAMMPS file and the data files named by the
//             parameter file's lammpsfile / datafile keys, as the program does
//             at the end of a run (src/meamzilla.cpp:149-150)
#include "meamzilla.h"
#include "universe.h"
#include "config.h"
#include "cell.h"
#include "atom.h"
#include "pot_list.h"
#include "potential.h"
#include "pots/pot_eam/pot_eam.h"

#include <mpi.h>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <string>
#include <vector>

using namespace MEAMZ_NS;

namespace {

// Pointer-to-member access to the protected density routine of PotEAM.
struct DensAccess : PotEAM {
  static void (PotEAM::*get())(const Comm&, Vector<double>*) { return &DensAccess::compute_densities; }
};

void print_vec(const char* name, const std::vector<double>& v, bool comma = true) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
  std::printf("]%s\n", comma ? "," : "");
}

}  // namespace

int main(int argc, char** argv) {
  MPI_Init(&argc, &argv);
  double fd_h = 0.0;
  int ntime = 0;
  bool dens = false, write = false;
  std::vector<char*> args;
  args.push_back(argv[0]);
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--fd") && i + 1 < argc) fd_h = std::atof(argv[++i]);
    else if (!std::strcmp(argv[i], "--time") && i + 1 < argc) ntime = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "--dens")) dens = true;
    else if (!std::strcmp(argv[i], "--write")) write = true;
    else args.push_back(argv[i]);
  }
  int rank = 0;
  MPI_Comm_rank(MPI_COMM_WORLD, &rank);
  {
    Meamzilla mmz(static_cast<int>(args.size()), args.data(), MPI_COMM_WORLD);
    const Comm& comm = mmz.universe->comm_inter;
    PotList& plist = *mmz.potlist;
    if (rank == 0) std::printf("{\"pots\": [\n");
    for (int p = 0; p < plist.get_group_npots(); ++p) {
      Potential* pot = plist[p];
      ErrorVec ev;
      double F = pot->compute(comm, &ev);
      std::vector<double> ce, cs, fr;
      for (Cell* c : mmz.config->cells) {
        ce.push_back(c->energy);
        cs.push_back(c->stress.xx); cs.push_back(c->stress.yy); cs.push_back(c->stress.zz);
        cs.push_back(c->stress.xy); cs.push_back(c->stress.yz); cs.push_back(c->stress.zx);
        for (Atom* a : c->atoms) { fr.push_back(a->force.x); fr.push_back(a->force.y); fr.push_back(a->force.z); }
      }
      std::vector<double> dv;
      if (dens) {
        PotEAM* pe = dynamic_cast<PotEAM*>(pot);
        if (pe) (pe->*DensAccess::get())(comm, &dv);
      }
      std::vector<double> coeffs;
      for (int i = 0; i < pot->get_ncoeff(); ++i) coeffs.push_back(pot->coeff(i));
      std::vector<double> grad;
      if (fd_h > 0) {
        for (int i = 0; i < pot->get_ncoeff(); ++i) {
          double c = pot->coeff(i);
          double h = fd_h * (std::abs(c) > 1.0 ? std::abs(c) : 1.0);
          pot->coeff(i) = c + h;
          double fp = pot->compute(comm);
          pot->coeff(i) = c - h;
          double fm = pot->compute(comm);
          pot->coeff(i) = c;
          grad.push_back((fp - fm) / (2 * h));
        }
      }
      double evals_per_s = 0.0;
      if (ntime > 0) {
        MPI_Barrier(MPI_COMM_WORLD);
        double t0 = MPI_Wtime();
        for (int i = 0; i < ntime; ++i) pot->compute(comm);
        MPI_Barrier(MPI_COMM_WORLD);
        double t1 = MPI_Wtime();
        evals_per_s = ntime / (t1 - t0);
      }
      if (rank == 0) {
        std::printf("{\"F\": %.17g,\n\"ncoeff\": %d,\n", F, pot->get_ncoeff());
        print_vec("errvec", ev);
        print_vec("cell_energy", ce);
        print_vec("cell_stress", cs);
        print_vec("forces", fr);
        print_vec("densities", dv);
        print_vec("coeffs", coeffs);
        print_vec("fd_grad", grad);
        std::printf("\"fd_h\": %.17g, \"nranks\": %d, \"evals_per_s\": %.17g}%s\n", fd_h,
                    comm.get_size(), evals_per_s, p + 1 < plist.get_group_npots() ? "," : "");
      }
    }
    if (rank == 0) std::printf("]}\n");
    std::fflush(stdout);
    if (write) {
      if (!plist.get_lmp_file().empty()) plist.write_lmps(plist.get_lmp_file());
      if (!plist.get_dat_file().empty()) plist.write_data(plist.get_dat_file());
    }
  }
  MPI_Finalize();
  return 0;
}
