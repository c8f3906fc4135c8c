// The reference fitting program with its spline potentials evaluated by
// libmeamz on the GPU (see pot_gpu.h).
//
//   mpiexec -n 1 meamz_gpu -p <param file>           the reference's own run
//       (Meamzilla::run, src/meamzilla.cpp:106-167: initial objective,
//       rescale, the optimizer list of the parameter file -- the reference's
//       Powell, simulated annealing and genetic algorithm --, end potential,
//       LAMMPS and data files, final objective)
//   mpiexec -n 1 meamz_gpu -p <param file> --eval    one evaluation of every
//       potential through Potential::compute, printed as JSON (F, error
//       vector, per-cell energies and stresses, forces, densities)
//
// The potential file keeps its reference style (meam/spline or eam/spline):
// the reference tree is read-only here, so instead of the factory line a
// maintainer would add (PotentialStyle(meam/gpu, PotMEAMGpu)), the potentials
// the reference read are replaced in the list by GPU-evaluated copies before
// the run starts.
#include <mpi.h>

#include <cstdio>
#include <cstring>
#include <typeinfo>
#include <vector>

#include "meamzilla.h"
#include "pot_gpu.h"
#include "pot_list.h"
#include "pots/pot_eam_spline/pot_eam_spline.h"
#include "pots/pot_meam_spline/pot_meam_spline.h"
#include "universe.h"

using namespace MEAMZ_NS;

namespace {

typedef PotGpu<PotMEAMSpline, MZ_STYLE_MEAM_SPLINE> PotMEAMGpu;
typedef PotGpu<PotEAMSpline, MZ_STYLE_EAM_SPLINE> PotEAMGpu;

// Densities through the protected PotEAM hook, as write_extras calls it.
struct DensHook : PotEAM {
  static void (PotEAM::*get())(const Comm&, Vector<double>*) { return &DensHook::compute_densities; }
};

Potential* gpu_copy(Potential* p) {
  if (typeid(*p) == typeid(PotMEAMSpline)) return new PotMEAMGpu(*static_cast<PotMEAMSpline*>(p));
  if (typeid(*p) == typeid(PotEAMSpline)) return new PotEAMGpu(*static_cast<PotEAMSpline*>(p));
  return nullptr;
}

void print_vec(const char* name, const std::vector<double>& v) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
  std::printf("],\n");
}

}  // namespace

int main(int argc, char** argv) {
  MPI_Init(&argc, &argv);
  bool eval = false;
  std::vector<char*> args;
  for (int i = 0; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--eval")) eval = true;
    else args.push_back(argv[i]);
  }
  int status = 0;
  {
    Meamzilla* mmz = new Meamzilla(static_cast<int>(args.size()), args.data(), MPI_COMM_WORLD);
    PotList& plist = *mmz->potlist;
    for (int p = 0; p < plist.get_group_npots(); ++p) {
      Potential* g = gpu_copy(plist[p]);
      if (!g) {
        std::fprintf(stderr, "ERROR: meamz_gpu evaluates meam/spline and eam/spline potentials\n");
        status = 1;
        break;
      }
      delete plist[p];
      plist[p] = g;
    }
    if (!status) {
      try {
        if (eval) {
          const Comm& comm = mmz->universe->comm_inter;
          std::printf("{\"pots\": [\n");
          for (int p = 0; p < plist.get_group_npots(); ++p) {
            Potential* pot = plist[p];
            ErrorVec ev;
            const double F = pot->compute(comm, &ev);
            std::vector<double> ce, cs, fr, dv;
            for (Cell* c : mmz->config->cells) {
              ce.push_back(c->energy);
              for (int j = 0; j < 6; ++j) cs.push_back(c->stress[j]);
              for (Atom* a : c->atoms) {
                fr.push_back(a->force.x);
                fr.push_back(a->force.y);
                fr.push_back(a->force.z);
              }
            }
            (static_cast<PotEAM*>(pot)->*DensHook::get())(comm, &dv);
            std::printf("{\"F\": %.17g,\n", F);
            print_vec("errvec", ev);
            print_vec("cell_energy", ce);
            print_vec("cell_stress", cs);
            print_vec("forces", fr);
            print_vec("densities", dv);
            std::printf("\"ncalls\": %llu}%s\n", Potential::get_ncalls(),
                        p + 1 < plist.get_group_npots() ? "," : "");
          }
          std::printf("]}\n");
          std::fflush(stdout);
        } else {
          mmz->run();
        }
      } catch (Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        status = 1;
      }
    }
    delete mmz;
  }
  gpu_detail::shared_ctx().reset();  // release the device context before MPI and HIP shut down
  MPI_Finalize();
  return status;
}
