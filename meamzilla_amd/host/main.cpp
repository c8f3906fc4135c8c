// meamz_amd: the fitting driver over libmeamz.
//
//   meamz_amd -p <parameter file>     initial objective, optimizer list
//                                     ("optstyle powell lbfgs simann genalg"),
//                                     end potential, final objective
//   meamz_amd --check <parameter file> reads the parameter, potential and
//                                     configuration files and prints their
//                                     sizes and coefficients (no device use)
#include <cstring>
#include <iomanip>
#include <iostream>

#include "mzhost.h"

int main(int argc, char** argv) {
  if (argc == 3 && (!std::strcmp(argv[1], "-p") || !std::strcmp(argv[1], "-param"))) {
    try {
      return mzh::run_param_file(argv[2], std::cout);
    } catch (const std::exception& e) {
      std::cerr << "ERROR: " << e.what() << std::endl;
      return 1;
    }
  }
  if (argc == 3 && !std::strcmp(argv[1], "--check")) {
    try {
      std::string param = argv[2];
      const auto slash = param.find_last_of('/');
      const std::string dir = slash == std::string::npos ? "" : param.substr(0, slash + 1);
      mzh::Params prm = mzh::Params::read(param);
      mzh::PotFile pf = mzh::read_potential_file(dir + prm.str("startpot"));
      mzh::ConfigDb db = mzh::read_config_file(dir + prm.str("config"));
      int ncoeff = 0;
      for (const auto& s : pf.pots[0]) ncoeff += s.ncoeff();
      std::cout << "style " << pf.style << " ntypes " << pf.ntypes << " splines " << pf.pots[0].size() << " ncoeff "
                << ncoeff << " cells " << db.ncells << " atoms " << db.types.size() << "\n";
      // coefficient values in the walk order, without a device context
      std::vector<double> vals;
      for (const auto& s : pf.pots[0]) {
        for (int k = 0; k < s.n; ++k)
          if (!s.fixed[k]) vals.push_back(s.y[k]);
        for (int side = 0; side < 2; ++side)
          if (s.bc[side] == 0) {
            vals.push_back(s.grad[side]);
            if (s.basis == 5) vals.push_back(s.grad2[side]);
          }
      }
      vals.resize(ncoeff);
      std::cout << "coeffs";
      for (double v : vals) std::cout << " " << std::setprecision(17) << v;
      std::cout << "\n";
      return 0;
    } catch (const std::exception& e) {
      std::cerr << "ERROR: " << e.what() << std::endl;
      return 1;
    }
  }
  std::cerr << "usage: " << argv[0] << " -p <parameter file> | --check <parameter file>" << std::endl;
  return 2;
}
