// ep2d_cli -- command line of the reference (main.cu:1-42), same argv order:
//   ep2d_cli <init load value> <load type x> <load type y> <load type xy> <time steps> [<load value>]
// Reads data/pa.dat, runs ComputeEffModuli (NL <= 3) or ComputeEffDamping (NL > 3)
// and writes data/Pc_<nX>_.dat, data/tauXXc_<nX>_.dat and EffPlast2D.log.
// Configuration macros (NGRID, NL, NITER, EITER, DEVICE_IDX, EP2D_MODE) are
// compile-time, as in the reference build (eff_props_2D.m:45).
// The optional sixth argument is read only when present (main.cu:17-18 reads
// argv[6] whenever argc >= 6, i.e. also when it is missing).
#include <iostream>
#include <string>

#include "EffPlast2D.h"

int main(int argc, char** argv) {
  try {
    const std::string err_info = "ERROR:  missing arguments\n USAGE: " + std::string(argv[0]) +
      " <init load value> <init load type>.x <init load type>.y <init load type>.xy <time steps> [<load value>]";
    if (argc < 6) throw std::invalid_argument(err_info);
    const double init_load_value = std::stod(argv[1]);
    const std::array<double, 3> load_type = {std::stod(argv[2]), std::stod(argv[3]), std::stod(argv[4])};
    const unsigned int time_steps = (unsigned int)std::stod(argv[5]);
    double load_value = 0.0;
    if (argc >= 7) load_value = std::stod(argv[6]);

    EffPlast2D eff_plast(time_steps);
    if (NL <= 3)
      eff_plast.ComputeEffModuli(init_load_value, load_type, load_value);
    else
      eff_plast.ComputeEffDamping(init_load_value, load_type, load_value);
    return 0;
  } catch (const std::exception& e) {
    std::cerr << e.what() << "\n";
    return 1;
  }
}
