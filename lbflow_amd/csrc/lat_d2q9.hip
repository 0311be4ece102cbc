// Step kernels of the D2Q9 velocity set.
#include "lattice_launch.hpp"

LBF_LATTICE_TU(lbf::D2Q9, d2q9)
