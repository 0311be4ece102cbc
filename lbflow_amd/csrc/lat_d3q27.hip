// Step kernels of the D3Q27 velocity set.
#include "lattice_launch.hpp"

LBF_LATTICE_TU(lbf::D3Q27, d3q27)
