// Step kernels of the D3Q19 velocity set.
#include "lattice_launch.hpp"

LBF_LATTICE_TU(lbf::D3Q19, d3q19)
