// Step kernels of the D3Q15 velocity set.
#include "lattice_launch.hpp"

LBF_LATTICE_TU(lbf::D3Q15, d3q15)
