// derive2coh / derive2com on any grid (see generic_fam2d.hip).
#pragma once
#include "common.h"

namespace tara {

// True where the fused 2D compressible passes take the grid (powers of two, 16..256).
bool f2_fast_grid(int nx, int ny);
// The routines on host arrays in the reference layout, any grid; synchronous.
int gen_derive2coh_host(int nx, int ny, const tara_coh_params* prm, const double* const state[3],
                        double* const rhs[3]);
int gen_derive2com_host(int nx, int ny, const tara_com_params* prm, const double* const state[6],
                        double* const rhs[6]);

}  // namespace tara
