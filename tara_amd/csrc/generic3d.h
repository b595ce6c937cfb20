// derive3com on any grid (see generic3d.hip).
#pragma once
#include "common.h"

namespace tara {

// True where the fused 3D passes take the grid (powers of two, 16..256 per axis).
bool mhd_fast_grid(int nx, int ny, int nz);
// derive3com on host arrays in the reference layout, any grid; synchronous.
int gen_derive3com_host(int nx, int ny, int nz, const tara_mhd_params* prm, const double* const state[8],
                        double* const rhs[8]);

}  // namespace tara
