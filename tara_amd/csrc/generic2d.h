// derive2csh / derive2css and their RK4 step on any grid (see generic2d.hip).
#pragma once
#include "common.h"
#include "csh2d.h"

namespace tara {

// True where the fused passes take the grid (powers of two, 32..4096 x 16..4096).
bool csh_fast_grid(int nx, int ny);
// RHS of derive2csh (ph.screened == 0) or derive2css on device half spectra
// of pitch spec_ld(nx), kx fastest; synchronous on st.
int gen_csh_rhs(int nx, int ny, const CshPhys& ph, const double2* w, double2* f, hipStream_t st);
// One classic RK4 step of that RHS, in place; synchronous on st.
int gen_csh_rk4(int nx, int ny, const CshPhys& ph, double dt, double2* w, hipStream_t st);

}  // namespace tara
