// Host interface of the 3D compressible isothermal MHD pipeline.
#pragma once
#include <vector>

#include "common.h"
#include "csh2d.h"  // EventTimer

namespace tara {

enum MhdMode { MHD_RHS = 0, MHD_AB2 = 1 };

// Plan + scratch of one (nx, ny, nz) grid.  Device spectra use the reference's
// kx-fastest layout with the row length padded to ld = Nh rounded up to 8,
// element (i, j, k) at (k*Ny + j)*ld + i; padding columns are kept at zero.
struct MhdPlan {
    int nx = 0, ny = 0, nz = 0, nh = 0, ld = 0;
    double lx = 0, ly = 0, lz = 0, mu = 0, eta = 0, cs = 0;
    size_t fsz = 0;            // elements per field (ld * ny * nz)
    DevBuf a;                  // 7 fields: inverse-transformed state (rho, rho u, B)
    DevBuf b;                  // 14 fields: u (3), momentum fluxes (8), induction fluxes (3)
    DevBuf kx, ky, kz, imin;
    DevBuf bad;                // int: set when an AB2 step's new state holds a non-finite value
    const double2* tw_y = nullptr;
    const double2* tw_z = nullptr;
    const double2* tw_z4 = nullptr;  // radix cap 4 (z forward pass at 4 elements per thread)
    const double2* tw_x = nullptr;  // length nx, radix cap 4 (x pass)
};

int mhd_plan_init(MhdPlan& p, int nx, int ny, int nz, double lx, double ly, double lz, double mu, double eta,
                  double cs);
// One right-hand side on the 7 live state fields st[0..6] = rho, rho ux, rho uy,
// rho uz, Bx, By, Bz (device, padded layout).  mode MHD_RHS writes the 7 RHS
// fields to out[0..6]; MHD_AB2 advances st in place with history hist[0..6]
// (f_{n-1} in, f_n out).
int mhd_rhs_stage(const MhdPlan& p, double2* const st[7], int mode, double dt, double2* const out[7],
                  double2* const hist[7], hipStream_t st_, EventTimer* tm);

}  // namespace tara
