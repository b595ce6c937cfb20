// Host interface of the 2D compressible families: isothermal hydro
// (derive2coh) and compressible MHD (derive2com).
#pragma once
#include "common.h"
#include "csh2d.h"  // EventTimer

namespace tara {

enum F2Mode { F2_RHS = 0, F2_AB2 = 1 };

// Plan + scratch of one (nx, ny) grid.  Device spectra are kx-fastest with the
// row padded to ld = Nh rounded up to 8: element (i, j) at j*ld + i; padding
// columns are kept at zero.  kx, ky follow the reference's expression
// 2 pi (i-1)/Lx with the caller's pi and box (src/derive2coh.f95:59-60).
struct F2Plan {
    int nx = 0, ny = 0, nh = 0, ld = 0;
    size_t fsz = 0;            // elements per field (ld * ny)
    double pi = 0, lx = 0, ly = 0;
    DevBuf a;                  // inverse-y-transformed fields
    DevBuf b;                  // products after the x pass
    DevBuf c, d, k;            // com only: first-round x output, derivative fields, u / forcing spectra
    DevBuf kx, ky, imin;       // imin[j]: modes with i >= imin[j] are dealiased away
    DevBuf bad;                // one int: set when an AB2 update stores a NaN or Inf
    const double2* tw_y = nullptr;
    const double2* tw_x = nullptr;  // length nx, radix cap 4
};

struct CohPhys {
    double mux = 0, muy = 0, cs = 0;
};
struct ComPhys {
    double mu = 0, eta = 0, cs = 0, kf = 0, amp = 0;
};

// family: 0 = coh (3 fields), 1 = com (6 fields)
int f2_plan_init(F2Plan& p, int family, int nx, int ny, double pi, double lx, double ly);
// One RHS on st (device, padded).  F2_RHS writes out[]; F2_AB2 advances st in
// place with history hist[] (f_{n-1} in, f_n out).
int coh_rhs_stage(const F2Plan& p, const CohPhys& ph, double2* const st[3], int mode, double dt,
                  double2* const out[3], hipStream_t s, EventTimer* tm);
int com_rhs_stage(const F2Plan& p, const ComPhys& ph, double2* const st[6], int mode, double dt,
                  double2* const out[6], hipStream_t s, EventTimer* tm);
// Pointwise updates over n complex elements: AB2, Euler, and the degenerate
// "RK4" / predictor-corrector forms of the coh steppers.
enum F2Point { F2P_AB2 = 0, F2P_EULER = 1, F2P_RK4_DEGEN = 2, F2P_PC_DEGEN = 3 };
int f2_pointwise(int kind, const double2* w, double2* wn, const double2* fo, const double2* fn, double dt, size_t n,
                 hipStream_t st);

}  // namespace tara
