// Host interface of the 2D incompressible (csh) pipeline.
#pragma once
#include <vector>

#include "common.h"

namespace tara {

enum CshMode { CSH_RHS = 0, CSH_AB2 = 1, CSH_RK4_1 = 2, CSH_RK4_2 = 3, CSH_RK4_3 = 4, CSH_RK4_4 = 5 };

// Brackets launches with HIP events on the launch stream, per kernel class.
// Events are recycled from a pool; durations are summed when collected.
struct EventTimer {
    static constexpr int kClasses = 8;
    struct Pair {
        hipEvent_t a, b;
        int cls;
    };
    std::vector<Pair> live;
    std::vector<Pair> pool;
    Pair cur[kClasses];
    long counts[kClasses] = {0};
    double ms[kClasses] = {0};
    void begin(int cls, hipStream_t st);
    void end(int cls, hipStream_t st);
    int collect();  // synchronises the recorded events, accumulates counts/ms
    ~EventTimer();
};

// Elements per thread (also the radix cap) of the y transforms for column
// length ny: 4 at ny = 1024, where a single run's 513 columns at 8 elements
// per thread would fill one wave per SIMD; 8 otherwise.
__host__ __device__ constexpr int csh_yem(int ny) { return ny == 1024 ? 4 : 8; }

// Shape of the inverse y pass (k_csh_yinv) for column length ny: whether two
// columns share a workgroup interleaved across its lanes (ny = 4096, 512
// threads per column), and how many columns a workgroup holds (otherwise
// contiguous, ny/csh_yem(ny) threads each, 256 threads at least).
__host__ __device__ constexpr bool csh_yinv_interleaved(int ny, bool screened)
{
    return !screened && ny == 4096;
}
__host__ __device__ constexpr int csh_yinv_cols(int ny, bool screened)
{
    const int t = ny / (ny < csh_yem(ny) ? ny : csh_yem(ny));
    return csh_yinv_interleaved(ny, screened) ? 2 : (t >= 256 ? 1 : 256 / t);
}
// Layout of the intermediates G0..G3 (see blkg), a function of nh = nx/2 + 1
// only, so that the x pass (templated on nx) has it at compile time: log2 of
// the rows per group, and columns per group.  Nx = 4096 (the interleaved
// inverse y pass at Ny = 4096): 2-row groups of 2050 columns, an odd multiple
// of 64 bytes (2052 columns measured 0.206 against 0.1945 ms for the 4096^2
// inverse y pass).  Otherwise 4-row groups of nh columns.
__host__ __device__ constexpr int csh_gshift(int nh) { return nh == 4096 / 2 + 1 ? 1 : 2; }
__host__ __device__ constexpr int csh_gld(int nh) { return nh == 4096 / 2 + 1 ? nh + 1 : nh; }

// Plan + scratch of one (nx, ny) grid: tables and the intermediate fields
// G0..G3 (after the inverse y pass) and H (after x), in the row-group layouts
// of csh2d.hip (blkg with gld columns and 2^gsh rows per group; blk8).
struct CshPlan {
    int nx = 0, ny = 0, nh = 0;
    int batch = 1;                   // independent runs of this grid (blockIdx.y), fields packed
    int gld = 0;                     // G0..G3 columns per row group (csh_gld)
    int gsh = 2;                     // G0..G3 log2 rows per group (csh_gshift)
    size_t bstride = 0;              // per run: gld * ny complex elements
    DevBuf g[4], h;
    DevBuf kx, ky, imin;
    const double2* tw_y = nullptr;
    const double2* tw_x = nullptr;   // length nx/2 (forward half-length x transform)
    const double2* tw_xn = nullptr;  // length nx (paired backward x transforms)
    const double2* tw_x16 = nullptr;   // the same two tables for radix-16 schedules
    const double2* tw_xn16 = nullptr;
    const double2* tw_x4 = nullptr;    // and for radix-4 schedules
    const double2* tw_xn4 = nullptr;
    int xcfg = 0;                    // x-pass shape (see csh_xshape_ok / the autotuner)
    const double2* rctw = nullptr;
    DevBuf bad;                      // int per run: set when a step's new state holds a non-finite value
};
// The per-run non-finite flags of a plan: csh_nonfinite synchronises st and
// returns the number of flagged runs (flags[r] = 0 / 1 when flags is not null);
// csh_clear_nonfinite clears run r's flag (every run's for r < 0), ordered on st.
int csh_nonfinite(const CshPlan& p, int* flags, hipStream_t st);
int csh_clear_nonfinite(const CshPlan& p, int r, hipStream_t st);

// Physics of one right-hand side.  screened = 0: derive2csh (psi = omega/k^2,
// derivatives and viscous term on omega).  screened = 1: derive2css
// (psi = omega/(k^2 + lambda), omega_d = k^2 psi carries the derivatives and the
// viscous term, the dealiased nonlinear term is scaled by alpha).
struct CshPhys {
    double nu = 0.0;
    double alpha = 1.0;
    double lambda = 0.0;
    int screened = 0;
    const double* nus = nullptr;  // batched runs: device array of nu per run (overrides nu)
};

// batch: number of independent runs of this grid the launches cover; every
// per-run field (state, G0..G3, H) is packed run after run, bstride apart.
int csh_plan_init(CshPlan& p, int nx, int ny, int batch = 1);
// Host tables of a plan: derive2csh's wavenumbers kx (nh), ky (ny) and the
// dealias bound imin (ny; modes i >= imin[j] drop the nonlinear term).
int csh_host_tables(int nx, int ny, std::vector<double>& kx, std::vector<double>& ky, std::vector<int>& imin);
// nsteps AB2 or RK4 steps (scheme: TARA_SCHEME_*) of derive2csh on the resident
// state w, with each forward y pass fused with the inverse y pass that follows
// it (k_csh_yfwd_inv): one stand-alone inverse y pass per call.  Timing classes
// as csh_rhs_stage (class 2 holds the fused passes).
int csh_advance(const CshPlan& p, int scheme, int nsteps, const CshPhys& ph, double dt, double2* w,
                double2* aux, double2* next, hipStream_t st, EventTimer* tm);
// The parts of csh_advance: the stand-alone inverse y pass of w into G0..G3,
// and one step that expects G0..G3 to hold it (x pass, fused forward + inverse
// y pass per stage; plain_last: the last stage's forward y pass alone, else
// G0..G3 are left holding the inverse y pass of the new state).  For Ny < 1024
// the passes are not fused: csh_yinv_pass does nothing and the step runs
// csh_rhs_stage per stage.
int csh_yinv_pass(const CshPlan& p, const double2* w, const CshPhys& ph, hipStream_t st);
int csh_fused_step(const CshPlan& p, int scheme, const CshPhys& ph, double dt, double2* w, double2* aux,
                   double2* next, bool plain_last, hipStream_t st, EventTimer* tm);
// One right-hand-side evaluation on stage input ws followed by the fused
// update selected by mode (see k_csh_yfwd).  Timing classes: 0 inverse y,
// 1 x pass, 2 forward y + update.
int csh_rhs_stage(const CshPlan& p, const double2* ws, int mode, const CshPhys& ph, double dt, double2* base,
                  double2* aux, double2* next, hipStream_t st, EventTimer* tm);
// x-pass shapes: the default for nx, how many exist, whether shape c runs on
// this plan's grid, and the x pass alone (for timing by the autotuner).
int csh_default_xshape(int nx);
int csh_num_xshapes();
int csh_xshape_ok(const CshPlan& p, int c);
int csh_xpass(const CshPlan& p, hipStream_t st);
// Random-phase vorticity (ky-major state of an nx x ny run): amplitude
// sqrt(k E(k)/pi), E(k) = k^4 exp(-(k/k0)^2), scaled to kinetic energy `energy`.
int csh_random_phase(double2* w, int nx, int ny, unsigned seed, double k0, double energy, hipStream_t st);
// Access to a tara_csh context's device state (capi_csh.hip): the ky-major
// omega_hat, the plan and the stream; and zeroing of its AB2 history.
int csh_ctx_state(tara_csh* c, double2** w, const CshPlan** p, hipStream_t* st);
int csh_ctx_zero_history(tara_csh* c);
int launch_ab2(const double2* w, double2* wn, const double2* fo, const double2* fn, double dt, size_t n,
               hipStream_t st);

}  // namespace tara
