/*
 * libtara_hip.so -- C ABI of the MI355X (gfx950) pseudo-spectral hot path.
 *
 * Plain C: ints, doubles and pointers only.  Spectral arrays use the reference's
 * Fortran layout: complex(8) (re, im) interleaved, column-major half spectra
 * (Nh, Ny[, Nz]) with Nh = Nx/2 + 1, i.e. element (i, j[, k]) (1-based) at
 * double offset 2*((i-1) + (j-1)*Nh [+ (k-1)*Nh*Ny]).  Real fields are
 * (Nx, Ny[, Nz]) column-major.  Grid sizes: each entry point states its own.
 * The transforms and the 2D incompressible drop-ins (derive2csh, derive2css,
 * rk42css, the ab2 steppers) take any size, as FFTW and the reference do;
 * powers of two (32..4096 for the 2D drop-ins) run on the fused kernels, other
 * sizes on mixed-radix transforms and separate passes (slower, same results).
 *
 * Every function returns 0 (TARA_OK) or a negative status; tara_last_error()
 * gives the message (per thread).  Functions taking host pointers are
 * synchronous: they copy in, compute on the GPU and copy out before returning.
 * There is no CPU fallback: without a usable gfx950 device every compute entry
 * point fails with TARA_ERR_HIP.
 */
#ifndef TARA_HIP_H
#define TARA_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define TARA_OK 0
#define TARA_ERR_ARG (-1)
#define TARA_ERR_HIP (-2)
#define TARA_ERR_ALLOC (-3)
#define TARA_ERR_UNSUPPORTED (-4)
#define TARA_ERR_STATE (-5)
#define TARA_ERR_NONFINITE (-6) /* a resident state became NaN or Inf (see the *_sync calls) */

#define TARA_SCHEME_AB2 1
#define TARA_SCHEME_RK4 2

const char* tara_last_error(void);
int tara_version(void);

/* ------------------------------------------------------------------ transforms
 * FFTW 3.3.9 semantics of the calls the reference makes
 * (lib/fftw/api/f77funcs.h:200-218, 294-312: dfftw_plan_dft_{r2c,c2r}_{1d,2d,3d}_
 * followed by dfftw_execute_, FFTW_ESTIMATE): unnormalised, forward e^{-i},
 * halved axis = Fortran first index.  c2r runs the complex passes first and the
 * x pass last, reading only the real part of the kx = 0 and kx = Nx/2 entries
 * (lib/fftw/rdft/rank-geq2-rdft2.c:40-68).  Unlike FFTW, c2r leaves `in` intact.
 */
int tara_fft_r2c_1d(int nx, const double* in, double* out);   /* any sizes >= 1, at most 2^31 points */
int tara_fft_c2r_1d(int nx, const double* in, double* out);
int tara_fft_r2c_2d(int nx, int ny, const double* in, double* out);
int tara_fft_c2r_2d(int nx, int ny, const double* in, double* out);
int tara_fft_r2c_3d(int nx, int ny, int nz, const double* in, double* out);
int tara_fft_c2r_3d(int nx, int ny, int nz, const double* in, double* out);

/* --------------------------------------------- 2D incompressible hydro (csh)
 * Replaces src/derive2csh.f95:3 derive2csh(Nx,Ny,Nh,nu,time,omegak,omegak_new,
 * dt_omegak_old,dt_omegak_new).  Only omegak (in) and dt_omegak_new (out) are
 * live in the reference; time, omegak_new and dt_omegak_old are unused there
 * and are not passed.  nh must equal nx/2+1.
 */
int tara_derive2csh(int nx, int ny, int nh, double nu, const double* omegak, double* dt_omegak_new);

/* No counterpart in the reference: tara_derive2csh for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  omegak and dt_omegak_new are DEVICE
 * arrays in the reference layout; the work is enqueued on `stream` (a
 * hipStream_t; NULL = the null stream) and the call returns without waiting.
 * Every intermediate lives in the library's static 64 MiB per-device scratch,
 * so grids are limited to nx, ny <= 1024 (larger ones: TARA_ERR_UNSUPPORTED,
 * before anything is enqueued); no hipMalloc / hipFree is made.  The first
 * call for a grid on a device writes that grid's tables into the scratch
 * (waiting for `stream` first); later calls on the same grid only launch
 * kernels.  Calls that overlap on one GPU race: they share its scratch.
 * Results equal tara_derive2csh's bit for bit. */
int tara_derive2csh_noalloc(int nx, int ny, int nh, double nu, const double* omegak, double* dt_omegak_new,
                            void* stream);

/* Replaces src/ab2csh.f95:3 ab2csh(Nx,Ny,Nh,nu,time,dt,omegak,omegak_new,
 * dt_omegak_old,dt_omegak_new): omegak_new = omegak + (3/2 f_n - 1/2 f_{n-1}) dt. */
int tara_ab2csh(int nx, int ny, int nh, double dt, const double* omegak, double* omegak_new,
                const double* dt_omegak_old, const double* dt_omegak_new);

/* Device-resident stepping of the same equations (performance mode).  Replaces
 * the time loop of src/main2csh.f95:114-126 (derive2csh + ab2csh + history
 * rotation) when the caller does not need the field on the host every step.
 * scheme: TARA_SCHEME_AB2 (the reference's scheme; history starts at zero) or
 * TARA_SCHEME_RK4 (classic RK4, the intended scheme of src/rk42css.f95:15-48).
 */
typedef struct tara_csh tara_csh;
int tara_csh_create(int nx, int ny, double nu, double dt, int scheme, tara_csh** out);
int tara_csh_upload(tara_csh* ctx, const double* omegak);   /* zeroes the AB2 history */
int tara_csh_download(tara_csh* ctx, double* omegak);
int tara_csh_advance(tara_csh* ctx, int nsteps);             /* asynchronous */
/* Waits for the context's work.  Guard: the last stage of every step flags a
 * new state value that is NaN or Inf; sync (and download, which still copies
 * the state out) then return TARA_ERR_NONFINITE until a new state is set
 * (upload or an initial condition).  The reference has no such check. */
int tara_csh_sync(tara_csh* ctx);
int tara_csh_destroy(tara_csh* ctx);
/* Kernel timing: when on, each launch is bracketed by HIP events on the
 * context's stream.  tara_csh_kernel_stats returns, per kernel class
 * (0 = inverse y pass, 1 = x pass, 2 = forward y pass + update), the number of
 * timed launches and their summed duration in milliseconds, then resets.  For
 * the derive2csh solver at Ny >= 1024, class 2 launches (all but each advance
 * call's last) also run the next stage's inverse y pass, and class 0 holds one
 * launch per advance call. */
int tara_csh_set_timing(tara_csh* ctx, int on);
int tara_csh_kernel_stats(tara_csh* ctx, int nclass, long* counts, double* ms);
/* Autotuner (no counterpart in the reference, whose FFTW planner runs in
 * FFTW_ESTIMATE mode): times every x-pass shape that runs on this grid (reps
 * launches each, on the context's stream) and keeps the fastest.  ms, if not
 * NULL, receives per-shape milliseconds (-1 for shapes that do not run);
 * tara_csh_num_xshapes() gives its length.  tara_csh_set_xshape forces one. */
int tara_csh_num_xshapes(void);
int tara_csh_tune(tara_csh* ctx, int reps, int* chosen, double* ms);
int tara_csh_set_xshape(tara_csh* ctx, int shape);

/* Per-step work of the reference's driver on the resident state
 * (src/main2csh.f95:128-178): real-space omega, ux, uy (each (Nx, Ny), divided
 * by Nx*Ny as the driver does; any pointer may be NULL).  tara_csh_record
 * advances nframes x every steps and stores omegak after every `every` steps
 * (frames: nframes consecutive (Nh, Ny) half spectra). */
int tara_csh_fields(tara_csh* ctx, double* omega, double* ux, double* uy);
int tara_csh_record(tara_csh* ctx, int nframes, int every, double* frames);
/* The driver's tracer particles (src/main2csh.f95:187-265) on the device:
 * nsteps times, one step of the state followed by the particle update with
 * the new state's velocity (ux = i ky psi, uy = -i kx psi with the driver's
 * kx = 2 pi (i-1)/lx, ky likewise, psi = omega/k^2; c2r divided by Nx*Ny),
 * bilinear interpolation as the driver writes it, x <- x + u dt and the wrap
 * into (0, lx] x (0, ly].  xp, yp: host arrays of np positions in
 * [0, lx] x [0, ly], updated in place; lx, ly: the driver's box (the input
 * values times pi); dt: the driver's step.  periodic == 0 reads the corners
 * as the driver addresses its column-major arrays (the last column's right
 * neighbour is the next row's first element; reads past the end, undefined
 * there, read 0, as the driver's do when glibc allocates its arrays with
 * mmap, from 128 KiB); periodic != 0 takes every corner's periodic image.  Synchronous; TARA_ERR_NONFINITE if a position became non-finite. */
int tara_csh_advance_tracers(tara_csh* ctx, int nsteps, double dt, double lx, double ly, int periodic, int np,
                             double* xp, double* yp);
/* Initial conditions generated on the device into the context's state (the AB2
 * history is zeroed).  tara_csh_ic_shear: the shear layer of
 * src/main2csh.f95:84-110 (box and d as input.ini gives them, multiplied by pi
 * inside).  tara_csh_ic_random: random-phase omega with
 * E(k) ~ k^4 exp(-(k/k0)^2), kinetic energy `energy`, Nyquist modes and k = 0
 * zero, Hermitian kx = 0 column; the phase of mode (i, j) is 2 pi u with u built
 * from the first two words of Philox-4x32-10(counter = {i + j*Nh, 0, 0, 0},
 * key = {seed, 0}).  No counterpart in the reference (its driver draws no
 * random fields). */
int tara_csh_ic_shear(tara_csh* ctx, double lx_in, double ly_in, double w0, double m0, double d_in);
int tara_csh_ic_random(tara_csh* ctx, unsigned seed, double k0, double energy);
/* The counter-based generator used above (host evaluation). */
int tara_philox4x32_10(const unsigned ctr[4], const unsigned key[2], unsigned out[4]);

/* ------------------------------------------------ batched ensemble (csh)
 * No counterpart in the reference (one run per process there): nruns
 * independent 2D incompressible runs (derive2csh physics, grid n[r] x n[r],
 * viscosity nu[r], common dt, classic RK4 or AB2).  Runs of one grid size form a
 * bucket whose passes are single batched launches; one step of every bucket is
 * captured once in a hipGraph (tara_ens_use_graph(ctx, 0) launches eagerly
 * instead) and replayed per step.  upload/download take one run's (Nh, N) half
 * spectrum in the reference layout; upload zeroes that run's AB2 history.
 * tara_ens_tune runs the x-pass autotuner per bucket.
 */
typedef struct tara_ens tara_ens;
int tara_ens_create(int nruns, const int* n, const double* nu, double dt, int scheme, tara_ens** out);
int tara_ens_upload(tara_ens* ctx, int run, const double* omegak);
int tara_ens_download(tara_ens* ctx, int run, double* omegak);
int tara_ens_advance(tara_ens* ctx, int nsteps);
/* sync returns TARA_ERR_NONFINITE when any run's state became NaN or Inf (the
 * guard of tara_csh_sync, per run); tara_ens_nonfinite gives the number of such
 * runs and, if flags (nruns ints) is not NULL, 1 for each of them.  A run's flag
 * clears when its state is set again (upload, ic_random, set_omega); the other
 * runs are not affected by one run's failure. */
int tara_ens_sync(tara_ens* ctx);
int tara_ens_nonfinite(tara_ens* ctx, int* flags);
int tara_ens_use_graph(tara_ens* ctx, int on);
int tara_ens_tune(tara_ens* ctx, int reps);
int tara_ens_destroy(tara_ens* ctx);
/* The captured step graph: total nodes and kernel nodes (0 before the first
 * advance call; tara_ens_advance(ctx, 0) captures it without stepping). */
int tara_ens_graph_nodes(tara_ens* ctx, long* nodes, long* kernels);
/* Device-side data paths of one run, ordered on the ensemble's stream:
 * ic_random sets the random-phase vorticity of tara_csh_ic_random (synchronous);
 * set_omega / get_omega move real-space vorticity (device float, n*n, x fastest)
 * into / out of the FP64 state through the library's FFTs (asynchronous;
 * set_omega zeroes the run's AB2 history; get_omega's pointer must stay valid
 * until tara_ens_sync). */
int tara_ens_ic_random(tara_ens* ctx, int run, unsigned seed, double k0, double energy);
int tara_ens_set_omega(tara_ens* ctx, int run, const float* omega);
int tara_ens_get_omega(tara_ens* ctx, int run, float* omega);

/* ------------------------------------- 2D screened vorticity (css family)
 * Replaces src/derive2css.f95:6 derive2css(Nx,Ny,Nh,alpha,nu,time,omegak,
 * omegak_new,dt_omegak_old,dt_omegak_new): psi = omega/(k^2 + lambda) with
 * psi(0,0) = 0, omega_d = k^2 psi, f = -(alpha mask(u . grad omega_d) +
 * nu k^2 omega_d).  The reference leaves Lx, Ly and lambda uninitialised
 * (src/derive2css.f95:11); this takes Lx = Ly = 2 pi as derive2csh does
 * (integer wavenumbers) and lambda as an argument (src/main2css.f95:87 sets 10).
 * lambda = 0, alpha = 1 gives derive2csh's right-hand side.
 */
int tara_derive2css(int nx, int ny, int nh, double alpha, double nu, double lambda, const double* omegak,
                    double* dt_omegak_new);
/* Replaces src/ab2css.f95:6 (the same AB2 update as ab2csh). */
int tara_ab2css(int nx, int ny, int nh, double dt, const double* omegak, double* omegak_new,
                const double* dt_omegak_old, const double* dt_omegak_new);
/* Replaces src/rk42css.f95:6 rk42css with the classic RK4 step it intends
 * (stages at dt/2, dt/2, dt; the reference reads an uninitialised slope,
 * src/rk42css.f95:15-16, and never uses its derive2css results). */
int tara_rk42css(int nx, int ny, int nh, double alpha, double nu, double lambda, double dt, const double* omegak,
                 double* omegak_new);
/* Device-resident css stepping; the context is driven by tara_csh_upload /
 * advance / download / sync / destroy. */
int tara_css_create(int nx, int ny, double alpha, double nu, double lambda, double dt, int scheme,
                    tara_csh** out);

/* ------------------------- 2D compressible families (coh hydro, com MHD)
 * Box and wavenumbers as the reference forms them: kx = 2 pi (i-1)/Lx with the
 * caller's pi (an argument of derive2coh / derive2com) and Lx, Ly already
 * multiplied by pi (src/main2coh.f95:101).  Circular dealias
 * sqrt(kx^2+ky^2) >= (Nx+Ny)/2/3 on every term.  Grids: powers of two,
 * 16..256 per axis.
 */
typedef struct {
    double pi;       /* the pi argument */
    double lx, ly;   /* COMMON /comm/ Lx, Ly (src/main2coh.f95:29) */
    double mux, muy; /* viscosities of the x and y second derivatives */
    double cs;       /* isothermal sound speed */
    double dt;       /* resident stepping only */
} tara_coh_params;

/* Replaces src/derive2coh.f95:5 derive2coh(Nx,Ny,Nh,pi,time,nk,ukx,uky,
 * dn_dt_old,dux_dt_old,duy_dt_old,dn_dt_new,dux_dt_new,duy_dt_new):
 * state = {nk, ukx, uky} (in), rhs = {dn_dt_new, dux_dt_new, duy_dt_new} (out).
 * The _old arrays are unused by the reference and not passed. */
int tara_derive2coh(int nx, int ny, int nh, const tara_coh_params* p, const double* const state[3],
                    double* const rhs[3]);
/* Replaces src/ab2coh.f95:5: new = old + (3/2 f_n - 1/2 f_{n-1}) dt per field. */
int tara_ab2coh(int nx, int ny, int nh, double dt, const double* const state[3], double* const state_new[3],
                const double* const f_old[3], const double* const f_new[3]);
/* Replaces src/euler2coh.f95:5: new = old + slope dt per field. */
int tara_euler2coh(int nx, int ny, int nh, double dt, const double* const state[3], double* const state_new[3],
                   const double* const slope[3]);
/* Replaces rk4_nk2coh / rk4_ukx2coh / rk4_uky2coh (src/rk4nk2coh.f95:6 and
 * its twins) on the field each one advances.  Their derive2coh calls write
 * local arrays that are never read, so all four stage slopes are the input
 * slope: new = old + dt/6 (((s + 2s) + 2s) + s). */
int tara_rk4coh_field(int nx, int ny, int nh, double dt, const double* field, double* field_new,
                      const double* slope);
/* Replaces pc_nk2coh / pc_ukx2coh / pc_uky2coh (src/pcnk2coh.f95:6): predictor
 * and corrector slopes are both the input slope: new = old + ((s + s) dt)/2. */
int tara_pccoh_field(int nx, int ny, int nh, double dt, const double* field, double* field_new,
                     const double* slope);

typedef struct {
    double pi;
    double lx, ly;   /* COMMON /comm/ Lx, Ly (src/main2com.f95:43) */
    double mu, eta;  /* viscosity, resistivity */
    double cs;       /* P = CS^2 rho */
    double kf, a;    /* forcing F = rho A (-sin(kf y), sin(kf x)) */
    double dt;       /* resident stepping only */
} tara_com_params;

/* Replaces src/derive2com.f95:5 derive2com(Nx,Ny,Nh,pi,time,rho_k,rho_ux_k,
 * rho_uy_k,E_k,Bx_k,By_k, 12 interleaved (d_*_dt_old, d_*_dt_new)):
 * state = {rho, rho ux, rho uy, E, Bx, By} (in), rhs = the six d_*_dt_new (out).
 * The forcing uses x(i) = (i-1) Lx/Nx, y(j) = (j-1) Ly/Ny, the grid of
 * src/main2com.f95:231-233 (the routine's own x, y are uninitialised,
 * src/derive2com.f95:15).  curl B is formed as the reference writes it
 * (from the velocity derivatives divided by Nx Ny once more,
 * src/derive2com.f95:186-187). */
int tara_derive2com(int nx, int ny, int nh, const tara_com_params* p, const double* const state[6],
                    double* const rhs[6]);
/* Replaces src/ab2com.f95:6 (AB2 on all six fields). */
int tara_ab2com(int nx, int ny, int nh, double dt, const double* const state[6], double* const state_new[6],
                const double* const f_old[6], const double* const f_new[6]);

/* Device-resident AB2 stepping of either family (the derive + ab2 + rotation
 * loops of src/main2coh.f95:278-300 and src/main2com.f95:364-380, without the
 * diagnostics).  upload/download take 3 (coh) or 6 (com) arrays; the AB2
 * history starts at zero on upload.  Timing classes: 0 = prep + inverse y,
 * 1 = x passes (+ the com middle pass), 2 = forward y + assembly.  sync and
 * download return TARA_ERR_NONFINITE once an update has stored a NaN or Inf
 * (the flag clears on upload). */
typedef struct tara_f2 tara_f2;
int tara_coh_create(int nx, int ny, const tara_coh_params* p, tara_f2** out);
int tara_com_create(int nx, int ny, const tara_com_params* p, tara_f2** out);
int tara_f2_upload(tara_f2* ctx, const double* const state[]);
int tara_f2_download(tara_f2* ctx, double* const state[]);
int tara_f2_advance(tara_f2* ctx, int nsteps);
int tara_f2_sync(tara_f2* ctx);
int tara_f2_destroy(tara_f2* ctx);
int tara_f2_set_timing(tara_f2* ctx, int on);
int tara_f2_kernel_stats(tara_f2* ctx, int nclass, long* counts, double* ms);

/* ------------------------------------ 3D compressible isothermal MHD (3com)
 * Fields, in the order of the reference's argument lists: rho, rho_ux, rho_uy,
 * rho_uz, E, Bx, By, Bz; each a (Nh, Ny, Nz) half spectrum.  Physical box
 * lengths are the values after the reference multiplies the input by pi
 * (src/main3com.f95:203); mu and eta are the viscosity and resistivity of the
 * diffusion terms (the reference hard-wires both to 1/450,
 * src/derive3com.f95:1023-1037); cs is the isothermal sound speed (P = cs^2 rho,
 * src/derive3com.f95:395).  Grids: powers of two, 16..256 per axis.
 */
typedef struct {
    double lx, ly, lz;
    double mu, eta;
    double cs;
    double dt;
} tara_mhd_params;

/* Replaces src/derive3com.f95:5 derive3com(Nx,Ny,Nz,Nh,pi,time, 8 state arrays,
 * 8 interleaved (d_*_dt_old, d_*_dt_new) pairs).  rhs[f] receives d_*_dt_new.
 * The energy right-hand side rhs[4] is returned as zero: ab3com never reads it
 * (the energy is frozen, src/ab3com.f95:51). */
int tara_derive3com(int nx, int ny, int nz, int nh, const tara_mhd_params* p, const double* const state[8],
                    double* const rhs[8]);

/* Replaces src/ab3com.f95:6 ab3com: new = old + (3/2 f_n - 1/2 f_{n-1}) dt for
 * rho, rho u, B; E is copied unchanged. */
int tara_ab3com(int nx, int ny, int nz, int nh, double dt, const double* const state[8],
                double* const state_new[8], const double* const f_old[8], const double* const f_new[8]);

/* Device-resident AB2 stepping of the same equations (the derive3com + ab3com
 * + rotation of src/main3com.f95:536-617, without the diagnostics).  The AB2
 * history starts at zero on upload. */
typedef struct tara_mhd tara_mhd;
int tara_mhd_create(int nx, int ny, int nz, const tara_mhd_params* p, tara_mhd** out);
int tara_mhd_upload(tara_mhd* ctx, const double* const state[8]);
int tara_mhd_download(tara_mhd* ctx, double* const state[8]);
int tara_mhd_advance(tara_mhd* ctx, int nsteps);
/* Waits for the context's work; TARA_ERR_NONFINITE (also from download, which
 * still copies the state out) when an AB2 step produced a NaN or Inf state
 * value, until a new state is set (upload or an initial condition). */
int tara_mhd_sync(tara_mhd* ctx);
int tara_mhd_destroy(tara_mhd* ctx);
/* The reference's initial state, src/main3com.f95:290-324 (ABC flow): uniform
 * rho0 and pressure p0, u = u0 (a sin(kf z) + cc cos(kf y), b sin(kf x) + a cos(kf z),
 * cc sin(kf y) + b cos(kf x)), uniform B = b0 on each axis, E = p0 + rho0 u^2/2 +
 * B^2/2, on the grid x(i) = (i-1) Lx/Nx; transformed on the device with the
 * library's FFTs into the context's state (history zeroed).  The main's values:
 * rho0 = 1, u0 = 0.1, b0 = 0.1 sqrt(mu_0), p0 = 1, a = b = cc = kf = 1 (:228-268). */
int tara_mhd_ic_abc(tara_mhd* ctx, double rho0, double u0, double b0, double p0, double a, double b, double cc,
                    double kf);
/* Random-phase initial state generated on the device (history zeroed): rho = 1,
 * solenoidal u and B with shell spectrum ~ k^4 exp(-2 (k/k0)^2), scaled to
 * 1/2 <u^2> = e_kin and 1/2 <B^2> = e_mag, rho u = u, E = 1 + u^2/2 + B^2/2.
 * Each mode's three phases come from Philox-4x32-10 with key {seed, 1} (u) or
 * {seed, 2} (B), counters {m, 0, 0, 0} and {m, 1, 0, 0}, m = i + Nh (j + Ny l);
 * the kx = 0 plane is Hermitian, k = 0 and the Nyquist planes are zero, and the
 * raw vector is projected with its own wavevector.  No counterpart in the
 * reference (its 3D initial state is the ABC flow of tara_mhd_ic_abc). */
int tara_mhd_ic_random(tara_mhd* ctx, unsigned seed, double k0, double e_kin, double e_mag);
/* The reference driver's per-step diagnostics (src/main3com.f95:619-1071,
 * the columns after `time` of its Energy.dat row, :1070) on the resident
 * state, computed on the device: out[0..9] = Pressure <CS^2 rho>, Energy
 * <|u|^2>/2, T_Energy <E>, B_Field <|B|^2>/(2 mu_0) (mu_0 = 1), C1 <|omega|>,
 * C2 <|omega|^2>, Hf <u.omega>, Hm <A.B>, HT <(u+A).(omega+B)>, div_B_Tot
 * <div B>, with u = rho u/rho, A = curl B/k^2 and omega formed as that loop
 * forms them (its omega_y is i kx uz - i kz ux).  tara_mhd_record_energy
 * advances nrec x every steps and writes the ten values after every `every`
 * steps into rows[nrec][10] (one device-to-host copy at the end).  Both are
 * synchronous; the first call allocates the work space (15 real and 11
 * spectral fields). */
int tara_mhd_energy(tara_mhd* ctx, double out[10]);
int tara_mhd_record_energy(tara_mhd* ctx, int nrec, int every, double* rows);
/* Kernel timing as for tara_csh: classes 0 = inverse z+y passes, 1 = x pass,
 * 2 = forward y pass, 3 = forward z pass + assembly + update. */
int tara_mhd_set_timing(tara_mhd* ctx, int on);
int tara_mhd_kernel_stats(tara_mhd* ctx, int nclass, long* counts, double* ms);

/* ------------------------------------------------- matrix-core GEMM (FNO)
 * No counterpart in the reference.  C[z](m, n) = (or +=) alpha * sum_k
 * A[z](m, k) B[z](k, n) for z = (z1, z2), z1 < nb1, z2 < nb2, on device
 * pointers; element (m, k) of A at a + m*sam + k*sak + z1*sa1 + z2*sa2 (elements),
 * likewise B and C.  dtype 0: A, B bfloat16 (MFMA 32x32x8, FP32 accumulation);
 * 1: float (MFMA 32x32x2).  epi 0: store float, 1: add into float, 2: store
 * bfloat16 (round to nearest even).  Synchronous. */
int tara_gemm(int dtype, int epi, long m, long n, long k, const void* a, long sam, long sak, const void* b, long sbk,
              long sbn, void* c, long scm, long scn, int nb1, long sa1, long sb1, long sc1, int nb2, long sa2,
              long sb2, long sc2, float alpha);

/* ------------------------------------------- Fourier-neural-operator surrogate
 * No counterpart in the reference.  Maps real-space vorticity omega(t) (device
 * float [batch][n][n], x fastest) to omega(t + T): lift, `layers` Fourier
 * layers of width `width` keeping kx < mx and -my/2 <= ky < my/2, GELU,
 * projection through 2*width hidden channels, per-sample amplitude scaling and
 * a residual connection (parameter layout in tara_amd/csrc/fno.hip).
 * precision 0: bfloat16 operands with FP32 accumulation on the matrix cores;
 * 1: FP32 operands throughout.  Training minimises the batch mean of the
 * relative L2 error |out - target| / |target| with Adam (0.9, 0.999, 1e-8).
 * forward, grads, train_step and rel_err are synchronous; get/set_params and
 * grads' grad_host use host float arrays of tara_fno_num_params entries.
 * rel_err: err[b] = |a_b - ref_b| / |ref_b| over device arrays. */
typedef struct tara_fno tara_fno;
int tara_fno_create(int n, int width, int mx, int my, int layers, int max_batch, int precision, unsigned seed,
                    tara_fno** out);
int tara_fno_num_params(tara_fno* f, long* n);
int tara_fno_get_params(tara_fno* f, float* host);
int tara_fno_set_params(tara_fno* f, const float* host);
int tara_fno_forward(tara_fno* f, int batch, const float* omega, float* out);
int tara_fno_grads(tara_fno* f, int batch, const float* omega, const float* target, float* loss, float* grad_host);
int tara_fno_train_step(tara_fno* f, int batch, const float* omega, const float* target, float lr, float* loss);
int tara_fno_rel_err(tara_fno* f, int batch, const float* a, const float* ref, float* err_host);
/* Training batches: xb[b] = inp[idx[b]], yb[b] = tgt[idx[b]] (device arrays of
 * n x n floats per sample; idx a host array of batch indices in [0, npairs),
 * batch <= 64), enqueued on the network's stream before its next step. */
int tara_fno_gather_pairs(tara_fno* f, int batch, const long* idx, long npairs, const float* inp, const float* tgt,
                          float* xb, float* yb);
/* Surrogate rollout with the FP64 solver as checker.  traj: device float
 * [nsteps+1][batch][n][n], traj[0] set by the caller; step t+1 = network(step t).
 * Every `every`-th step t+1 = (j+1)*every (j < nchk = nsteps / every) is
 * re-checked: the solver (the checker ensembles' viscosity, dt and scheme, grid
 * n) advances traj[t] by solver_steps steps and the check's error is the
 * relative L2 distance of step t+1 from that.  nens = nchk: ensemble j (batch
 * runs) runs check j on its own streams as soon as traj[t] exists, concurrently
 * with the rollout; nens = 1: one ensemble of nchk*batch runs (run j*batch + b)
 * advances every check together once the rollout's steps are done.
 * anchor (batch runs, same solver; needed when nchk > 0) holds the last state
 * known to be the FP64 solution, traj[0] to begin with.  When a check's error
 * exceeds tol for any sample, the FP64 solver takes the trajectory over: the
 * anchor advances from its step to the failing step T = (j+1)*every, its fields
 * replace traj[T] for every sample, T becomes the anchor's step and the rollout
 * resumes from there.  So every state after a take-over is the FP64 solution
 * from traj[0] (with tol = 0 every check fails and each traj[(j+1)*every] is
 * the plain solver's).  check_err (host, [nchk][batch]) receives each check's
 * error, fallbacks the number of take-overs.  use_graph replays a captured
 * network step.  Synchronous. */
int tara_fno_rollout(tara_fno* f, int batch, float* traj, int nsteps, int every, int nens, tara_ens* const* checkers,
                     tara_ens* anchor, int solver_steps, float tol, int use_graph, float* check_err, int* fallbacks);
int tara_fno_destroy(tara_fno* f);

/* ------------------------------------------------------ snapshots and data
 * Host I/O, no GPU needed.
 * Checkpoints (FP64): a 64-byte header ("TARACKPT", version 1, kind 1, nx, ny,
 * nz, nfields as uint32, step as int64, time as double, 16 reserved bytes), then
 * nfields (Nh, Ny, Nz) complex(8) half spectra in the reference layout (nz = 1
 * for 2D).
 * tara_fort_write: the reference's text snapshot (src/main2csh.f95:170-178):
 * per grid point, i outer and j inner, one line "x y omega ux uy" with
 * x = (i-1) Lx/Nx, y = (j-1) Ly/Ny; values printed %24.16E (list-directed
 * output's spacing is compiler-specific; the columns are the reference's).
 * Trajectory pairs for surrogate training: a 64-byte header ("TARAPAIR",
 * version 1, dtype 1 = float32 / 2 = bfloat16, nx, ny as uint32, npairs as
 * uint64, dt between the fields of a pair, nu, 16 reserved bytes), then per pair
 * the real-space (Nx, Ny) omega before and after, Fortran order.  BF16 values are
 * rounded to nearest even from the FP32 value.  No counterpart in the reference.
 */
int tara_ckpt_write(const char* path, int nx, int ny, int nz, int nfields, long step, double time,
                    const double* const fields[]);
int tara_ckpt_read_header(const char* path, int* nx, int* ny, int* nz, int* nfields, long* step, double* time);
int tara_ckpt_read(const char* path, double* const fields[]);
int tara_fort_write(const char* path, int nx, int ny, double lx, double ly, const double* omega, const double* ux,
                    const double* uy);
typedef struct tara_pairs tara_pairs;
int tara_pairs_create(const char* path, int nx, int ny, int dtype, double dt_pair, double nu, tara_pairs** out);
int tara_pairs_append(tara_pairs* p, const double* omega_in, const double* omega_out);
int tara_pairs_close(tara_pairs* p);

#ifdef __cplusplus
}
#endif
#endif /* TARA_HIP_H */
