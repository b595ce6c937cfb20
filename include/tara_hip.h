/*
 * libtara_hip.so -- C ABI of the MI355X (gfx950) pseudo-spectral hot path.
 *
 * Plain C: ints, doubles and pointers only.  Spectral arrays use the reference's
 * Fortran layout: complex(8) (re, im) interleaved, column-major half spectra
 * (Nh, Ny[, Nz]) with Nh = Nx/2 + 1, i.e. element (i, j[, k]) (1-based) at
 * double offset 2*((i-1) + (j-1)*Nh [+ (k-1)*Nh*Ny]).  Real fields are
 * (Nx, Ny[, Nz]) column-major.  Grid sizes must be powers of two, 32..4096.
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

#define TARA_SCHEME_AB2 1
#define TARA_SCHEME_RK4 2

const char* tara_last_error(void);
int tara_version(void);

/* ------------------------------------------------------------------ transforms
 * FFTW 3.3.9 semantics of the calls the reference makes
 * (lib/fftw/api/f77funcs.h:206-218, 300-312: dfftw_plan_dft_{r2c,c2r}_{2d,3d}_
 * followed by dfftw_execute_, FFTW_ESTIMATE): unnormalised, forward e^{-i},
 * halved axis = Fortran first index.  c2r runs the complex passes first and the
 * x pass last, reading only the real part of the kx = 0 and kx = Nx/2 entries
 * (lib/fftw/rdft/rank-geq2-rdft2.c:40-68).  Unlike FFTW, c2r leaves `in` intact.
 */
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
int tara_csh_sync(tara_csh* ctx);
int tara_csh_destroy(tara_csh* ctx);
/* Kernel timing: when on, each launch is bracketed by HIP events on the
 * context's stream.  tara_csh_kernel_stats returns, per kernel class
 * (0 = inverse y pass, 1 = x pass, 2 = forward y pass + update), the number of
 * timed launches and their summed duration in milliseconds, then resets. */
int tara_csh_set_timing(tara_csh* ctx, int on);
int tara_csh_kernel_stats(tara_csh* ctx, int nclass, long* counts, double* ms);

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
int tara_mhd_sync(tara_mhd* ctx);
int tara_mhd_destroy(tara_mhd* ctx);
/* Kernel timing as for tara_csh: classes 0 = inverse z+y passes, 1 = x pass,
 * 2 = forward y pass, 3 = forward z pass + assembly + update. */
int tara_mhd_set_timing(tara_mhd* ctx, int on);
int tara_mhd_kernel_stats(tara_mhd* ctx, int nclass, long* counts, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* TARA_HIP_H */
