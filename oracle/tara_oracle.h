/*
 * TEST INFRASTRUCTURE ONLY.  This is the CPU oracle: a plain-C restatement of
 * the reference's arithmetic (TARA's Fortran routines and the FFTW transforms
 * they call), used by tests/, by __graft_entry__.smoke() as the checker and by
 * bench.py's cpu_baseline leg.  The product (tara_amd/, libtara_hip.so) never
 * links, loads or calls anything here.
 *
 * Layout conventions are the reference's: Fortran column-major, complex(8)
 * interleaved (re, im), real fields (Nx,Ny[,Nz]) and half spectra
 * (Nh = Nx/2+1, Ny[, Nz]).  All arithmetic is IEEE double, no FMA contraction.
 */
#ifndef TARA_ORACLE_H
#define TARA_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- transforms (FFTW 3.3.9 semantics: unnormalised, e^{-i} forward) ---- */
void orc_dft(int n, double *z, long stride, int sign);
void orc_r2c_2d(int nx, int ny, const double *in, double *out);
void orc_c2r_2d(int nx, int ny, const double *in, double *out);
void orc_r2c_3d(int nx, int ny, int nz, const double *in, double *out);
void orc_c2r_3d(int nx, int ny, int nz, const double *in, double *out);

/* ---- 2D incompressible hydro (derive2csh / ab2csh family) ---- */
void orc_derive2csh(int nx, int ny, double nu, const double *omegak, double *dt_new);
void orc_ab2csh(int nx, int ny, double dt, const double *omegak, double *omegak_new,
                const double *dt_old, const double *dt_new);
void orc_ab2csh_steps(int nx, int ny, double nu, double dt, int nsteps,
                      double *omegak, double *hist);
void orc_rk4csh_steps(int nx, int ny, double nu, double dt, int nsteps, double *omegak);
/* ---- 2D screened vorticity (derive2css family; Lx = Ly = 2 pi, lambda given) ---- */
void orc_derive2css(int nx, int ny, double alpha, double nu, double lambda, const double *omegak, double *dt_new);
void orc_rk4css_steps(int nx, int ny, double alpha, double nu, double lambda, double dt, int nsteps,
                      double *omegak);
void orc_philox4x32_10(const unsigned ctr[4], const unsigned key[2], unsigned out[4]);
void orc_ic_random(int nx, int ny, unsigned seed, double k0, double energy, double *omegak);
void orc_ic2csh(int nx, int ny, double lx_in, double ly_in, double w0, double m0,
                double d_in, double *omega, double *omegak);

/* ---- 2D compressible families (derive2coh, derive2com) ---- */
typedef struct {
    double pi, lx, ly;   /* pi argument; box already multiplied by pi */
    double mux, muy, cs;
} orc_coh_params;
typedef struct {
    double pi, lx, ly;
    double mu, eta, cs, kf, a;
} orc_com_params;
void orc_derive2coh(int nx, int ny, const orc_coh_params *p, const double *const state[3],
                    double *const rhs[3]);
void orc_derive2com(int nx, int ny, const orc_com_params *p, const double *const state[6],
                    double *const rhs[6]);

/* ---- 3D compressible isothermal MHD (derive3com / ab3com) ---- */
typedef struct {
    double lx, ly, lz;   /* physical box (already multiplied by pi) */
    double mu, eta;      /* viscosity / resistivity used in the diffusion terms */
    double cs;           /* isothermal sound speed, P = CS^2 rho */
    double dt;
} orc_mhd_params;
/* state: 8 spectra (rho, rho_ux, rho_uy, rho_uz, E, Bx, By, Bz) each (Nh,Ny,Nz).
   rhs: 8 spectra; E's slot is written with zero (it is dead in the reference). */
void orc_derive3com(int nx, int ny, int nz, const orc_mhd_params *p,
                    const double *const state[8], double *const rhs[8]);
void orc_ab3com(int nx, int ny, int nz, double dt, const double *const state[8],
                double *const state_new[8], const double *const f_old[8],
                const double *const f_new[8]);
/* Random-phase initial state (no reference counterpart): rho = 1, solenoidal
   u and B with shell spectrum ~ k^4 exp(-2 (k/k0)^2), 1/2 <u^2> = e_kin,
   1/2 <B^2> = e_mag (Philox-4x32-10 phases, keys {seed, 1} and {seed, 2}),
   rho u = u, E = 1 + u^2/2 + B^2/2. */
void orc_ic_random3d(int nx, int ny, int nz, unsigned seed, double k0, double e_kin, double e_mag,
                     double *const st[8]);

#ifdef __cplusplus
}
#endif
#endif
