/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle.  Plain-C restatement of the
 * reference arithmetic on the hot path, used only as the checker by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The product never
 * links or calls it.
 *
 * Each routine cites the reference file:line it restates.  Arithmetic is IEEE
 * double with FMA contraction disabled at compile time (Makefile), evaluated
 * in the reference's expression order where that order is visible in the
 * Fortran source.  The transforms restate FFTW 3.3.9's definitions (unnormalised,
 * e^{-2 pi i jk/n} forward) and its multi-dimensional ordering
 * (lib/fftw/rdft/rank-geq2-rdft2.c:40-68): r2c = real pass along the halved
 * (Fortran-first) axis then complex passes along the others; c2r = complex
 * passes first, then the 1-D c2r along x, which reads only the real part of
 * the kx = 0 and kx = Nx/2 entries.
 */
#include "tara_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define RE(p, i) ((p)[2 * (size_t)(i)])
#define IM(p, i) ((p)[2 * (size_t)(i) + 1])

/* ------------------------------------------------------------------------ */
/* 1-D complex DFT: radix-2 iterative for powers of two, direct O(n^2) with   */
/* long-double accumulation otherwise.  Twiddles from long-double cosl/sinl. */
/* ------------------------------------------------------------------------ */

static void twiddle(int n, long k, int sign, double *c, double *s)
{
    long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)(k % n) / (long double)n;
    *c = (double)cosl(a);
    *s = (double)(sign * sinl(a));
}

void orc_dft(int n, double *z, long stride, int sign)
{
    if (n <= 1) return;
    double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int j = 0; j < n; ++j) {
        buf[2 * j] = z[2 * (size_t)j * stride];
        buf[2 * j + 1] = z[2 * (size_t)j * stride + 1];
    }
    if ((n & (n - 1)) == 0) {
        /* bit reversal */
        for (int i = 1, j = 0; i < n; ++i) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                double tr = buf[2 * i], ti = buf[2 * i + 1];
                buf[2 * i] = buf[2 * j]; buf[2 * i + 1] = buf[2 * j + 1];
                buf[2 * j] = tr; buf[2 * j + 1] = ti;
            }
        }
        double *tw = (double *)malloc(sizeof(double) * (size_t)n);   /* n/2 complex twiddles */
        for (int k = 0; k < n / 2; ++k) twiddle(n, k, sign, &tw[2 * k], &tw[2 * k + 1]);
        for (int len = 2; len <= n; len <<= 1) {
            int step = n / len;
            for (int i = 0; i < n; i += len) {
                for (int k = 0; k < len / 2; ++k) {
                    double wr = tw[2 * k * step], wi = tw[2 * k * step + 1];
                    double *a = &buf[2 * (i + k)], *b = &buf[2 * (i + k + len / 2)];
                    double br = b[0] * wr - b[1] * wi;
                    double bi = b[0] * wi + b[1] * wr;
                    b[0] = a[0] - br; b[1] = a[1] - bi;
                    a[0] = a[0] + br; a[1] = a[1] + bi;
                }
            }
        }
        free(tw);
        for (int j = 0; j < n; ++j) {
            z[2 * (size_t)j * stride] = buf[2 * j];
            z[2 * (size_t)j * stride + 1] = buf[2 * j + 1];
        }
    } else {
        for (int k = 0; k < n; ++k) {
            long double sr = 0.0L, si = 0.0L;
            for (int j = 0; j < n; ++j) {
                long double a = 2.0L * 3.141592653589793238462643383279502884L *
                                (long double)(((long)j * k) % n) / (long double)n;
                long double c = cosl(a), s = sign * sinl(a);
                sr += buf[2 * j] * c - buf[2 * j + 1] * s;
                si += buf[2 * j] * s + buf[2 * j + 1] * c;
            }
            z[2 * (size_t)k * stride] = (double)sr;
            z[2 * (size_t)k * stride + 1] = (double)si;
        }
    }
    free(buf);
}

/* 1-D r2c along a contiguous real row of length n -> n/2+1 complex. */
static void r2c_row(int n, const double *x, double *out)
{
    double *z = (double *)calloc(2 * (size_t)n, sizeof(double));
    for (int j = 0; j < n; ++j) z[2 * j] = x[j];
    orc_dft(n, z, 1, -1);
    memcpy(out, z, sizeof(double) * 2 * (size_t)(n / 2 + 1));
    free(z);
}

/* 1-D c2r: n/2+1 complex -> n real.  Imaginary parts of DC and (even n)
   Nyquist are not read, as FFTW's halfcomplex c2r codelets. */
static void c2r_row(int n, const double *in, double *x)
{
    double *z = (double *)calloc(2 * (size_t)n, sizeof(double));
    int nh = n / 2 + 1;
    z[0] = in[0];
    z[1] = 0.0;
    for (int k = 1; k < nh; ++k) {
        if (2 * k == n) { z[2 * k] = in[2 * k]; z[2 * k + 1] = 0.0; continue; }
        z[2 * k] = in[2 * k];
        z[2 * k + 1] = in[2 * k + 1];
        z[2 * (n - k)] = in[2 * k];
        z[2 * (n - k) + 1] = -in[2 * k + 1];
    }
    orc_dft(n, z, 1, +1);
    for (int j = 0; j < n; ++j) x[j] = z[2 * j];
    free(z);
}

void orc_r2c_2d(int nx, int ny, const double *in, double *out)
{
    int nh = nx / 2 + 1;
    for (int j = 0; j < ny; ++j) r2c_row(nx, in + (size_t)j * nx, out + 2 * (size_t)j * nh);
    for (int i = 0; i < nh; ++i) orc_dft(ny, out + 2 * (size_t)i, nh, -1);
}

void orc_c2r_2d(int nx, int ny, const double *in, double *out)
{
    int nh = nx / 2 + 1;
    double *t = (double *)malloc(sizeof(double) * 2 * (size_t)nh * ny);
    memcpy(t, in, sizeof(double) * 2 * (size_t)nh * ny);
    for (int i = 0; i < nh; ++i) orc_dft(ny, t + 2 * (size_t)i, nh, +1);
    for (int j = 0; j < ny; ++j) c2r_row(nx, t + 2 * (size_t)j * nh, out + (size_t)j * nx);
    free(t);
}

void orc_r2c_3d(int nx, int ny, int nz, const double *in, double *out)
{
    int nh = nx / 2 + 1;
    size_t plane = (size_t)nh * ny;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            r2c_row(nx, in + ((size_t)k * ny + j) * nx, out + 2 * (k * plane + (size_t)j * nh));
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nh; ++i) orc_dft(ny, out + 2 * (k * plane + i), nh, -1);
    for (size_t p = 0; p < plane; ++p) orc_dft(nz, out + 2 * p, (long)plane, -1);
}

void orc_c2r_3d(int nx, int ny, int nz, const double *in, double *out)
{
    int nh = nx / 2 + 1;
    size_t plane = (size_t)nh * ny;
    double *t = (double *)malloc(sizeof(double) * 2 * plane * nz);
    memcpy(t, in, sizeof(double) * 2 * plane * nz);
    for (size_t p = 0; p < plane; ++p) orc_dft(nz, t + 2 * p, (long)plane, +1);
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nh; ++i) orc_dft(ny, t + 2 * (k * plane + i), nh, +1);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            c2r_row(nx, t + 2 * (k * plane + (size_t)j * nh), out + ((size_t)k * ny + j) * nx);
    free(t);
}

/* ------------------------------------------------------------------------ */
/* 2D incompressible hydro: src/derive2csh.f95, src/ab2csh.f95              */
/* ------------------------------------------------------------------------ */

/* Wavenumbers of derive2csh: pi is the default-real literal 3.1428
   (src/derive2csh.f95:16), Lx = Ly = 2 pi (:20-21), kx = 2 pi (i-1)/Lx (:30),
   ky = 2 pi (j-1)/Ly for j <= Ny/2 and 2 pi ((j-1)-Ny)/Ly above (:32, :48). */
static double csh_k(int idx0, int n, int halved)
{
    const double pi = (double)3.1428f;
    const double L = 2.0 * pi;
    double m = (halved || idx0 < n / 2) ? (double)idx0 : ((double)idx0 - (double)n);
    return 2.0 * pi * m / L;
}

void orc_derive2csh(int nx, int ny, double nu, const double *w, double *dt_new)
{
    int nh = nx / 2 + 1;
    size_t ns = (size_t)nh * ny, nr = (size_t)nx * ny;
    double *ukx = calloc(2 * ns, 8), *uky = calloc(2 * ns, 8), *ikxw = calloc(2 * ns, 8), *ikyw = calloc(2 * ns, 8);
    double *ux = malloc(8 * nr), *uy = malloc(8 * nr), *wx = malloc(8 * nr), *wy = malloc(8 * nr);
    double *nlkx = malloc(16 * ns), *nlky = malloc(16 * ns);
    /* psi, u spectra (src/derive2csh.f95:29-57) and i k omega (:59-71) */
    for (int j = 0; j < ny; ++j) {
        double ky = csh_k(j, ny, 0);
        for (int i = 0; i < nh; ++i) {
            double kx = csh_k(i, nx, 1);
            size_t p = (size_t)j * nh + i;
            double wr = RE(w, p), wi = IM(w, p);
            double pr, pim;
            if (i == 0 && j == 0) { pr = 0.0; pim = 0.0; }
            else { double k2 = kx * kx + ky * ky; pr = wr / k2; pim = wi / k2; }
            if (i == 0 && j == 0) { RE(ukx, p) = 0; IM(ukx, p) = 0; RE(uky, p) = 0; IM(uky, p) = 0; }
            else {
                RE(ukx, p) = -(ky * pim); IM(ukx, p) = ky * pr;       /* +i ky psi */
                RE(uky, p) = kx * pim;    IM(uky, p) = -(kx * pr);    /* -i kx psi */
            }
            RE(ikxw, p) = -(kx * wi); IM(ikxw, p) = kx * wr;
            RE(ikyw, p) = -(ky * wi); IM(ikyw, p) = ky * wr;
        }
    }
    /* four c2r (:73-87), normalise and products (:89-98) */
    orc_c2r_2d(nx, ny, ukx, ux);
    orc_c2r_2d(nx, ny, uky, uy);
    orc_c2r_2d(nx, ny, ikxw, wx);
    orc_c2r_2d(nx, ny, ikyw, wy);
    double norm = (double)nx * (double)ny;
    for (size_t q = 0; q < nr; ++q) {
        double a = ux[q] / norm, b = uy[q] / norm, c = wx[q] / norm, d = wy[q] / norm;
        wx[q] = a * c;
        wy[q] = b * d;
    }
    /* two r2c (:100-106) */
    orc_r2c_2d(nx, ny, wx, nlkx);
    orc_r2c_2d(nx, ny, wy, nlky);
    /* circular dealias (:108-125), viscous term (:127-139), negate (:141-145) */
    double cut = ((double)(nx + ny) / 2.0) / 3.0;
    for (int j = 0; j < ny; ++j) {
        double ky = csh_k(j, ny, 0);
        for (int i = 0; i < nh; ++i) {
            double kx = csh_k(i, nx, 1);
            size_t p = (size_t)j * nh + i;
            double ar = RE(nlkx, p), ai = IM(nlkx, p), br = RE(nlky, p), bi = IM(nlky, p);
            if (sqrt(kx * kx + ky * ky) >= cut) { ar = ai = br = bi = 0.0; }
            double vr = nu * (kx * kx + ky * ky);
            double nr_ = (ar + br) + vr * RE(w, p);
            double ni_ = (ai + bi) + vr * IM(w, p);
            RE(dt_new, p) = -nr_;
            IM(dt_new, p) = -ni_;
        }
    }
    free(ukx); free(uky); free(ikxw); free(ikyw); free(ux); free(uy); free(wx); free(wy);
    free(nlkx); free(nlky);
}

/* src/ab2csh.f95:10-14: new = w + (3/2 f_n - 1/2 f_{n-1}) dt */
void orc_ab2csh(int nx, int ny, double dt, const double *w, double *wn,
                const double *f_old, const double *f_new)
{
    size_t n2 = 2 * (size_t)(nx / 2 + 1) * ny;
    for (size_t q = 0; q < n2; ++q) wn[q] = w[q] + (1.5 * f_new[q] - 0.5 * f_old[q]) * dt;
}

/* Loop of src/main2csh.f95:114-126 (derive, ab2, rotate).  hist is the
   f_{n-1} array; the caller chooses its initial contents (zero by contract). */
void orc_ab2csh_steps(int nx, int ny, double nu, double dt, int nsteps, double *w, double *hist)
{
    size_t ns = (size_t)(nx / 2 + 1) * ny;
    double *fn = malloc(16 * ns), *wn = malloc(16 * ns);
    for (int t = 0; t < nsteps; ++t) {
        orc_derive2csh(nx, ny, nu, w, fn);
        orc_ab2csh(nx, ny, dt, w, wn, hist, fn);
        memcpy(hist, fn, 16 * ns);
        memcpy(w, wn, 16 * ns);
    }
    free(fn); free(wn);
}

/* Classic RK4 over derive2csh: the intended scheme of src/rk42css.f95:15-48
   (k1..k4 at stage offsets dt/2, dt/2, dt; w += dt/6 (k1 + 2k2 + 2k3 + k4)). */
void orc_rk4csh_steps(int nx, int ny, double nu, double dt, int nsteps, double *w)
{
    size_t n2 = 2 * (size_t)(nx / 2 + 1) * ny;
    double *k1 = malloc(8 * n2), *k2 = malloc(8 * n2), *k3 = malloc(8 * n2), *k4 = malloc(8 * n2), *s = malloc(8 * n2);
    for (int t = 0; t < nsteps; ++t) {
        orc_derive2csh(nx, ny, nu, w, k1);
        for (size_t q = 0; q < n2; ++q) s[q] = w[q] + (0.5 * dt) * k1[q];
        orc_derive2csh(nx, ny, nu, s, k2);
        for (size_t q = 0; q < n2; ++q) s[q] = w[q] + (0.5 * dt) * k2[q];
        orc_derive2csh(nx, ny, nu, s, k3);
        for (size_t q = 0; q < n2; ++q) s[q] = w[q] + dt * k3[q];
        orc_derive2csh(nx, ny, nu, s, k4);
        for (size_t q = 0; q < n2; ++q)
            w[q] = w[q] + (dt / 6.0) * (((k1[q] + 2.0 * k2[q]) + 2.0 * k3[q]) + k4[q]);
    }
    free(k1); free(k2); free(k3); free(k4); free(s);
}

/* src/derive2css.f95:6-136, screened vorticity.  The reference leaves Lx, Ly
   and lambda uninitialised (:11); this restatement takes Lx = Ly = 2 pi with
   the routine's own pi = 3.1428 (:17), as derive2csh does, so kx, ky are the
   integers of csh_k, and lambda as an argument.  psi(0,0) = 0 (:24-27);
   omega_d = k^2 psi (:36, 48) carries the derivatives (:37-38) and the viscous
   term (:124); the dealiased sum of the two products is scaled by alpha (:123). */
void orc_derive2css(int nx, int ny, double alpha, double nu, double lambda, const double *w, double *dt_new)
{
    int nh = nx / 2 + 1;
    size_t ns = (size_t)nh * ny, nr = (size_t)nx * ny;
    double *ukx = calloc(2 * ns, 8), *uky = calloc(2 * ns, 8), *ikxw = calloc(2 * ns, 8), *ikyw = calloc(2 * ns, 8);
    double *wd = calloc(2 * ns, 8);
    double *ux = malloc(8 * nr), *uy = malloc(8 * nr), *wx = malloc(8 * nr), *wy = malloc(8 * nr);
    double *nlkx = malloc(16 * ns), *nlky = malloc(16 * ns);
    for (int j = 0; j < ny; ++j) {
        double ky = csh_k(j, ny, 0);
        for (int i = 0; i < nh; ++i) {
            double kx = csh_k(i, nx, 1);
            size_t p = (size_t)j * nh + i;
            double k2 = kx * kx + ky * ky;
            double pr = 0.0, pim = 0.0;
            if (!(i == 0 && j == 0)) { double den = k2 + lambda; pr = RE(w, p) / den; pim = IM(w, p) / den; }
            RE(ukx, p) = -(ky * pim); IM(ukx, p) = ky * pr;       /* +i ky psi */
            RE(uky, p) = kx * pim;    IM(uky, p) = -(kx * pr);    /* -i kx psi */
            double dr = k2 * pr, di = k2 * pim;                   /* omega_d */
            RE(wd, p) = dr; IM(wd, p) = di;
            RE(ikxw, p) = -(kx * di); IM(ikxw, p) = kx * dr;
            RE(ikyw, p) = -(ky * di); IM(ikyw, p) = ky * dr;
        }
    }
    orc_c2r_2d(nx, ny, ukx, ux);
    orc_c2r_2d(nx, ny, uky, uy);
    orc_c2r_2d(nx, ny, ikxw, wx);
    orc_c2r_2d(nx, ny, ikyw, wy);
    double norm = (double)nx * (double)ny;
    for (size_t q = 0; q < nr; ++q) {
        double a = ux[q] / norm, b = uy[q] / norm, c = wx[q] / norm, d = wy[q] / norm;
        wx[q] = a * c;
        wy[q] = b * d;
    }
    orc_r2c_2d(nx, ny, wx, nlkx);
    orc_r2c_2d(nx, ny, wy, nlky);
    double cut = ((double)(nx + ny) / 2.0) / 3.0;
    for (int j = 0; j < ny; ++j) {
        double ky = csh_k(j, ny, 0);
        for (int i = 0; i < nh; ++i) {
            double kx = csh_k(i, nx, 1);
            size_t p = (size_t)j * nh + i;
            double ar = RE(nlkx, p), ai = IM(nlkx, p), br = RE(nlky, p), bi = IM(nlky, p);
            if (sqrt(kx * kx + ky * ky) >= cut) { ar = ai = br = bi = 0.0; }
            double vr = nu * (kx * kx + ky * ky);
            RE(dt_new, p) = -(alpha * (ar + br) + vr * RE(wd, p));
            IM(dt_new, p) = -(alpha * (ai + bi) + vr * IM(wd, p));
        }
    }
    free(ukx); free(uky); free(ikxw); free(ikyw); free(wd); free(ux); free(uy); free(wx); free(wy);
    free(nlkx); free(nlky);
}

/* Classic RK4 over derive2css (the scheme src/rk42css.f95:15-48 intends). */
void orc_rk4css_steps(int nx, int ny, double alpha, double nu, double lambda, double dt, int nsteps, double *w)
{
    size_t n2 = 2 * (size_t)(nx / 2 + 1) * ny;
    double *k1 = malloc(8 * n2), *k2 = malloc(8 * n2), *k3 = malloc(8 * n2), *k4 = malloc(8 * n2), *s = malloc(8 * n2);
    for (int t = 0; t < nsteps; ++t) {
        orc_derive2css(nx, ny, alpha, nu, lambda, w, k1);
        for (size_t q = 0; q < n2; ++q) s[q] = w[q] + (0.5 * dt) * k1[q];
        orc_derive2css(nx, ny, alpha, nu, lambda, s, k2);
        for (size_t q = 0; q < n2; ++q) s[q] = w[q] + (0.5 * dt) * k2[q];
        orc_derive2css(nx, ny, alpha, nu, lambda, s, k3);
        for (size_t q = 0; q < n2; ++q) s[q] = w[q] + dt * k3[q];
        orc_derive2css(nx, ny, alpha, nu, lambda, s, k4);
        for (size_t q = 0; q < n2; ++q)
            w[q] = w[q] + (dt / 6.0) * (((k1[q] + 2.0 * k2[q]) + 2.0 * k3[q]) + k4[q]);
    }
    free(k1); free(k2); free(k3); free(k4); free(s);
}

/* ------------------------------------------------------------------------ */
/* 2D compressible families: src/derive2coh.f95, src/derive2com.f95          */
/* ------------------------------------------------------------------------ */

/* kx = 2 pi (i-1)/Lx, ky = 2 pi (j-1)/Ly or 2 pi ((j-1)-Ny)/Ly with the
   caller's pi (src/derive2coh.f95:59-60, 74-75). */
static double f2_k(int i, int n, double pi, double l, int half)
{
    int m = (half || i < n / 2) ? i : i - n;
    return 2.0 * pi * (double)m / l;
}

static int f2_cut(double kx, double ky, int nx, int ny)
{
    return sqrt(kx * kx + ky * ky) >= ((double)(nx + ny) / 2.0) / 3.0 + 0;
}

/* spectral multiply helpers on interleaved complex arrays */
static void spec_ik(const double *z, double *o, size_t p, double k)
{
    double re = z[2 * p], im = z[2 * p + 1];
    o[2 * p] = -(k * im);
    o[2 * p + 1] = k * re;
}

static void spec_k2(const double *z, double *o, size_t p, double k)
{
    double kk = k * k;
    o[2 * p] = kk * z[2 * p];
    o[2 * p + 1] = kk * z[2 * p + 1];
}

/* src/derive2coh.f95:5-265.  state = {nk, ukx, uky}, rhs = {dn, dux, duy}. */
void orc_derive2coh(int nx, int ny, const orc_coh_params *prm, const double *const state[3],
                    double *const rhs[3])
{
    int nh = nx / 2 + 1;
    size_t ns = (size_t)nh * ny, nr = (size_t)nx * ny;
    /* 13 spectra to invert (:56-129) */
    double *sp[13], *re[13];
    for (int f = 0; f < 13; ++f) { sp[f] = calloc(2 * ns, 8); re[f] = malloc(8 * nr); }
    const double *nk = state[0], *ukx = state[1], *uky = state[2];
    for (int j = 0; j < ny; ++j) {
        double ky = f2_k(j, ny, prm->pi, prm->ly, 0);
        for (int i = 0; i < nh; ++i) {
            double kx = f2_k(i, nx, prm->pi, prm->lx, 1);
            size_t p = (size_t)j * nh + i;
            sp[0][2 * p] = nk[2 * p]; sp[0][2 * p + 1] = nk[2 * p + 1];
            sp[1][2 * p] = ukx[2 * p]; sp[1][2 * p + 1] = ukx[2 * p + 1];
            sp[2][2 * p] = uky[2 * p]; sp[2][2 * p + 1] = uky[2 * p + 1];
            spec_ik(nk, sp[3], p, kx);   /* i kx n  */
            spec_ik(nk, sp[4], p, ky);   /* i ky n  */
            spec_ik(ukx, sp[5], p, kx);  /* i kx ux */
            spec_ik(ukx, sp[6], p, ky);  /* i ky ux */
            spec_ik(uky, sp[7], p, kx);  /* i kx uy */
            spec_ik(uky, sp[8], p, ky);  /* i ky uy */
            spec_k2(ukx, sp[9], p, kx);  /* kx^2 ux */
            spec_k2(ukx, sp[10], p, ky); /* ky^2 ux */
            spec_k2(uky, sp[11], p, kx); /* kx^2 uy */
            spec_k2(uky, sp[12], p, ky); /* ky^2 uy */
        }
    }
    for (int f = 0; f < 13; ++f) orc_c2r_2d(nx, ny, sp[f], re[f]);
    double norm = (double)nx * (double)ny;
    /* 9 real-space terms (:131-167) */
    double *t[9];
    for (int f = 0; f < 9; ++f) t[f] = malloc(8 * nr);
    for (size_t q = 0; q < nr; ++q) {
        double n = re[0][q] / norm, ux = re[1][q] / norm, uy = re[2][q] / norm;
        double dnx = re[3][q] / norm, dny = re[4][q] / norm;
        double duxx = re[5][q] / norm, duxy = re[6][q] / norm, duyx = re[7][q] / norm, duyy = re[8][q] / norm;
        double d2uxx = re[9][q] / norm, d2uxy = re[10][q] / norm, d2uyx = re[11][q] / norm, d2uyy = re[12][q] / norm;
        t[0][q] = dnx / n;                  /* Px */
        t[1][q] = dny / n;                  /* Py */
        t[2][q] = d2uxx / n;                /* Viscxx */
        t[3][q] = d2uxy / n;                /* Viscxy */
        t[4][q] = d2uyx / n;                /* Viscyx */
        t[5][q] = d2uyy / n;                /* Viscyy */
        t[6][q] = ux * dnx + uy * dny + n * (duxx + duyy);  /* NLn */
        t[7][q] = ux * duxx + uy * duxy;    /* NLx */
        t[8][q] = ux * duyx + uy * duyy;    /* NLy */
    }
    double *tk[9];
    for (int f = 0; f < 9; ++f) { tk[f] = malloc(16 * ns); orc_r2c_2d(nx, ny, t[f], tk[f]); }
    /* dealias all (:201-242), assemble (:246-258) */
    double cs2 = prm->cs * prm->cs;
    for (int j = 0; j < ny; ++j) {
        double ky = f2_k(j, ny, prm->pi, prm->ly, 0);
        for (int i = 0; i < nh; ++i) {
            double kx = f2_k(i, nx, prm->pi, prm->lx, 1);
            size_t p = (size_t)j * nh + i;
            int cut = f2_cut(kx, ky, nx, ny);
            for (int c = 0; c < 2; ++c) {
                double v[9];
                for (int f = 0; f < 9; ++f) v[f] = cut ? 0.0 : tk[f][2 * p + c];
                double nlx = v[7] + (prm->mux * v[2] + prm->muy * v[3]) + cs2 * v[0];
                double nly = v[8] + (prm->mux * v[4] + prm->muy * v[5]) + cs2 * v[1];
                rhs[0][2 * p + c] = -v[6];
                rhs[1][2 * p + c] = -nlx;
                rhs[2][2 * p + c] = -nly;
            }
        }
    }
    for (int f = 0; f < 13; ++f) { free(sp[f]); free(re[f]); }
    for (int f = 0; f < 9; ++f) { free(t[f]); free(tk[f]); }
}

/* src/derive2com.f95:5-426.  state = {rho, rho ux, rho uy, E, Bx, By}.  The
   forcing uses x(i) = (i-1) Lx/Nx, y(j) = (j-1) Ly/Ny (src/main2com.f95:231-233;
   the routine's own x, y are uninitialised).  curl B follows the reference's
   lines :172-175 as written. */
void orc_derive2com(int nx, int ny, const orc_com_params *prm, const double *const state[6],
                    double *const rhs[6])
{
    int nh = nx / 2 + 1;
    size_t ns = (size_t)nh * ny, nr = (size_t)nx * ny;
    double norm = (double)nx * (double)ny;
    double *r[6];
    for (int f = 0; f < 6; ++f) { r[f] = malloc(8 * nr); orc_c2r_2d(nx, ny, state[f], r[f]); }
    double *ux = malloc(8 * nr), *uy = malloc(8 * nr), *fx = malloc(8 * nr), *fy = malloc(8 * nr);
    double dx = prm->lx / (double)nx, dy = prm->ly / (double)ny;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t q = (size_t)j * nx + i;
            for (int f = 0; f < 6; ++f) r[f][q] = r[f][q] / norm;
            ux[q] = r[1][q] / r[0][q];
            uy[q] = r[2][q] / r[0][q];
            fx[q] = -(r[0][q] * prm->a * sin(prm->kf * ((double)j * dy)));
            fy[q] = r[0][q] * prm->a * sin(prm->kf * ((double)i * dx));
        }
    double *uxk = malloc(16 * ns), *uyk = malloc(16 * ns), *fxk = malloc(16 * ns), *fyk = malloc(16 * ns);
    orc_r2c_2d(nx, ny, ux, uxk);
    orc_r2c_2d(nx, ny, uy, uyk);
    orc_r2c_2d(nx, ny, fx, fxk);
    orc_r2c_2d(nx, ny, fy, fyk);
    double *a = calloc(2 * ns, 8), *b = calloc(2 * ns, 8);
    for (int j = 0; j < ny; ++j) {
        double ky = f2_k(j, ny, prm->pi, prm->ly, 0);
        for (int i = 0; i < nh; ++i) {
            double kx = f2_k(i, nx, prm->pi, prm->lx, 1);
            size_t p = (size_t)j * nh + i;
            spec_ik(uxk, a, p, kx);
            spec_ik(uyk, b, p, ky);
        }
    }
    double *duxdx = malloc(8 * nr), *duydy = malloc(8 * nr);
    orc_c2r_2d(nx, ny, a, duxdx);
    orc_c2r_2d(nx, ny, b, duydy);
    /* fluxes (:169-226) */
    double *fl[9];
    for (int f = 0; f < 9; ++f) fl[f] = malloc(8 * nr);
    double cs2 = prm->cs * prm->cs;
    for (size_t q = 0; q < nr; ++q) {
        double rho = r[0][q], rux = r[1][q], ruy = r[2][q], E = r[3][q], bx = r[4][q], by = r[5][q];
        double vx = ux[q], vy = uy[q];
        double b2 = bx * bx + by * by;
        double dudx = duxdx[q] / norm, dvdy = duydy[q] / norm;
        double dbxdy = dudx / norm, dbydx = dvdy / norm;
        double curl = dbydx - dbxdy;
        double P = cs2 * rho;
        fl[0][q] = rux * vx + P + b2 / 2.0 - bx * bx;   /* Mom_x_1 */
        fl[1][q] = rux * vy - bx * by;                  /* Mom_x_2 */
        fl[2][q] = rux * vy - bx * by;                  /* Mom_y_1 */
        fl[3][q] = ruy * vy + P + b2 / 2.0 - by * by;   /* Mom_y_2 */
        double ex = (E + P + b2 / 2.0) * vx - vx * bx * bx - vy * bx * by;
        fl[4][q] = ex - prm->eta * by * curl;           /* Energy_x */
        double ey = (E + P + b2 / 2.0) * vy - vx * bx * by - vy * by * by;
        fl[5][q] = ey + prm->eta * bx * curl;           /* Energy_y */
        double dv = dudx + dvdy;
        fl[6][q] = dv * dv;                             /* E_Visc */
        fl[7][q] = vy * bx - by * vx;                   /* Mag_x */
        fl[8][q] = vx * by - bx * vy;                   /* Mag_y */
    }
    double *fk[9];
    for (int f = 0; f < 9; ++f) { fk[f] = malloc(16 * ns); orc_r2c_2d(nx, ny, fl[f], fk[f]); }
    /* spectral assembly (:252-420) */
    for (int j = 0; j < ny; ++j) {
        double ky = f2_k(j, ny, prm->pi, prm->ly, 0);
        for (int i = 0; i < nh; ++i) {
            double kx = f2_k(i, nx, prm->pi, prm->lx, 1);
            size_t p = (size_t)j * nh + i;
            int cut = f2_cut(kx, ky, nx, ny);
            double t[19][2];
#define IK(src, k, slot) do { double re_ = (src)[2 * p], im_ = (src)[2 * p + 1]; \
            t[slot][0] = -((k) * im_); t[slot][1] = (k) * re_; } while (0)
#define K2(src, k, slot) do { double kk_ = (k) * (k); t[slot][0] = kk_ * (src)[2 * p]; t[slot][1] = kk_ * (src)[2 * p + 1]; } while (0)
            IK(state[1], kx, 0);  IK(state[2], ky, 1);
            IK(fk[0], kx, 2);     IK(fk[1], ky, 3);     IK(fk[2], kx, 4);   IK(fk[3], ky, 5);
            K2(uxk, kx, 6);       K2(uxk, ky, 7);       K2(uyk, kx, 8);     K2(uyk, ky, 9);
            IK(fk[4], kx, 10);    IK(fk[5], ky, 11);
            t[12][0] = fk[6][2 * p]; t[12][1] = fk[6][2 * p + 1];
            IK(fk[7], ky, 13);    IK(fk[8], kx, 14);
            K2(state[4], kx, 15); K2(state[4], ky, 16); K2(state[5], kx, 17); K2(state[5], ky, 18);
#undef IK
#undef K2
            if (cut) for (int s_ = 0; s_ < 19; ++s_) t[s_][0] = t[s_][1] = 0.0;
            for (int c = 0; c < 2; ++c) {
                rhs[0][2 * p + c] = -(t[0][c] + t[1][c]);
                double d1 = -(t[2][c] + t[3][c]);
                rhs[1][2 * p + c] = d1 - prm->mu * (t[6][c] + t[7][c]) + fxk[2 * p + c];
                double d2 = -(t[4][c] + t[5][c]);
                rhs[2][2 * p + c] = d2 - prm->mu * (t[8][c] + t[9][c]) + fyk[2 * p + c];
                double d3 = -(t[10][c] + t[11][c]);
                rhs[3][2 * p + c] = d3 + prm->mu * t[12][c];
                rhs[4][2 * p + c] = -t[13][c] - prm->eta * (t[15][c] + t[16][c]);
                rhs[5][2 * p + c] = -t[14][c] - prm->eta * (t[17][c] + t[18][c]);
            }
        }
    }
    for (int f = 0; f < 6; ++f) free(r[f]);
    for (int f = 0; f < 9; ++f) { free(fl[f]); free(fk[f]); }
    free(ux); free(uy); free(fx); free(fy); free(uxk); free(uyk); free(fxk); free(fyk);
    free(a); free(b); free(duxdx); free(duydy);
}

/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
   easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies by
   0xD2511F53 and 0xCD9E8D57 with a Weyl key schedule (0x9E3779B9, 0xBB67AE85). */
void orc_philox4x32_10(const unsigned ctr[4], const unsigned key[2], unsigned out[4])
{
    unsigned c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        unsigned long long p0 = (unsigned long long)0xD2511F53u * c0;
        unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c2;
        unsigned n0 = (unsigned)(p1 >> 32) ^ c1 ^ k0, n1 = (unsigned)p1;
        unsigned n2 = (unsigned)(p0 >> 32) ^ c3 ^ k1, n3 = (unsigned)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Random-phase vorticity half spectrum (Fortran layout (Nh, Ny)): amplitude
   sqrt(k E(k)/pi), E(k) = k^4 exp(-(k/k0)^2), scaled to kinetic energy
   `energy`; phase 2 pi u, u = 53-bit uniform from the first two Philox words at
   counter {i + j*Nh, 0, 0, 0}, key {seed, 0}; Nyquist row/column and k = 0 zero;
   kx = 0 column Hermitian.  (The device generator this checks has no
   counterpart in the reference.) */
void orc_ic_random(int nx, int ny, unsigned seed, double k0, double energy, double *wk)
{
    const double pi = 3.14159265358979323846;
    int nh = nx / 2 + 1;
    double s = 0.0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nh; ++i) {
            size_t p = (size_t)j * nh + i;
            RE(wk, p) = 0.0; IM(wk, p) = 0.0;
            int jj = j, conj = 0;
            if (i == 0 && j > ny / 2) { jj = ny - j; conj = 1; }
            if (i == nh - 1 || j == ny / 2 || (i == 0 && jj == 0)) continue;
            double kx = i, ky = jj < ny / 2 ? jj : jj - ny;
            double k = sqrt(kx * kx + ky * ky);
            double e = k * k * k * k * exp(-(k / k0) * (k / k0));
            double amp = sqrt(k * e / pi);
            unsigned ctr[4] = {(unsigned)(i + (size_t)jj * nh), 0, 0, 0}, key[2] = {seed, 0}, r[4];
            orc_philox4x32_10(ctr, key, r);
            double u = ((double)(r[0] >> 5) * 67108864.0 + (double)(r[1] >> 6)) / 9007199254740992.0;
            double ph = 2.0 * pi * u;
            RE(wk, p) = amp * cos(ph);
            IM(wk, p) = conj ? -amp * sin(ph) : amp * sin(ph);
            if (i < nh - 1) s += (i == 0 ? 1.0 : 2.0) * amp * amp / (k * k);
        }
    /* kinetic energy of the unscaled field: the kx = 0 column above counted
       both +ky and -ky once each, the other columns twice */
    double ke = 0.5 * s / ((double)nx * ny * (double)nx * ny);
    double scale = ke > 0 ? sqrt(energy / ke) : 0.0;
    for (size_t p = 0; p < 2 * (size_t)nh * ny; ++p) wk[p] *= scale;
}

/* Uniform double in [0, 1) from two Philox words (53 random bits). */
static double orc_u53(unsigned a, unsigned b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0;
}

/* One random-phase solenoidal vector field (three half spectra v[c], each
 * (Nh,Ny,Nz)) with |v(k)| per component k exp(-(k/k0)^2), i.e. shell
 * spectrum ~ k^4 exp(-2 (k/k0)^2), scaled to 1/2 <|v|^2> = energy.  Mode
 * (i,j,l) takes the three phases 2 pi u_c from Philox-4x32-10 with key
 * {seed, tag} and counters {m, 0, 0, 0}, {m, 1, 0, 0}, m = i + Nh (j + Ny l):
 * u_0 from words 0-1 of the first draw, u_1 from words 2-3, u_2 from words
 * 0-1 of the second.  On the kx = 0 plane the mode whose mirror (-ky, -kz)
 * has the smaller index j + Ny l takes the conjugate of the mirror's draw
 * (Hermitian).  The raw vector is projected with its own wavevector,
 * v <- v - k (k . v)/k^2.  k = 0 and the Nyquist planes (i = Nh-1,
 * j = Ny/2, l = Nz/2) are zero.  No counterpart in the reference, whose 3D
 * initial state is the deterministic ABC flow (src/main3com.f95:290-324). */
static void ic_solenoidal(int nx, int ny, int nz, unsigned seed, unsigned tag, double k0, double energy,
                          double *const v[3])
{
    const double pi = 3.14159265358979323846;
    const int nh = nx / 2 + 1;
    double s = 0.0;
    for (int l = 0; l < nz; ++l)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nh; ++i) {
                const size_t p = (size_t)i + (size_t)nh * ((size_t)j + (size_t)ny * l);
                for (int c = 0; c < 3; ++c) { RE(v[c], p) = 0.0; IM(v[c], p) = 0.0; }
                if (i == nh - 1 || j == ny / 2 || l == nz / 2 || (i == 0 && j == 0 && l == 0)) continue;
                int jc = j, lc = l, conj = 0;
                if (i == 0) {
                    const int jm = (ny - j) % ny, lm = (nz - l) % nz;
                    if (jm + ny * lm < j + ny * l) { jc = jm; lc = lm; conj = 1; }
                }
                const double kv[3] = {(double)i, (double)(j < ny / 2 ? j : j - ny), (double)(l < nz / 2 ? l : l - nz)};
                const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
                const double k = sqrt(k2);
                const double amp = k * exp(-(k / k0) * (k / k0));
                const unsigned m = (unsigned)((size_t)i + (size_t)nh * ((size_t)jc + (size_t)ny * lc));
                unsigned c0[4] = {m, 0, 0, 0}, c1[4] = {m, 1, 0, 0}, key[2] = {seed, tag}, r[4], q[4];
                orc_philox4x32_10(c0, key, r);
                orc_philox4x32_10(c1, key, q);
                const double u[3] = {orc_u53(r[0], r[1]), orc_u53(r[2], r[3]), orc_u53(q[0], q[1])};
                double re[3], im[3];
                for (int c = 0; c < 3; ++c) {
                    const double ph = 2.0 * pi * u[c];
                    re[c] = amp * cos(ph);
                    im[c] = conj ? -(amp * sin(ph)) : amp * sin(ph);
                }
                const double dr = (kv[0] * re[0] + kv[1] * re[1] + kv[2] * re[2]) / k2;
                const double di = (kv[0] * im[0] + kv[1] * im[1] + kv[2] * im[2]) / k2;
                const double w = i == 0 ? 1.0 : 2.0;
                for (int c = 0; c < 3; ++c) {
                    RE(v[c], p) = re[c] - kv[c] * dr;
                    IM(v[c], p) = im[c] - kv[c] * di;
                    s += w * (RE(v[c], p) * RE(v[c], p) + IM(v[c], p) * IM(v[c], p));
                }
            }
    /* 1/2 <|v|^2> = 1/2 sum over the full spectrum |v_k|^2 / (Nx Ny Nz)^2 */
    const double nn = (double)nx * ny * nz;
    const double e = 0.5 * s / (nn * nn);
    const double scale = e > 0 ? sqrt(energy / e) : 0.0;
    for (int c = 0; c < 3; ++c)
        for (size_t p = 0; p < 2 * (size_t)nh * ny * nz; ++p) v[c][p] *= scale;
}

void orc_ic_random3d(int nx, int ny, int nz, unsigned seed, double k0, double e_kin, double e_mag,
                     double *const st[8])
{
    const int nh = nx / 2 + 1;
    const size_t ns = (size_t)nh * ny * nz, nr = (size_t)nx * ny * nz;
    double *u[3] = {st[1], st[2], st[3]}, *b[3] = {st[5], st[6], st[7]};
    ic_solenoidal(nx, ny, nz, seed, 1u, k0, e_kin, u);
    ic_solenoidal(nx, ny, nz, seed, 2u, k0, e_mag, b);
    /* rho = 1: its unnormalised transform is Nx Ny Nz at k = 0; rho u = u */
    memset(st[0], 0, 2 * ns * sizeof(double));
    st[0][0] = (double)nr;
    /* E = P + rho u^2/2 + B^2/2 with P = 1 (as src/main3com.f95:324-325 forms it) */
    double *re = (double *)malloc(nr * sizeof(double)), *acc = (double *)calloc(nr, sizeof(double));
    double *tmp = (double *)malloc(2 * ns * sizeof(double));
    const double inv = 1.0 / (double)nr;
    for (int f = 0; f < 6; ++f) {
        memcpy(tmp, f < 3 ? u[f] : b[f - 3], 2 * ns * sizeof(double));
        orc_c2r_3d(nx, ny, nz, tmp, re);
        for (size_t q = 0; q < nr; ++q) {
            const double x = re[q] * inv;
            acc[q] += x * x;
        }
    }
    for (size_t q = 0; q < nr; ++q) re[q] = 1.0 + 0.5 * acc[q];
    orc_r2c_3d(nx, ny, nz, re, st[4]);
    free(re);
    free(acc);
    free(tmp);
}

/* Shear-layer IC of src/main2csh.f95:84-110 (x-fastest real field, then r2c). */
void orc_ic2csh(int nx, int ny, double lx_in, double ly_in, double w0, double m0,
                double d_in, double *omega, double *omegak)
{
    const double pi = 3.14159265358979323846;
    double lx = lx_in * pi, ly = ly_in * pi;
    double dx = lx / (double)nx, dy = ly / (double)ny, d = d_in * pi;
    for (int j = 0; j < ny; ++j) {
        double y = (double)j * dy;
        for (int i = 0; i < nx; ++i) {
            double x = (double)i * dx;
            double c1 = cosh((y + 0.5 * pi) / d), c2 = cosh((y - 0.5 * pi) / d);
            double c3 = cosh((y + 1.5 * pi) / d), c4 = cosh((y - 1.5 * pi) / d);
            double v = w0 / (c1 * c1) - w0 / (c2 * c2);
            v = v - w0 / (c3 * c3) + w0 / (c4 * c4);
            v = v + 0.10 * cos(m0 * x);
            omega[(size_t)j * nx + i] = v;
        }
    }
    orc_r2c_2d(nx, ny, omega, omegak);
}

/* ------------------------------------------------------------------------ */
/* 3D compressible isothermal MHD: src/derive3com.f95, src/ab3com.f95       */
/* Only the live dataflow is restated: E is frozen by ab3com (ab3com.f95:51) */
/* and P = CS^2 rho (derive3com.f95:395), so E, curl B, the energy fluxes    */
/* and the (commented-out) forcing never reach rho, rho u or B.             */
/* ------------------------------------------------------------------------ */

/* k = 2 pi (i-1)/L with true pi (derive3com.f95:245-247, 260-262); the
   second half of y and z uses (i-1)-N.  Returned in the reference's
   evaluation order ((2 pi) m) / L. */
static double com_k(int idx0, int n, int halved, double L)
{
    const double pi = 3.14159265358979323846;
    double m = (halved || idx0 < n / 2) ? (double)idx0 : (double)(idx0 - n);
    return 2.0 * pi * m / L;
}

void orc_derive3com(int nx, int ny, int nz, const orc_mhd_params *p,
                    const double *const st[8], double *const rhs[8])
{
    int nh = nx / 2 + 1;
    size_t ns = (size_t)nh * ny * nz, nr = (size_t)nx * ny * nz;
    double norm = (double)nx * (double)ny * (double)nz;
    /* c2r of rho, rho u, B (derive3com.f95:94-140); E is skipped (dead) */
    static const int live[7] = {0, 1, 2, 3, 5, 6, 7};
    double *r[8] = {0};
    for (int f = 0; f < 7; ++f) {
        r[live[f]] = malloc(8 * nr);
        orc_c2r_3d(nx, ny, nz, st[live[f]], r[live[f]]);
        for (size_t q = 0; q < nr; ++q) r[live[f]][q] = r[live[f]][q] / norm;
    }
    double *rho = r[0], *rux = r[1], *ruy = r[2], *ruz = r[3], *bx = r[5], *by = r[6], *bz = r[7];
    /* velocity, B^2 (:176-186), pressure and fluxes (:395-436) */
    enum { UX, UY, UZ, MX1, MX2, MX3, MY2, MY3, MZ1, MZ2, MZ3, MG1, MG2, MG3, NF };
    double *fl[NF];
    for (int f = 0; f < NF; ++f) fl[f] = malloc(8 * nr);
    double cs2 = p->cs * p->cs;
    for (size_t q = 0; q < nr; ++q) {
        double ux = rux[q] / rho[q], uy = ruy[q] / rho[q], uz = ruz[q] / rho[q];
        double b2 = bx[q] * bx[q] + by[q] * by[q] + bz[q] * bz[q];
        double P = cs2 * rho[q];
        fl[UX][q] = ux; fl[UY][q] = uy; fl[UZ][q] = uz;
        fl[MX1][q] = rux[q] * ux + P + b2 / 2.0 - bx[q] * bx[q];
        fl[MX2][q] = rux[q] * uy - bx[q] * by[q];             /* == Mom_y_1 */
        fl[MX3][q] = rux[q] * uz - bx[q] * bz[q];
        fl[MY2][q] = ruy[q] * uy + P + b2 / 2.0 - by[q] * by[q];
        fl[MY3][q] = ruy[q] * uz - by[q] * bz[q];
        fl[MZ1][q] = ruz[q] * ux - bz[q] * bx[q];
        fl[MZ2][q] = ruz[q] * uy - bz[q] * by[q];
        fl[MZ3][q] = ruz[q] * uz + P + b2 / 2.0 - bz[q] * bz[q];
        fl[MG1][q] = ux * by[q] - uy * bx[q];                 /* Mag_x_1 == Mag_y_1 */
        fl[MG2][q] = ux * bz[q] - uz * bx[q];                 /* Mag_x_2 == Mag_z_1 */
        fl[MG3][q] = uy * bz[q] - uz * by[q];                 /* Mag_y_2 == Mag_z_2 */
    }
    double *fk[NF];
    for (int f = 0; f < NF; ++f) { fk[f] = malloc(16 * ns); orc_r2c_3d(nx, ny, nz, fl[f], fk[f]); free(fl[f]); }
    for (int f = 0; f < 7; ++f) free(r[live[f]]);

    /* i k . flux, k^2 diffusion, spherical mask (:574-843), assembly (:1018-1037) */
    double cut = ((double)(nx + ny + nz) / 3.0) / 3.0 + 0;
    const double *rk[3] = {st[1], st[2], st[3]};
    for (int k = 0; k < nz; ++k) {
        double kz = com_k(k, nz, 0, p->lz);
        for (int j = 0; j < ny; ++j) {
            double ky = com_k(j, ny, 0, p->ly);
            for (int i = 0; i < nh; ++i) {
                double kx = com_k(i, nx, 1, p->lx);
                size_t q = ((size_t)k * ny + j) * nh + i;
                int m = sqrt(kx * kx + ky * ky + kz * kz) >= cut;
                double kk[3] = {kx, ky, kz};
                /* i k * z = (-k zi, k zr) */
#define IKR(kv, a) (m ? 0.0 : -((kv) * IM(a, q)))
#define IKI(kv, a) (m ? 0.0 : ((kv) * RE(a, q)))
#define K2R(kv, a) (m ? 0.0 : ((kv) * (kv)) * RE(a, q))
#define K2I(kv, a) (m ? 0.0 : ((kv) * (kv)) * IM(a, q))
                /* density: -(ikx rux + iky ruy + ikz ruz) on the state spectra */
                double sr = (IKR(kk[0], rk[0]) + IKR(kk[1], rk[1])) + IKR(kk[2], rk[2]);
                double si = (IKI(kk[0], rk[0]) + IKI(kk[1], rk[1])) + IKI(kk[2], rk[2]);
                RE(rhs[0], q) = -sr; IM(rhs[0], q) = -si;
                /* momentum: -(ik.Mom_row) - (k^2 u)/visc */
                const double *mom[3][3] = {{fk[MX1], fk[MX2], fk[MX3]},
                                           {fk[MX2], fk[MY2], fk[MY3]},
                                           {fk[MZ1], fk[MZ2], fk[MZ3]}};
                const double *uk[3] = {fk[UX], fk[UY], fk[UZ]};
                for (int c = 0; c < 3; ++c) {
                    double ar = (IKR(kx, mom[c][0]) + IKR(ky, mom[c][1])) + IKR(kz, mom[c][2]);
                    double ai = (IKI(kx, mom[c][0]) + IKI(ky, mom[c][1])) + IKI(kz, mom[c][2]);
                    double dr = (K2R(kx, uk[c]) + K2R(ky, uk[c])) + K2R(kz, uk[c]);
                    double di = (K2I(kx, uk[c]) + K2I(ky, uk[c])) + K2I(kz, uk[c]);
                    RE(rhs[1 + c], q) = -ar - dr * p->mu;
                    IM(rhs[1 + c], q) = -ai - di * p->mu;
                }
                RE(rhs[4], q) = 0.0; IM(rhs[4], q) = 0.0;
                /* induction (:1030-1037) */
                const double *bk[3] = {st[5], st[6], st[7]};
                double d2r[3], d2i[3];
                for (int c = 0; c < 3; ++c) {
                    d2r[c] = (K2R(kx, bk[c]) + K2R(ky, bk[c])) + K2R(kz, bk[c]);
                    d2i[c] = (K2I(kx, bk[c]) + K2I(ky, bk[c])) + K2I(kz, bk[c]);
                }
                double xr = IKR(ky, fk[MG1]) + IKR(kz, fk[MG2]);
                double xi = IKI(ky, fk[MG1]) + IKI(kz, fk[MG2]);
                RE(rhs[5], q) = xr - d2r[0] * p->eta; IM(rhs[5], q) = xi - d2i[0] * p->eta;
                double yr = IKR(kx, fk[MG1]) - IKR(kz, fk[MG3]);
                double yi = IKI(kx, fk[MG1]) - IKI(kz, fk[MG3]);
                RE(rhs[6], q) = -yr - d2r[1] * p->eta; IM(rhs[6], q) = -yi - d2i[1] * p->eta;
                double zr = IKR(kx, fk[MG2]) + IKR(ky, fk[MG3]);
                double zi = IKI(kx, fk[MG2]) + IKI(ky, fk[MG3]);
                RE(rhs[7], q) = -zr - d2r[2] * p->eta; IM(rhs[7], q) = -zi - d2i[2] * p->eta;
#undef IKR
#undef IKI
#undef K2R
#undef K2I
            }
        }
    }
    for (int f = 0; f < NF; ++f) free(fk[f]);
}

/* src/ab3com.f95:39-59: AB2 for rho, rho u, B; E is copied (frozen, :51). */
void orc_ab3com(int nx, int ny, int nz, double dt, const double *const st[8],
                double *const sn[8], const double *const fo[8], const double *const fn[8])
{
    size_t n2 = 2 * (size_t)(nx / 2 + 1) * ny * nz;
    for (int f = 0; f < 8; ++f)
        for (size_t q = 0; q < n2; ++q)
            sn[f][q] = (f == 4) ? st[f][q] : st[f][q] + (1.5 * fn[f][q] - 0.5 * fo[f][q]) * dt;
}
