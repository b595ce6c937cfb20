// derive2coh and derive2com on grids the fused 2D compressible passes do not
// take (any Nx >= 2, Ny >= 1), built from the any-length transforms.  Each
// routine's own sequence, one device kernel per step over half spectra of
// pitch spec_ld(nx) or real planes of pitch nx:
//   derive2coh (src/derive2coh.f95:56-258): thirteen spectra (the fields, i k
//     of each, k^2 of the velocity), thirteen c2r, nine real-space terms,
//     nine r2c, dealias and assembly;
//   derive2com (src/derive2com.f95:94-420): six c2r, velocity and forcing,
//     four r2c, the two velocity derivatives and their c2r, nine fluxes, nine
//     r2c, assembly.
// Wavenumbers 2 pi m / L with the caller's pi and box; the circular cut at
// (Nx+Ny)/2/3.  Slower than the fused path; the arithmetic is the same term
// by term.
#include "common.h"
#include "fft_api.h"
#include "generic_fam2d.h"

namespace tara {

namespace {

__device__ __forceinline__ double f2k(int i, int n, double pi, double l, bool half)
{
    const int m = (half || i < n / 2) ? i : i - n;
    return 2.0 * pi * (double)m / l;
}

__device__ __forceinline__ bool f2cut(double kx, double ky, int nx, int ny)
{
    return sqrt(kx * kx + ky * ky) >= ((double)(nx + ny) / 2.0) / 3.0;
}

__device__ __forceinline__ double2 ik(double2 z, double k) { return make_double2(-(k * z.y), k * z.x); }
__device__ __forceinline__ double2 k2m(double2 z, double k)
{
    const double kk = k * k;
    return make_double2(kk * z.x, kk * z.y);
}

unsigned blocks_for(size_t n)
{
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 16384 ? 16384 : (b ? b : 1));
}

struct Grid {
    int nx, ny, nh, ld;
    size_t plane, nr;
    double pi, lx, ly;
};

// ---------------------------------------------------------------- coh
__global__ void k_coh_spec(const double2* __restrict__ S, double2* __restrict__ SP, Grid g)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)g.nh * g.ny;
         q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % g.nh), j = (int)(q / g.nh);
        const double kx = f2k(i, g.nx, g.pi, g.lx, true), ky = f2k(j, g.ny, g.pi, g.ly, false);
        const size_t p = (size_t)j * g.ld + i;
        const double2 n = S[p], ux = S[g.plane + p], uy = S[2 * g.plane + p];
        const double2 v[13] = {n,          ux,         uy,          ik(n, kx),   ik(n, ky),
                               ik(ux, kx), ik(ux, ky), ik(uy, kx),  ik(uy, ky),  k2m(ux, kx),
                               k2m(ux, ky), k2m(uy, kx), k2m(uy, ky)};
        for (int f = 0; f < 13; ++f) SP[f * g.plane + p] = v[f];
    }
}

__global__ void k_coh_terms(const double* __restrict__ R, double* __restrict__ Tm, size_t nr, double norm)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < nr; q += (size_t)gridDim.x * blockDim.x) {
        const double n = R[q] / norm, ux = R[nr + q] / norm, uy = R[2 * nr + q] / norm;
        const double dnx = R[3 * nr + q] / norm, dny = R[4 * nr + q] / norm;
        const double duxx = R[5 * nr + q] / norm, duxy = R[6 * nr + q] / norm, duyx = R[7 * nr + q] / norm,
                     duyy = R[8 * nr + q] / norm;
        const double d2uxx = R[9 * nr + q] / norm, d2uxy = R[10 * nr + q] / norm, d2uyx = R[11 * nr + q] / norm,
                     d2uyy = R[12 * nr + q] / norm;
        Tm[q] = dnx / n;
        Tm[nr + q] = dny / n;
        Tm[2 * nr + q] = d2uxx / n;
        Tm[3 * nr + q] = d2uxy / n;
        Tm[4 * nr + q] = d2uyx / n;
        Tm[5 * nr + q] = d2uyy / n;
        Tm[6 * nr + q] = ux * dnx + uy * dny + n * (duxx + duyy);
        Tm[7 * nr + q] = ux * duxx + uy * duxy;
        Tm[8 * nr + q] = ux * duyx + uy * duyy;
    }
}

__global__ void k_coh_final(const double2* __restrict__ TK, double2* __restrict__ O, Grid g, double mux, double muy,
                            double cs2)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)g.nh * g.ny;
         q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % g.nh), j = (int)(q / g.nh);
        const double kx = f2k(i, g.nx, g.pi, g.lx, true), ky = f2k(j, g.ny, g.pi, g.ly, false);
        const size_t p = (size_t)j * g.ld + i;
        const bool cut = f2cut(kx, ky, g.nx, g.ny);
        double2 v[9];
        for (int f = 0; f < 9; ++f) v[f] = cut ? make_double2(0.0, 0.0) : TK[f * g.plane + p];
        const double nlxr = v[7].x + (mux * v[2].x + muy * v[3].x) + cs2 * v[0].x;
        const double nlxi = v[7].y + (mux * v[2].y + muy * v[3].y) + cs2 * v[0].y;
        const double nlyr = v[8].x + (mux * v[4].x + muy * v[5].x) + cs2 * v[1].x;
        const double nlyi = v[8].y + (mux * v[4].y + muy * v[5].y) + cs2 * v[1].y;
        O[p] = make_double2(-v[6].x, -v[6].y);
        O[g.plane + p] = make_double2(-nlxr, -nlxi);
        O[2 * g.plane + p] = make_double2(-nlyr, -nlyi);
    }
}

// ---------------------------------------------------------------- com
// R: the six fields after c2r (normalised here, in place); U: ux, uy, fx, fy.
__global__ void k_com_prim(double* __restrict__ R, double* __restrict__ U, Grid g, double norm, double a, double kf)
{
    const double dx = g.lx / (double)g.nx, dy = g.ly / (double)g.ny;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < g.nr; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % g.nx), j = (int)(q / g.nx);
        for (int f = 0; f < 6; ++f) R[f * g.nr + q] = R[f * g.nr + q] / norm;
        const double rho = R[q];
        U[q] = R[g.nr + q] / rho;
        U[g.nr + q] = R[2 * g.nr + q] / rho;
        U[2 * g.nr + q] = -(rho * a * sin(kf * ((double)j * dy)));
        U[3 * g.nr + q] = rho * a * sin(kf * ((double)i * dx));
    }
}

// i kx ux_k and i ky uy_k into AB (planes 0, 1).
__global__ void k_com_grad(const double2* __restrict__ UK, double2* __restrict__ AB, Grid g)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)g.nh * g.ny;
         q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % g.nh), j = (int)(q / g.nh);
        const double kx = f2k(i, g.nx, g.pi, g.lx, true), ky = f2k(j, g.ny, g.pi, g.ly, false);
        const size_t p = (size_t)j * g.ld + i;
        AB[p] = ik(UK[p], kx);
        AB[g.plane + p] = ik(UK[g.plane + p], ky);
    }
}

__global__ void k_com_flux(const double* __restrict__ R, const double* __restrict__ U, const double* __restrict__ D,
                           double* __restrict__ Fl, size_t nr, double norm, double cs2, double eta)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < nr; q += (size_t)gridDim.x * blockDim.x) {
        const double rho = R[q], rux = R[nr + q], ruy = R[2 * nr + q], E = R[3 * nr + q], bx = R[4 * nr + q],
                     by = R[5 * nr + q];
        const double vx = U[q], vy = U[nr + q];
        const double b2 = bx * bx + by * by;
        const double dudx = D[q] / norm, dvdy = D[nr + q] / norm;
        const double dbxdy = dudx / norm, dbydx = dvdy / norm;
        const double curl = dbydx - dbxdy;
        const double P = cs2 * rho;
        Fl[q] = rux * vx + P + b2 / 2.0 - bx * bx;
        Fl[nr + q] = rux * vy - bx * by;
        Fl[2 * nr + q] = rux * vy - bx * by;
        Fl[3 * nr + q] = ruy * vy + P + b2 / 2.0 - by * by;
        const double ex = (E + P + b2 / 2.0) * vx - vx * bx * bx - vy * bx * by;
        Fl[4 * nr + q] = ex - eta * by * curl;
        const double ey = (E + P + b2 / 2.0) * vy - vx * bx * by - vy * by * by;
        Fl[5 * nr + q] = ey + eta * bx * curl;
        const double dv = dudx + dvdy;
        Fl[6 * nr + q] = dv * dv;
        Fl[7 * nr + q] = vy * bx - by * vx;
        Fl[8 * nr + q] = vx * by - bx * vy;
    }
}

// S: state spectra (6); FK: flux spectra (9); UK: ux, uy, fx, fy spectra.
__global__ void k_com_final(const double2* __restrict__ S, const double2* __restrict__ FK,
                            const double2* __restrict__ UK, double2* __restrict__ O, Grid g, double mu, double eta)
{
    const size_t pl = g.plane;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)g.nh * g.ny;
         q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % g.nh), j = (int)(q / g.nh);
        const double kx = f2k(i, g.nx, g.pi, g.lx, true), ky = f2k(j, g.ny, g.pi, g.ly, false);
        const size_t p = (size_t)j * g.ld + i;
        double2 t[19];
        t[0] = ik(S[pl + p], kx);
        t[1] = ik(S[2 * pl + p], ky);
        t[2] = ik(FK[p], kx);
        t[3] = ik(FK[pl + p], ky);
        t[4] = ik(FK[2 * pl + p], kx);
        t[5] = ik(FK[3 * pl + p], ky);
        t[6] = k2m(UK[p], kx);
        t[7] = k2m(UK[p], ky);
        t[8] = k2m(UK[pl + p], kx);
        t[9] = k2m(UK[pl + p], ky);
        t[10] = ik(FK[4 * pl + p], kx);
        t[11] = ik(FK[5 * pl + p], ky);
        t[12] = FK[6 * pl + p];
        t[13] = ik(FK[7 * pl + p], ky);
        t[14] = ik(FK[8 * pl + p], kx);
        t[15] = k2m(S[4 * pl + p], kx);
        t[16] = k2m(S[4 * pl + p], ky);
        t[17] = k2m(S[5 * pl + p], kx);
        t[18] = k2m(S[5 * pl + p], ky);
        if (f2cut(kx, ky, g.nx, g.ny))
            for (int s = 0; s < 19; ++s) t[s] = make_double2(0.0, 0.0);
        const double2 fx = UK[2 * pl + p], fy = UK[3 * pl + p];
        double o[6][2];
        for (int c = 0; c < 2; ++c) {
            auto v = [&](int s) { return c ? t[s].y : t[s].x; };
            o[0][c] = -(v(0) + v(1));
            const double d1 = -(v(2) + v(3));
            o[1][c] = d1 - mu * (v(6) + v(7)) + (c ? fx.y : fx.x);
            const double d2 = -(v(4) + v(5));
            o[2][c] = d2 - mu * (v(8) + v(9)) + (c ? fy.y : fy.x);
            const double d3 = -(v(10) + v(11));
            o[3][c] = d3 + mu * v(12);
            o[4][c] = -v(13) - eta * (v(15) + v(16));
            o[5][c] = -v(14) - eta * (v(17) + v(18));
        }
        for (int f = 0; f < 6; ++f) O[f * pl + p] = make_double2(o[f][0], o[f][1]);
    }
}

// c2r / r2c of nf consecutive planes.
int c2r_planes(double2* s, double* r, const Grid& g, int nf, hipStream_t st)
{
    int rc;
    if ((rc = col_fft(s, g.ny, g.ld, g.ld, (long)g.plane, nf, +1, st))) return rc;
    return row_c2r(s, r, g.nx, g.ld, nf * g.ny, 1.0, st);
}

int r2c_planes(const double* r, double2* s, const Grid& g, int nf, hipStream_t st)
{
    int rc;
    if ((rc = row_r2c(r, s, g.nx, g.ld, nf * g.ny, st))) return rc;
    return col_fft(s, g.ny, g.ld, g.ld, (long)g.plane, nf, -1, st);
}

int upload(const Grid& g, double2* d, const double* const h[], int nf)
{
    for (int f = 0; f < nf; ++f)
        TARA_HIP(hipMemcpy2D(d + f * g.plane, g.ld * sizeof(double2), h[f], g.nh * sizeof(double2),
                             g.nh * sizeof(double2), g.ny, hipMemcpyHostToDevice));
    return TARA_OK;
}

int download(const Grid& g, double* const h[], const double2* d, int nf)
{
    for (int f = 0; f < nf; ++f)
        TARA_HIP(hipMemcpy2D(h[f], g.nh * sizeof(double2), d + f * g.plane, g.ld * sizeof(double2),
                             g.nh * sizeof(double2), g.ny, hipMemcpyDeviceToHost));
    return TARA_OK;
}

int make_grid(Grid& g, int nx, int ny, double pi, double lx, double ly)
{
    TARA_CHECK(nx >= 2 && ny >= 1 && (double)nx * ny <= (double)(1L << 28), TARA_ERR_UNSUPPORTED,
               "2D compressible: grid %dx%d unsupported", nx, ny);
    TARA_CHECK(lx > 0 && ly > 0 && pi > 0, TARA_ERR_ARG, "2D compressible: pi and box lengths must be positive");
    g.nx = nx;
    g.ny = ny;
    g.nh = nx / 2 + 1;
    g.ld = spec_ld(nx);
    g.plane = (size_t)g.ld * ny;
    g.nr = (size_t)nx * ny;
    g.pi = pi;
    g.lx = lx;
    g.ly = ly;
    return TARA_OK;
}

}  // namespace

bool f2_fast_grid(int nx, int ny)
{
    return is_pow2(nx) && is_pow2(ny) && nx >= 16 && ny >= 16 && nx <= 256 && ny <= 256;
}

int gen_derive2coh_host(int nx, int ny, const tara_coh_params* prm, const double* const state[3],
                        double* const rhs[3])
{
    Grid g;
    int rc = make_grid(g, nx, ny, prm->pi, prm->lx, prm->ly);
    if (rc) return rc;
    DevBuf sb, spb, rb, tb, tkb, ob;
    if ((rc = sb.alloc(3 * g.plane * sizeof(double2))) || (rc = spb.alloc(13 * g.plane * sizeof(double2))) ||
        (rc = rb.alloc(13 * g.nr * sizeof(double))) || (rc = tb.alloc(9 * g.nr * sizeof(double))) ||
        (rc = tkb.alloc(9 * g.plane * sizeof(double2))) || (rc = ob.alloc(3 * g.plane * sizeof(double2))))
        return rc;
    hipStream_t st = nullptr;
    TARA_HIP(hipMemset(sb.p, 0, sb.bytes));
    TARA_HIP(hipMemset(spb.p, 0, spb.bytes));
    if ((rc = upload(g, sb.as<double2>(), state, 3))) return rc;
    const size_t nq = (size_t)g.nh * ny;
    hipLaunchKernelGGL(k_coh_spec, dim3(blocks_for(nq)), dim3(256), 0, st, sb.as<double2>(), spb.as<double2>(), g);
    TARA_HIP(hipGetLastError());
    if ((rc = c2r_planes(spb.as<double2>(), rb.as<double>(), g, 13, st))) return rc;
    hipLaunchKernelGGL(k_coh_terms, dim3(blocks_for(g.nr)), dim3(256), 0, st, rb.as<double>(), tb.as<double>(), g.nr,
                       (double)nx * (double)ny);
    TARA_HIP(hipGetLastError());
    if ((rc = r2c_planes(tb.as<double>(), tkb.as<double2>(), g, 9, st))) return rc;
    hipLaunchKernelGGL(k_coh_final, dim3(blocks_for(nq)), dim3(256), 0, st, tkb.as<double2>(), ob.as<double2>(), g,
                       prm->mux, prm->muy, prm->cs * prm->cs);
    TARA_HIP(hipGetLastError());
    return download(g, rhs, ob.as<double2>(), 3);
}

int gen_derive2com_host(int nx, int ny, const tara_com_params* prm, const double* const state[6],
                        double* const rhs[6])
{
    Grid g;
    int rc = make_grid(g, nx, ny, prm->pi, prm->lx, prm->ly);
    if (rc) return rc;
    DevBuf sb, cb, rb, ub, ukb, abb, db, fb, fkb, ob;
    if ((rc = sb.alloc(6 * g.plane * sizeof(double2))) || (rc = cb.alloc(6 * g.plane * sizeof(double2))) ||
        (rc = rb.alloc(6 * g.nr * sizeof(double))) || (rc = ub.alloc(4 * g.nr * sizeof(double))) ||
        (rc = ukb.alloc(4 * g.plane * sizeof(double2))) || (rc = abb.alloc(2 * g.plane * sizeof(double2))) ||
        (rc = db.alloc(2 * g.nr * sizeof(double))) || (rc = fb.alloc(9 * g.nr * sizeof(double))) ||
        (rc = fkb.alloc(9 * g.plane * sizeof(double2))) || (rc = ob.alloc(6 * g.plane * sizeof(double2))))
        return rc;
    hipStream_t st = nullptr;
    TARA_HIP(hipMemset(sb.p, 0, sb.bytes));
    TARA_HIP(hipMemset(abb.p, 0, abb.bytes));
    if ((rc = upload(g, sb.as<double2>(), state, 6))) return rc;
    TARA_HIP(hipMemcpy(cb.p, sb.p, sb.bytes, hipMemcpyDeviceToDevice));
    const double norm = (double)nx * (double)ny;
    const size_t nq = (size_t)g.nh * ny;
    if ((rc = c2r_planes(cb.as<double2>(), rb.as<double>(), g, 6, st))) return rc;
    hipLaunchKernelGGL(k_com_prim, dim3(blocks_for(g.nr)), dim3(256), 0, st, rb.as<double>(), ub.as<double>(), g, norm,
                       prm->a, prm->kf);
    TARA_HIP(hipGetLastError());
    if ((rc = r2c_planes(ub.as<double>(), ukb.as<double2>(), g, 4, st))) return rc;
    hipLaunchKernelGGL(k_com_grad, dim3(blocks_for(nq)), dim3(256), 0, st, ukb.as<double2>(), abb.as<double2>(), g);
    TARA_HIP(hipGetLastError());
    if ((rc = c2r_planes(abb.as<double2>(), db.as<double>(), g, 2, st))) return rc;
    hipLaunchKernelGGL(k_com_flux, dim3(blocks_for(g.nr)), dim3(256), 0, st, rb.as<double>(), ub.as<double>(),
                       db.as<double>(), fb.as<double>(), g.nr, norm, prm->cs * prm->cs, prm->eta);
    TARA_HIP(hipGetLastError());
    if ((rc = r2c_planes(fb.as<double>(), fkb.as<double2>(), g, 9, st))) return rc;
    hipLaunchKernelGGL(k_com_final, dim3(blocks_for(nq)), dim3(256), 0, st, sb.as<double2>(), fkb.as<double2>(),
                       ukb.as<double2>(), ob.as<double2>(), g, prm->mu, prm->eta);
    TARA_HIP(hipGetLastError());
    return download(g, rhs, ob.as<double2>(), 6);
}

}  // namespace tara
