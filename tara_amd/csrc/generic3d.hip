// derive3com on grids the fused 3D passes do not take (any Nx >= 2, Ny, Nz
// >= 1), built from the any-length transforms: c2r of the seven live state
// fields and the 1/N normalisation, the pointwise velocity, pressure,
// momentum-flux and induction-flux fields, fourteen r2c, and the spherically
// masked divergence / diffusion assembly (src/derive3com.f95:94-140,
// :176-186, :395-436, :574-843, :1018-1037), each a device kernel over
// half spectra of pitch spec_ld(nx) or real fields of pitch nx.  Slower than
// the fused path; the arithmetic is the same term by term.
#include "common.h"
#include "fft_api.h"
#include "generic3d.h"

namespace tara {

namespace {

constexpr int kLive[7] = {0, 1, 2, 3, 5, 6, 7};
enum { UX, UY, UZ, MX1, MX2, MX3, MY2, MY3, MZ1, MZ2, MZ3, MG1, MG2, MG3, NF };

// k = 2 pi (i-1)/L with the true pi, the upper half of y and z at (i-1)-N
// (src/derive3com.f95:245-247, 260-262), evaluated as ((2 pi) m) / L.
__device__ __forceinline__ double com_k(int idx0, int n, bool halved, double L)
{
    const double pi = 3.14159265358979323846;
    const double m = (halved || idx0 < n / 2) ? (double)idx0 : (double)(idx0 - n);
    return 2.0 * pi * m / L;
}

unsigned blocks_for(size_t n)
{
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 16384 ? 16384 : (b ? b : 1));
}

__global__ void k3_div(double* __restrict__ r, size_t n, double norm)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        r[q] = r[q] / norm;
}

// R: rho, rho u (3), B (3) in real space; F: the fourteen flux fields.
__global__ void k3_flux(const double* __restrict__ R, double* __restrict__ F, size_t nr, double cs2)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < nr; q += (size_t)gridDim.x * blockDim.x) {
        const double rho = R[q], rux = R[nr + q], ruy = R[2 * nr + q], ruz = R[3 * nr + q];
        const double bx = R[4 * nr + q], by = R[5 * nr + q], bz = R[6 * nr + q];
        const double ux = rux / rho, uy = ruy / rho, uz = ruz / rho;
        const double b2 = bx * bx + by * by + bz * bz;
        const double P = cs2 * rho;
        F[UX * nr + q] = ux;
        F[UY * nr + q] = uy;
        F[UZ * nr + q] = uz;
        F[MX1 * nr + q] = rux * ux + P + b2 / 2.0 - bx * bx;
        F[MX2 * nr + q] = rux * uy - bx * by;
        F[MX3 * nr + q] = rux * uz - bx * bz;
        F[MY2 * nr + q] = ruy * uy + P + b2 / 2.0 - by * by;
        F[MY3 * nr + q] = ruy * uz - by * bz;
        F[MZ1 * nr + q] = ruz * ux - bz * bx;
        F[MZ2 * nr + q] = ruz * uy - bz * by;
        F[MZ3 * nr + q] = ruz * uz + P + b2 / 2.0 - bz * bz;
        F[MG1 * nr + q] = ux * by - uy * bx;
        F[MG2 * nr + q] = ux * bz - uz * bx;
        F[MG3 * nr + q] = uy * bz - uz * by;
    }
}

// S: the seven live state spectra (rho, rho u, B); K: the fourteen flux
// spectra; O: the seven live right-hand sides.  Planes of fsz entries.
__global__ void k3_assemble(const double2* __restrict__ S, const double2* __restrict__ K, double2* __restrict__ O,
                            int nx, int ny, int nz, int ld, size_t fsz, double lx, double ly, double lz, double mu,
                            double eta)
{
    const int nh = nx / 2 + 1;
    const double cut = ((double)(nx + ny + nz) / 3.0) / 3.0;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < fsz; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % ld);
        const size_t jk = q / ld;
        const int j = (int)(jk % ny), k = (int)(jk / ny);
        if (i >= nh) {
            for (int f = 0; f < 7; ++f) O[f * fsz + q] = make_double2(0.0, 0.0);
            continue;
        }
        const double kx = com_k(i, nx, true, lx), ky = com_k(j, ny, false, ly), kz = com_k(k, nz, false, lz);
        const bool m = sqrt(kx * kx + ky * ky + kz * kz) >= cut;
        auto ikr = [&](double kv, const double2* a) { return m ? 0.0 : -(kv * a[q].y); };
        auto iki = [&](double kv, const double2* a) { return m ? 0.0 : (kv * a[q].x); };
        auto k2r = [&](double kv, const double2* a) { return m ? 0.0 : (kv * kv) * a[q].x; };
        auto k2i = [&](double kv, const double2* a) { return m ? 0.0 : (kv * kv) * a[q].y; };
        const double2 *rux = S + fsz, *ruy = S + 2 * fsz, *ruz = S + 3 * fsz;
        // density: -(i kx rho ux + i ky rho uy + i kz rho uz) on the state spectra
        const double sr = (ikr(kx, rux) + ikr(ky, ruy)) + ikr(kz, ruz);
        const double si = (iki(kx, rux) + iki(ky, ruy)) + iki(kz, ruz);
        O[q] = make_double2(-sr, -si);
        // momentum: -(i k . flux row) - mu k^2 u
        const double2* mom[3][3] = {{K + MX1 * fsz, K + MX2 * fsz, K + MX3 * fsz},
                                    {K + MX2 * fsz, K + MY2 * fsz, K + MY3 * fsz},
                                    {K + MZ1 * fsz, K + MZ2 * fsz, K + MZ3 * fsz}};
        const double2* uk[3] = {K + UX * fsz, K + UY * fsz, K + UZ * fsz};
        for (int c = 0; c < 3; ++c) {
            const double ar = (ikr(kx, mom[c][0]) + ikr(ky, mom[c][1])) + ikr(kz, mom[c][2]);
            const double ai = (iki(kx, mom[c][0]) + iki(ky, mom[c][1])) + iki(kz, mom[c][2]);
            const double dr = (k2r(kx, uk[c]) + k2r(ky, uk[c])) + k2r(kz, uk[c]);
            const double di = (k2i(kx, uk[c]) + k2i(ky, uk[c])) + k2i(kz, uk[c]);
            O[(1 + c) * fsz + q] = make_double2(-ar - dr * mu, -ai - di * mu);
        }
        // induction
        double d2r[3], d2i[3];
        for (int c = 0; c < 3; ++c) {
            const double2* b = S + (4 + c) * fsz;
            d2r[c] = (k2r(kx, b) + k2r(ky, b)) + k2r(kz, b);
            d2i[c] = (k2i(kx, b) + k2i(ky, b)) + k2i(kz, b);
        }
        const double2 *g1 = K + MG1 * fsz, *g2 = K + MG2 * fsz, *g3 = K + MG3 * fsz;
        const double xr = ikr(ky, g1) + ikr(kz, g2), xi = iki(ky, g1) + iki(kz, g2);
        O[4 * fsz + q] = make_double2(xr - d2r[0] * eta, xi - d2i[0] * eta);
        const double yr = ikr(kx, g1) - ikr(kz, g3), yi = iki(kx, g1) - iki(kz, g3);
        O[5 * fsz + q] = make_double2(-yr - d2r[1] * eta, -yi - d2i[1] * eta);
        const double zr = ikr(kx, g2) + ikr(ky, g3), zi = iki(kx, g2) + iki(ky, g3);
        O[6 * fsz + q] = make_double2(-zr - d2r[2] * eta, -zi - d2i[2] * eta);
    }
}

int c2r3(double2* s, double* r, int nx, int ny, int nz, int ld, hipStream_t st)
{
    int rc;
    if ((rc = col_fft(s, nz, ld, (long)ld * ny, ld, ny, +1, st))) return rc;
    if ((rc = col_fft(s, ny, ld, ld, (long)ld * ny, nz, +1, st))) return rc;
    return row_c2r(s, r, nx, ld, ny * nz, 1.0, st);
}

int r2c3(const double* r, double2* s, int nx, int ny, int nz, int ld, hipStream_t st)
{
    int rc;
    if ((rc = row_r2c(r, s, nx, ld, ny * nz, st))) return rc;
    if ((rc = col_fft(s, ny, ld, ld, (long)ld * ny, nz, -1, st))) return rc;
    return col_fft(s, nz, ld, (long)ld * ny, ld, ny, -1, st);
}

}  // namespace

bool mhd_fast_grid(int nx, int ny, int nz)
{
    return is_pow2(nx) && is_pow2(ny) && is_pow2(nz) && nx >= 16 && ny >= 16 && nz >= 16 && nx <= 256 && ny <= 256 &&
           nz <= 256;
}

int gen_derive3com_host(int nx, int ny, int nz, const tara_mhd_params* prm, const double* const state[8],
                        double* const rhs[8])
{
    TARA_CHECK(nx >= 2 && ny >= 1 && nz >= 1 && (double)nx * ny * nz <= (double)(1L << 28), TARA_ERR_UNSUPPORTED,
               "derive3com: grid %dx%dx%d unsupported", nx, ny, nz);
    TARA_CHECK(prm->lx > 0 && prm->ly > 0 && prm->lz > 0, TARA_ERR_ARG, "3D MHD: box lengths must be positive");
    const int nh = nx / 2 + 1, ld = spec_ld(nx);
    const size_t fsz = (size_t)ld * ny * nz, nr = (size_t)nx * ny * nz, rows = (size_t)ny * nz;
    DevBuf sb, cb, kb, ob, rb, fb;
    int rc;
    if ((rc = sb.alloc(7 * fsz * sizeof(double2))) || (rc = cb.alloc(7 * fsz * sizeof(double2))) ||
        (rc = kb.alloc(NF * fsz * sizeof(double2))) || (rc = ob.alloc(7 * fsz * sizeof(double2))) ||
        (rc = rb.alloc(7 * nr * sizeof(double))) || (rc = fb.alloc(NF * nr * sizeof(double))))
        return rc;
    double2 *S = sb.as<double2>(), *Cs = cb.as<double2>(), *K = kb.as<double2>(), *O = ob.as<double2>();
    double *R = rb.as<double>(), *F = fb.as<double>();
    hipStream_t st = nullptr;
    TARA_HIP(hipMemset(S, 0, sb.bytes));
    for (int f = 0; f < 7; ++f)
        TARA_HIP(hipMemcpy2D(S + f * fsz, ld * sizeof(double2), state[kLive[f]], nh * sizeof(double2),
                             nh * sizeof(double2), rows, hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(Cs, S, sb.bytes, hipMemcpyDeviceToDevice));
    for (int f = 0; f < 7; ++f)
        if ((rc = c2r3(Cs + f * fsz, R + f * nr, nx, ny, nz, ld, st))) return rc;
    hipLaunchKernelGGL(k3_div, dim3(blocks_for(7 * nr)), dim3(256), 0, st, R, 7 * nr, (double)nx * ny * nz);
    TARA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k3_flux, dim3(blocks_for(nr)), dim3(256), 0, st, R, F, nr, prm->cs * prm->cs);
    TARA_HIP(hipGetLastError());
    for (int f = 0; f < NF; ++f)
        if ((rc = r2c3(F + f * nr, K + f * fsz, nx, ny, nz, ld, st))) return rc;
    hipLaunchKernelGGL(k3_assemble, dim3(blocks_for(fsz)), dim3(256), 0, st, S, K, O, nx, ny, nz, ld, fsz, prm->lx,
                       prm->ly, prm->lz, prm->mu, prm->eta);
    TARA_HIP(hipGetLastError());
    for (int f = 0; f < 7; ++f)
        TARA_HIP(hipMemcpy2D(rhs[kLive[f]], nh * sizeof(double2), O + f * fsz, ld * sizeof(double2),
                             nh * sizeof(double2), rows, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < 2 * (size_t)nh * rows; ++q) rhs[4][q] = 0.0;
    return TARA_OK;
}

}  // namespace tara
