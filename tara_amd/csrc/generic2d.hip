// derive2csh / derive2css on grids the fused passes do not take (any Nx >= 2,
// Ny >= 1, as FFTW and the reference accept), built from the any-length
// transforms: the routine's own sequence of spectral products, four c2r,
// the two real products, two r2c and the dealiased, viscous combination
// (src/derive2csh.f95:29-145; src/derive2css.f95:20-130), each a device
// kernel over half spectra of pitch spec_ld(nx) (kx fastest) or real planes
// of pitch nx.  Slower than the fused path (about ten launches and one
// workspace allocation per call) and used only where that path has no
// kernel; the arithmetic is the same, term by term.
#include "common.h"
#include "csh2d.h"
#include "fft_api.h"
#include "generic2d.h"

namespace tara {

namespace {

// derive2csh's wavenumbers: its pi is the default-real literal 3.1428
// (src/derive2csh.f95:16), Lx = Ly = 2 pi (:20-21), kx = 2 pi (i-1)/Lx (:30),
// ky = 2 pi (j-1)/Ly for j <= Ny/2, 2 pi ((j-1)-Ny)/Ly above (:32, :48).
__device__ __forceinline__ double csh_k(int idx0, int n, bool halved)
{
    const double pi = (double)3.1428f;
    const double L = 2.0 * pi;
    const double m = (halved || idx0 < n / 2) ? (double)idx0 : ((double)idx0 - (double)n);
    return 2.0 * pi * m / L;
}

unsigned blocks_for(size_t n)
{
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 16384 ? 16384 : (b ? b : 1));
}

// psi (psi(0,0) = 0), ukx = +i ky psi, uky = -i kx psi and i kx, i ky of
// omega (derive2csh) or of omega_d = k^2 psi with psi = omega/(k^2 + lambda)
// (derive2css), into planes 0..3 of S; omega_d into plane 4 (css).
__global__ void k_gen_spec(const double2* __restrict__ w, int nx, int ny, int ld, int screened, double lambda,
                           double2* __restrict__ S)
{
    const int nh = nx / 2 + 1;
    const size_t plane = (size_t)ld * ny;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)nh * ny;
         q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % nh), j = (int)(q / nh);
        const double kx = csh_k(i, nx, true), ky = csh_k(j, ny, false);
        const size_t p = (size_t)j * ld + i;
        const double wr = w[p].x, wi = w[p].y;
        const bool origin = i == 0 && j == 0;
        double2 ukx, uky, ikx, iky;
        if (!screened) {
            double pr = 0.0, pim = 0.0;
            if (!origin) {
                const double k2 = kx * kx + ky * ky;
                pr = wr / k2;
                pim = wi / k2;
            }
            ukx = origin ? make_double2(0.0, 0.0) : make_double2(-(ky * pim), ky * pr);
            uky = origin ? make_double2(0.0, 0.0) : make_double2(kx * pim, -(kx * pr));
            ikx = make_double2(-(kx * wi), kx * wr);
            iky = make_double2(-(ky * wi), ky * wr);
        } else {
            const double k2 = kx * kx + ky * ky;
            double pr = 0.0, pim = 0.0;
            if (!origin) {
                const double den = k2 + lambda;
                pr = wr / den;
                pim = wi / den;
            }
            ukx = make_double2(-(ky * pim), ky * pr);
            uky = make_double2(kx * pim, -(kx * pr));
            const double dr = k2 * pr, di = k2 * pim;
            S[4 * plane + p] = make_double2(dr, di);
            ikx = make_double2(-(kx * di), kx * dr);
            iky = make_double2(-(ky * di), ky * dr);
        }
        S[p] = ukx;
        S[plane + p] = uky;
        S[2 * plane + p] = ikx;
        S[3 * plane + p] = iky;
    }
}

// Normalise the four c2r outputs and form ux * omega_x, uy * omega_y into
// planes 2 and 3 (src/derive2csh.f95:89-98).
__global__ void k_gen_prod(double* __restrict__ R, size_t nr, double norm)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < nr; q += (size_t)gridDim.x * blockDim.x) {
        const double a = R[q] / norm, b = R[nr + q] / norm, c = R[2 * nr + q] / norm, d = R[3 * nr + q] / norm;
        R[2 * nr + q] = a * c;
        R[3 * nr + q] = b * d;
    }
}

// Circular dealias at (Nx+Ny)/2/3, viscous term and sign (src/derive2csh.f95:108-145;
// derive2css.f95:110-130 with alpha on the products and omega_d in the viscous term).
__global__ void k_gen_final(const double2* __restrict__ N, const double2* __restrict__ w,
                            const double2* __restrict__ wd, int nx, int ny, int ld, double nu, double alpha,
                            int screened, double2* __restrict__ f)
{
    const int nh = nx / 2 + 1;
    const size_t plane = (size_t)ld * ny;
    const double cut = ((double)(nx + ny) / 2.0) / 3.0;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)ld * ny;
         q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % ld), j = (int)(q / ld);
        if (i >= nh) {
            f[q] = make_double2(0.0, 0.0);
            continue;
        }
        const double kx = csh_k(i, nx, true), ky = csh_k(j, ny, false);
        double2 a = N[q], b = N[plane + q];
        if (sqrt(kx * kx + ky * ky) >= cut) a = b = make_double2(0.0, 0.0);
        const double vr = nu * (kx * kx + ky * ky);
        if (!screened) {
            const double nr_ = (a.x + b.x) + vr * w[q].x;
            const double ni_ = (a.y + b.y) + vr * w[q].y;
            f[q] = make_double2(-nr_, -ni_);
        } else {
            f[q] = make_double2(-(alpha * (a.x + b.x) + vr * wd[q].x), -(alpha * (a.y + b.y) + vr * wd[q].y));
        }
    }
}

// out = w + c k over n complex entries.
__global__ void k_gen_axpy(const double2* __restrict__ w, const double2* __restrict__ k, double c,
                           double2* __restrict__ out, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        out[q] = make_double2(w[q].x + c * k[q].x, w[q].y + c * k[q].y);
}

// w += dt/6 (((k1 + 2 k2) + 2 k3) + k4).
__global__ void k_gen_rk4(double2* __restrict__ w, const double2* __restrict__ k1, const double2* __restrict__ k2,
                          const double2* __restrict__ k3, const double2* __restrict__ k4, double dt, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const double h = dt / 6.0;
        w[q] = make_double2(w[q].x + h * (((k1[q].x + 2.0 * k2[q].x) + 2.0 * k3[q].x) + k4[q].x),
                            w[q].y + h * (((k1[q].y + 2.0 * k2[q].y) + 2.0 * k3[q].y) + k4[q].y));
    }
}

}  // namespace

bool csh_fast_grid(int nx, int ny)
{
    return is_pow2(nx) && is_pow2(ny) && nx >= 32 && ny >= 16 && nx <= 4096 && ny <= 4096;
}

int gen_csh_rhs(int nx, int ny, const CshPhys& ph, const double2* w, double2* f, hipStream_t st)
{
    TARA_CHECK(nx >= 2 && ny >= 1 && (double)nx * ny <= (double)(1L << 30), TARA_ERR_UNSUPPORTED,
               "derive2csh: grid %dx%d unsupported", nx, ny);
    const int ld = spec_ld(nx);
    const size_t plane = (size_t)ld * ny, nr = (size_t)nx * ny;
    DevBuf spec, real, nl;
    int rc;
    if ((rc = spec.alloc(5 * plane * sizeof(double2))) || (rc = real.alloc(4 * nr * sizeof(double))) ||
        (rc = nl.alloc(2 * plane * sizeof(double2))))
        return rc;
    double2* S = spec.as<double2>();
    double* R = real.as<double>();
    double2* N = nl.as<double2>();
    TARA_HIP(hipMemsetAsync(S, 0, spec.bytes, st));
    hipLaunchKernelGGL(k_gen_spec, dim3(blocks_for((size_t)(nx / 2 + 1) * ny)), dim3(256), 0, st, w, nx, ny, ld,
                       ph.screened, ph.lambda, S);
    TARA_HIP(hipGetLastError());
    if ((rc = col_fft(S, ny, ld, ld, (long)plane, 4, +1, st))) return rc;
    if ((rc = row_c2r(S, R, nx, ld, 4 * ny, 1.0, st))) return rc;
    hipLaunchKernelGGL(k_gen_prod, dim3(blocks_for(nr)), dim3(256), 0, st, R, nr, (double)nx * (double)ny);
    TARA_HIP(hipGetLastError());
    if ((rc = row_r2c(R + 2 * nr, N, nx, ld, 2 * ny, st))) return rc;
    if ((rc = col_fft(N, ny, ld, ld, (long)plane, 2, -1, st))) return rc;
    hipLaunchKernelGGL(k_gen_final, dim3(blocks_for(plane)), dim3(256), 0, st, N, w, S + 4 * plane, nx, ny, ld,
                       ph.nu, ph.alpha, ph.screened, f);
    TARA_HIP(hipGetLastError());
    TARA_HIP(hipStreamSynchronize(st));
    return TARA_OK;
}

int gen_csh_rk4(int nx, int ny, const CshPhys& ph, double dt, double2* w, hipStream_t st)
{
    const size_t n = (size_t)spec_ld(nx) * ny;
    DevBuf kb, sb;
    int rc;
    if ((rc = kb.alloc(4 * n * sizeof(double2))) || (rc = sb.alloc(n * sizeof(double2)))) return rc;
    double2* k = kb.as<double2>();
    double2* s = sb.as<double2>();
    if ((rc = gen_csh_rhs(nx, ny, ph, w, k, st))) return rc;
    hipLaunchKernelGGL(k_gen_axpy, dim3(blocks_for(n)), dim3(256), 0, st, w, k, 0.5 * dt, s, n);
    if ((rc = gen_csh_rhs(nx, ny, ph, s, k + n, st))) return rc;
    hipLaunchKernelGGL(k_gen_axpy, dim3(blocks_for(n)), dim3(256), 0, st, w, k + n, 0.5 * dt, s, n);
    if ((rc = gen_csh_rhs(nx, ny, ph, s, k + 2 * n, st))) return rc;
    hipLaunchKernelGGL(k_gen_axpy, dim3(blocks_for(n)), dim3(256), 0, st, w, k + 2 * n, dt, s, n);
    if ((rc = gen_csh_rhs(nx, ny, ph, s, k + 3 * n, st))) return rc;
    hipLaunchKernelGGL(k_gen_rk4, dim3(blocks_for(n)), dim3(256), 0, st, w, k, k + n, k + 2 * n, k + 3 * n, dt, n);
    TARA_HIP(hipGetLastError());
    TARA_HIP(hipStreamSynchronize(st));
    return TARA_OK;
}

}  // namespace tara
