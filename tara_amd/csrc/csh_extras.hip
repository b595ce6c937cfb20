// Work of the 2D time loop that the reference's driver does on the host
// every step, moved onto the device state of a tara_csh context:
//  - real-space omega, ux, uy (src/main2csh.f95:128-178: psi, u spectra, three
//    c2r, division by Nx*Ny), for snapshots and tracers at the caller's cadence;
//  - initial conditions: the reference's shear layer (src/main2csh.f95:84-110)
//    and a seeded random-phase field whose phases come from the counter-based
//    Philox-4x32-10 generator (Salmon, Moraes, Dror, Shaw, SC'11), so that any
//    mode's phase depends only on (seed, mode index), independent of the launch.
#include <cmath>
#include <vector>

#include "csh2d.h"
#include "fft_api.h"
#include "layout.h"
#include "philox.h"
#include "tara_hip.h"

namespace tara {

// ------------------------------------------------------------------ kernels
// Spectra of omega, ux = +i ky psi, uy = -i kx psi (psi = omega/k^2, psi(0,0) = 0)
// from the ky-major state into three padded kx-fastest planes (pitch ld).
__global__ void k_fields_spec(const double2* w, int nh, int ny, int ld, double2* s)
{
    const size_t plane = (size_t)ld * ny;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)nh * ny;
         q += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(q % ny), i = (int)(q / ny);
        const double2 z = w[q];
        const double kx = (double)i, ky = (double)(j < ny / 2 ? j : j - ny);
        const double k2 = kx * kx + ky * ky;
        const double2 ps = k2 != 0.0 ? make_double2(z.x / k2, z.y / k2) : make_double2(0.0, 0.0);
        const size_t o = (size_t)j * ld + i;
        s[o] = z;
        s[plane + o] = make_double2(-(ky * ps.y), ky * ps.x);
        s[2 * plane + o] = make_double2(kx * ps.y, -(kx * ps.x));
    }
}

// Shear layer of src/main2csh.f95:94-106 on the (Nx, Ny) grid, x = (i-1) dx.
__global__ void k_shear_layer(double* out, int nx, int ny, double dx, double dy, double pi, double w0, double m0,
                              double d)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)nx * ny;
         q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % nx), j = (int)(q / nx);
        const double x = (double)i * dx, y = (double)j * dy;
        const double c1 = cosh((y + 0.5 * pi) / d), c2 = cosh((y - 0.5 * pi) / d);
        const double c3 = cosh((y + 1.5 * pi) / d), c4 = cosh((y - 1.5 * pi) / d);
        double v = w0 / (c1 * c1) - w0 / (c2 * c2);
        v = v - w0 / (c3 * c3) + w0 / (c4 * c4);
        out[q] = v + 0.10 * cos(m0 * x);
    }
}

// Random-phase vorticity on the half plane, written ky-major: amplitude
// sqrt(k E(k)/pi) with E(k) = k^4 exp(-(k/k0)^2), times scale; phase 2 pi u
// with u from Philox(counter = (i + j*nh, 0, 0, 0), key = (seed, 0)).  Modes on
// the Nyquist row / column and k = 0 are zero; the kx = 0 column is Hermitian
// (ky < 0 entries are the conjugates of ky > 0 ones).
__global__ void k_random_phase(double2* w, int nh, int ny, unsigned seed, double k0, double scale)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)nh * ny;
         q += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(q % ny), i = (int)(q / ny);
        int jj = j;
        bool conj = false;
        if (i == 0 && j > ny / 2) {
            jj = ny - j;
            conj = true;
        }
        double2 v = make_double2(0.0, 0.0);
        if (!(i == nh - 1 || j == ny / 2 || (i == 0 && jj == 0))) {
            const double kx = (double)i, ky = (double)(jj < ny / 2 ? jj : jj - ny);
            const double k = sqrt(kx * kx + ky * ky);
            const double e = k * k * k * k * exp(-(k / k0) * (k / k0));
            const double amp = sqrt(k * e / 3.14159265358979323846) * scale;
            const U4 r = philox4x32_10(U4{(unsigned)(i + (size_t)jj * nh), 0u, 0u, 0u}, seed, 0u);
            const double ph = 2.0 * 3.14159265358979323846 * u53(r.x, r.y);
            v = make_double2(amp * cos(ph), amp * sin(ph));
            if (conj) v.y = -v.y;
        }
        w[q] = v;
    }
}

// ------------------------------------------------------------ tracers
// Velocity spectra of the driver (src/main2csh.f95:128-156): kx = 2 pi (i-1)/Lx,
// ky = 2 pi (j-1)/Ly or 2 pi ((j-1)-Ny)/Ly with the caller's box and true pi,
// psi = omega/(kx^2+ky^2) (psi(0,0) = 0), ux = +i ky psi, uy = -i kx psi, from
// the ky-major state into two padded kx-fastest planes (pitch ld).
__global__ void k_tracer_spec(const double2* w, int nh, int ny, int ld, double lx, double ly, double2* s)
{
    const double pi = 3.14159265358979323846;
    const size_t plane = (size_t)ld * ny;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)nh * ny;
         q += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(q % ny), i = (int)(q / ny);
        const double kx = 2.0 * pi * (double)i / lx;
        const double ky = 2.0 * pi * (double)(j < ny / 2 ? j : j - ny) / ly;
        const double2 z = w[q];
        const double k2 = kx * kx + ky * ky;
        const double2 ps = (i == 0 && j == 0) ? make_double2(0.0, 0.0) : make_double2(z.x / k2, z.y / k2);
        const size_t o = (size_t)j * ld + i;
        s[o] = make_double2(-(ky * ps.y), ky * ps.x);
        s[plane + o] = make_double2(kx * ps.y, -(kx * ps.x));
    }
}

// One step of the tracer particles (src/main2csh.f95:187-265): bilinear
// interpolation of ux, uy (unnormalised c2r output, divided by Nx*Ny as the
// driver does) at each particle with the driver's expression, then
// x <- x + u dt and the wrap into (0, L].  The corners are read as the
// driver addresses its column-major (Nx, Ny) arrays: with periodic == 0,
// ux(m+1, n) of the last column is the element after ux(Nx, n) in memory,
// ux(1, n+1), as there, and a read past the array's end (the last row's
// upper corners, the last column's in the row below it) reads 0: undefined
// in the language, it is what the driver reads where its arrays are
// allocated with mmap, as glibc does for arrays of 128 KiB and more (the
// page holding the end is zero-filled beyond it).  With periodic != 0 every
// corner is the periodic image.  The last cell uses x(Nx+1) = Lx,
// y(Ny+1) = Ly as the driver's edge branches do.  A position that is not
// finite or lies far outside the box sets *bad and is left alone.
__global__ void k_tracers(const double* ux, const double* uy, int nx, int ny, double lx, double ly, double dt,
                          int periodic, int np, double* xp, double* yp, int* bad)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const double dx = lx / (double)nx, dy = ly / (double)ny, h = 1.0 / (dx * dy), nn = (double)nx * (double)ny;
    double x = xp[p], y = yp[p];
    if (!(fabs(x) <= 4.0 * lx && fabs(y) <= 4.0 * ly)) {
        *bad = 1;
        return;
    }
    int m = (int)(x / dx), n = (int)(y / dy);  // 0-based cells: aint(x/dx) + 1 - 1
    m = m < 0 ? 0 : (m >= nx ? nx - 1 : m);
    n = n < 0 ? 0 : (n >= ny ? ny - 1 : n);
    const double xm = (double)m * dx, ym = (double)n * dy;
    const double xm1 = m + 1 < nx ? (double)(m + 1) * dx : lx, yn1 = n + 1 < ny ? (double)(n + 1) * dy : ly;
    const size_t total = (size_t)nx * ny, a = (size_t)n * nx + m;
    size_t b, c, d;
    if (periodic) {
        const size_t m1 = (size_t)((m + 1) % nx), n1 = (size_t)((n + 1) % ny);
        b = (size_t)n * nx + m1;
        c = n1 * nx + m;
        d = n1 * nx + m1;
    } else {
        b = a + 1;
        c = a + nx;
        d = a + nx + 1;
    }
    const double uxa = ux[a], uya = uy[a];
    const double uxb = b < total ? ux[b] : 0.0, uyb = b < total ? uy[b] : 0.0;
    const double uxc = c < total ? ux[c] : 0.0, uyc = c < total ? uy[c] : 0.0;
    const double uxd = d < total ? ux[d] : 0.0, uyd = d < total ? uy[d] : 0.0;
    const double vx = (uxa / nn) * (xm1 - x) * (yn1 - y) * h + (uxb / nn) * (x - xm) * (yn1 - y) * h +
                      (uxc / nn) * (xm1 - x) * (y - ym) * h + (uxd / nn) * (x - xm) * (y - ym) * h;
    const double vy = (uya / nn) * (xm1 - x) * (yn1 - y) * h + (uyb / nn) * (x - xm) * (yn1 - y) * h +
                      (uyc / nn) * (xm1 - x) * (y - ym) * h + (uyd / nn) * (x - xm) * (y - ym) * h;
    x = x + vx * dt;
    y = y + vy * dt;
    x = x - trunc(x / lx) * lx;  // int(x/Lx) without the integer overflow
    if (x <= 0.0) x = x + lx;
    y = y - trunc(y / ly) * ly;
    if (y <= 0.0) y = y + ly;
    xp[p] = x;
    yp[p] = y;
}

static unsigned grid_for(size_t n)
{
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b ? b : 1));
}

// Kinetic energy 1/2 <|u|^2> of the random-phase field before scaling: Parseval
// with half-plane weights (2 for kx > 0), |omega_k| = amp, |u_k| = amp / k.
static double random_phase_energy(int nx, int ny, double k0)
{
    const int nh = nx / 2 + 1;
    double s = 0.0;
    for (int i = 0; i < nh - 1; ++i)
        for (int j = 0; j < ny; ++j) {
            if (j == ny / 2 || (i == 0 && j == 0)) continue;
            const double kx = i, ky = j < ny / 2 ? j : j - ny;
            const double k2 = kx * kx + ky * ky, k = std::sqrt(k2);
            const double e = k2 * k2 * std::exp(-(k / k0) * (k / k0));
            s += (i == 0 ? 1.0 : 2.0) * (k * e / 3.14159265358979323846) / k2;
        }
    return 0.5 * s / ((double)nx * ny * (double)nx * ny);
}

}  // namespace tara

// ------------------------------------------------------------------ C ABI
using namespace tara;

extern "C" int tara_csh_fields(tara_csh* c, double* omega, double* ux, double* uy)
{
    double2* w = nullptr;
    const CshPlan* p = nullptr;
    hipStream_t st = nullptr;
    int rc = csh_ctx_state(c, &w, &p, &st);
    if (rc) return rc;
    const int nx = p->nx, ny = p->ny, ld = spec_ld(nx);
    TARA_CHECK(nx >= 16 && ny >= 8, TARA_ERR_UNSUPPORTED, "fields: grid too small");
    DevBuf spec, real;
    if ((rc = spec.alloc(3 * (size_t)ld * ny * sizeof(double2))) || (rc = real.alloc(3 * (size_t)nx * ny * 8)))
        return rc;
    TARA_HIP(hipMemsetAsync(spec.p, 0, spec.bytes, st));
    hipLaunchKernelGGL(k_fields_spec, dim3(grid_for((size_t)p->nh * ny)), dim3(256), 0, st, w, p->nh, ny, ld,
                       spec.as<double2>());
    TARA_HIP(hipGetLastError());
    if ((rc = col_fft(spec.as<double2>(), ny, ld, ld, (long)ld * ny, 3, +1, st))) return rc;
    if ((rc = row_c2r(spec.as<double2>(), real.as<double>(), nx, ld, 3 * ny, 1.0 / ((double)nx * ny), st)))
        return rc;
    double* outs[3] = {omega, ux, uy};
    const size_t nr = (size_t)nx * ny;
    for (int f = 0; f < 3; ++f)
        if (outs[f]) TARA_HIP(hipMemcpyAsync(outs[f], real.as<double>() + f * nr, nr * 8, hipMemcpyDeviceToHost, st));
    TARA_HIP(hipStreamSynchronize(st));
    return TARA_OK;
}

extern "C" int tara_csh_advance_tracers(tara_csh* c, int nsteps, double dt, double lx, double ly, int periodic,
                                        int np, double* xp, double* yp)
{
    TARA_CHECK(c && nsteps >= 0 && np >= 0 && (np == 0 || (xp && yp)) && lx > 0 && ly > 0 && std::isfinite(dt),
               TARA_ERR_ARG, "bad argument");
    for (int i = 0; i < np; ++i)
        TARA_CHECK(xp[i] >= 0.0 && xp[i] <= lx && yp[i] >= 0.0 && yp[i] <= ly, TARA_ERR_ARG,
                   "tracers: a position lies outside [0, lx] x [0, ly]");
    double2* w = nullptr;
    const CshPlan* p = nullptr;
    hipStream_t st = nullptr;
    int rc = csh_ctx_state(c, &w, &p, &st);
    if (rc) return rc;
    const int nx = p->nx, ny = p->ny, ld = spec_ld(nx);
    TARA_CHECK(nx >= 16 && ny >= 8, TARA_ERR_UNSUPPORTED, "tracers: grid too small");
    DevBuf spec, real, pos, flag;
    if ((rc = spec.alloc(2 * (size_t)ld * ny * sizeof(double2))) || (rc = real.alloc(2 * (size_t)nx * ny * 8)) ||
        (rc = pos.alloc(2 * (size_t)(np ? np : 1) * 8)) || (rc = flag.alloc(sizeof(int))))
        return rc;
    double* dxp = pos.as<double>();
    double* dyp = dxp + np;
    if (np) {
        TARA_HIP(hipMemcpyAsync(dxp, xp, (size_t)np * 8, hipMemcpyHostToDevice, st));
        TARA_HIP(hipMemcpyAsync(dyp, yp, (size_t)np * 8, hipMemcpyHostToDevice, st));
    }
    TARA_HIP(hipMemsetAsync(spec.p, 0, spec.bytes, st));
    TARA_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), st));
    const size_t nr = (size_t)nx * ny;
    for (int k = 0; k < nsteps; ++k) {
        if ((rc = tara_csh_advance(c, 1))) return rc;
        hipLaunchKernelGGL(k_tracer_spec, dim3(grid_for((size_t)p->nh * ny)), dim3(256), 0, st, w, p->nh, ny, ld, lx,
                           ly, spec.as<double2>());
        TARA_HIP(hipGetLastError());
        if ((rc = col_fft(spec.as<double2>(), ny, ld, ld, (long)ld * ny, 2, +1, st))) return rc;
        if ((rc = row_c2r(spec.as<double2>(), real.as<double>(), nx, ld, 2 * ny, 1.0, st))) return rc;
        if (np) {
            hipLaunchKernelGGL(k_tracers, dim3((np + 255) / 256), dim3(256), 0, st, real.as<double>(),
                               real.as<double>() + nr, nx, ny, lx, ly, dt, periodic, np, dxp, dyp, flag.as<int>());
            TARA_HIP(hipGetLastError());
        }
    }
    int bad = 0;
    if (np) {
        TARA_HIP(hipMemcpyAsync(xp, dxp, (size_t)np * 8, hipMemcpyDeviceToHost, st));
        TARA_HIP(hipMemcpyAsync(yp, dyp, (size_t)np * 8, hipMemcpyDeviceToHost, st));
    }
    TARA_HIP(hipMemcpyAsync(&bad, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    TARA_HIP(hipStreamSynchronize(st));
    TARA_CHECK(!bad, TARA_ERR_NONFINITE, "tracers: a particle position became non-finite");
    return TARA_OK;
}

extern "C" int tara_csh_record(tara_csh* c, int nframes, int every, double* frames)
{
    TARA_CHECK(c && frames && nframes >= 0 && every >= 1, TARA_ERR_ARG, "bad argument");
    double2* w = nullptr;
    const CshPlan* p = nullptr;
    hipStream_t st = nullptr;
    int rc = csh_ctx_state(c, &w, &p, &st);
    if (rc) return rc;
    const size_t per = (size_t)p->nh * p->ny * 2;
    for (int f = 0; f < nframes; ++f) {
        if ((rc = tara_csh_advance(c, every))) return rc;
        if ((rc = tara_csh_download(c, frames + f * per))) return rc;
    }
    return TARA_OK;
}

extern "C" int tara_csh_ic_shear(tara_csh* c, double lx_in, double ly_in, double w0, double m0, double d_in)
{
    double2* w = nullptr;
    const CshPlan* p = nullptr;
    hipStream_t st = nullptr;
    int rc = csh_ctx_state(c, &w, &p, &st);
    if (rc) return rc;
    const int nx = p->nx, ny = p->ny, nh = p->nh, ld = spec_ld(nx);
    TARA_CHECK(nx >= 16 && ny >= 8, TARA_ERR_UNSUPPORTED, "ic: grid too small");
    const double pi = 3.14159265358979323846;  // src/main2csh.f95:9
    const double lx = lx_in * pi, ly = ly_in * pi;
    DevBuf real, spec, flat;
    if ((rc = real.alloc((size_t)nx * ny * 8)) || (rc = spec.alloc((size_t)ld * ny * sizeof(double2))) ||
        (rc = flat.alloc((size_t)nh * ny * sizeof(double2))))
        return rc;
    hipLaunchKernelGGL(k_shear_layer, dim3(grid_for((size_t)nx * ny)), dim3(256), 0, st, real.as<double>(), nx, ny,
                       lx / nx, ly / ny, pi, w0, m0, d_in * pi);
    TARA_HIP(hipGetLastError());
    if ((rc = row_r2c(real.as<double>(), spec.as<double2>(), nx, ld, ny, st))) return rc;
    if ((rc = col_fft(spec.as<double2>(), ny, ld, ld, (long)ld * ny, 1, -1, st))) return rc;
    TARA_HIP(hipMemcpy2DAsync(flat.p, nh * sizeof(double2), spec.p, ld * sizeof(double2), nh * sizeof(double2), ny,
                              hipMemcpyDeviceToDevice, st));
    if ((rc = to_kymajor(flat.as<double2>(), w, nh, ny, 1, st))) return rc;
    TARA_HIP(hipStreamSynchronize(st));
    return csh_ctx_zero_history(c);
}

namespace tara {
int csh_random_phase(double2* w, int nx, int ny, unsigned seed, double k0, double energy, hipStream_t st)
{
    TARA_CHECK(k0 > 0 && energy >= 0, TARA_ERR_ARG, "ic: k0 must be positive, energy non-negative");
    const double e0 = random_phase_energy(nx, ny, k0);
    const double scale = e0 > 0 ? std::sqrt(energy / e0) : 0.0;
    const int nh = nx / 2 + 1;
    hipLaunchKernelGGL(k_random_phase, dim3(grid_for((size_t)nh * ny)), dim3(256), 0, st, w, nh, ny, seed, k0,
                       scale);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}
}  // namespace tara

extern "C" int tara_csh_ic_random(tara_csh* c, unsigned seed, double k0, double energy)
{
    double2* w = nullptr;
    const CshPlan* p = nullptr;
    hipStream_t st = nullptr;
    int rc = csh_ctx_state(c, &w, &p, &st);
    if (rc) return rc;
    if ((rc = csh_random_phase(w, p->nx, p->ny, seed, k0, energy, st))) return rc;
    TARA_HIP(hipStreamSynchronize(st));
    return csh_ctx_zero_history(c);
}

extern "C" int tara_philox4x32_10(const unsigned ctr[4], const unsigned key[2], unsigned out[4])
{
    TARA_CHECK(ctr && key && out, TARA_ERR_ARG, "null argument");
    const U4 r = philox4x32_10(U4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    out[3] = r.w;
    return TARA_OK;
}
