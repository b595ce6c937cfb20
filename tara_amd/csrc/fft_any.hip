// Transforms of any length: the fallback behind col_fft / row_r2c / row_c2r
// for lengths the register-resident kernels do not cover (FFTW accepts any
// size; a kept main may ask for Nx = 96 or 100).
//
// A complex transform of length n = r1 r2 ... rs runs as s Stockham
// (autosort, decimation in frequency) passes between two device buffers.
// The pass with radix r on the current sub-length L (stride s0 = n / L) maps
//   y[q + s0 (r p + u)] = w_L^{p u} * sum_j x[q + s0 (p + j m)] w_r^{j u},
//   m = L / r, p < m, q < s0, u < r,
// and the output of the last pass is in natural order.  One thread computes
// one output (O(r) work), so large prime factors still spread over the
// device.  Roots come from one table of e^{-2 pi i k/n} (long double, rounded
// once).  Radices: 4 while possible, then 2, 3, 5, 7, then the remaining
// primes.
//
// Real transforms keep FFTW's semantics: r2c is the complex transform of the
// real row, of which entries 0..n/2 are kept; c2r builds the Hermitian row
// from entries 0..n/2 with the imaginary parts of entry 0 and (even n) entry
// n/2 taken as zero, which is what FFTW's c2r reads, and keeps the real part.
// Workspace is allocated per call and the stream is synchronised before the
// call returns.
#include <vector>

#include "common.h"
#include "fft_api.h"

namespace tara {

namespace {

std::vector<int> radices(int n)
{
    std::vector<int> r;
    while (n % 4 == 0) {
        r.push_back(4);
        n /= 4;
    }
    for (int p = 2; p <= n; ++p)
        while (n % p == 0) {
            r.push_back(p);
            n /= p;
        }
    return r;
}

__device__ __forceinline__ double2 root(const double2* tw, long k, int dir)
{
    const double2 w = tw[k];
    return dir < 0 ? w : make_double2(w.x, -w.y);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// One pass over ntrans transforms; element e of transform t at
// buf[e * es + t * ts].  tfast: consecutive threads take consecutive t.
__global__ void k_any_pass(const double2* __restrict__ x, double2* __restrict__ y, int n, int L, int r, long ntrans,
                           long es, long ts, int tfast, const double2* __restrict__ tw, int dir)
{
    const long total = ntrans * n;
    const int s0 = n / L, m = L / r;
    const long nr = n / r, nl = n / L;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
        long t, e;
        if (tfast) {
            t = idx % ntrans;
            e = idx / ntrans;
        } else {
            t = idx / n;
            e = idx % n;
        }
        const long q = e % s0, rest = e / s0;
        const long u = rest % r, p = rest / r;
        const double2* xt = x + t * ts;
        double2 acc = make_double2(0.0, 0.0);
        for (int j = 0; j < r; ++j) {
            const double2 a = xt[(q + s0 * (p + (long)j * m)) * es];
            const double2 w = root(tw, ((j * u) % r) * nr, dir);
            acc.x += a.x * w.x - a.y * w.y;
            acc.y += a.x * w.y + a.y * w.x;
        }
        y[t * ts + e * es] = (p * u == 0) ? acc : cmul(acc, root(tw, (p * u) * nl, dir));
    }
}

// Strided columns <-> transform-fastest workspace: element j of column c of
// plane pl at a[pl * pstride + j * es + c] <-> W[j * ntrans + pl * ncols + c].
__global__ void k_any_gather(const double2* __restrict__ a, double2* __restrict__ w, int n, int ncols, long es,
                             long pstride, long ntrans)
{
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < ntrans * n; idx += (long)gridDim.x * blockDim.x) {
        const long t = idx % ntrans, j = idx / ntrans, c = t % ncols, pl = t / ncols;
        w[idx] = a[pl * pstride + j * es + c];
    }
}

__global__ void k_any_scatter(const double2* __restrict__ w, double2* __restrict__ a, int n, int ncols, long es,
                              long pstride, long ntrans)
{
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < ntrans * n; idx += (long)gridDim.x * blockDim.x) {
        const long t = idx % ntrans, j = idx / ntrans, c = t % ncols, pl = t / ncols;
        a[pl * pstride + j * es + c] = w[idx];
    }
}

// Real rows (pitch nx) -> complex rows (pitch nx).
__global__ void k_any_real_in(const double* __restrict__ x, double2* __restrict__ w, long total)
{
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x)
        w[idx] = make_double2(x[idx], 0.0);
}

// Complex rows (pitch nx) -> half spectra (pitch ld), entries nh..ld-1 zeroed.
__global__ void k_any_half_out(const double2* __restrict__ w, double2* __restrict__ X, int nx, int ld, long nrows)
{
    const int nh = nx / 2 + 1;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < nrows * ld; idx += (long)gridDim.x * blockDim.x) {
        const long row = idx / ld, k = idx % ld;
        X[idx] = k < nh ? w[row * nx + k] : make_double2(0.0, 0.0);
    }
}

// Half spectra (pitch ld) -> Hermitian complex rows (pitch nx).
__global__ void k_any_herm_in(const double2* __restrict__ X, double2* __restrict__ w, int nx, int ld, long nrows)
{
    const int nh = nx / 2 + 1;
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < nrows * nx; idx += (long)gridDim.x * blockDim.x) {
        const long row = idx / nx, k = idx % nx;
        const double2* src = X + row * ld;
        double2 v;
        if (k == 0 || 2 * k == nx)
            v = make_double2(src[k].x, 0.0);
        else if (k < nh)
            v = src[k];
        else {
            const double2 c = src[nx - k];
            v = make_double2(c.x, -c.y);
        }
        w[idx] = v;
    }
}

// Real part of complex rows (pitch nx), times s.
__global__ void k_any_real_out(const double2* __restrict__ w, double* __restrict__ x, double s, long total)
{
    for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x)
        x[idx] = w[idx].x * s;
}

unsigned blocks_for(long n)
{
    long b = (n + 255) / 256;
    return (unsigned)(b > 16384 ? 16384 : (b < 1 ? 1 : b));
}

// Runs the passes over w0 (w1: scratch of the same size); returns the buffer
// that holds the result.
int run_passes(double2* w0, double2* w1, int n, long ntrans, long es, long ts, int tfast, int dir, hipStream_t st,
               double2** result)
{
    *result = w0;
    if (n == 1) return TARA_OK;
    const double2* tw = unit_roots(n);
    TARA_CHECK(tw, TARA_ERR_ALLOC, "root table allocation failed (n = %d)", n);
    double2 *x = w0, *y = w1;
    int L = n;
    for (int r : radices(n)) {
        hipLaunchKernelGGL(k_any_pass, dim3(blocks_for(ntrans * n)), dim3(256), 0, st, x, y, n, L, r, ntrans, es, ts,
                           tfast, tw, dir);
        TARA_HIP(hipGetLastError());
        L /= r;
        std::swap(x, y);
    }
    *result = x;
    return TARA_OK;
}

}  // namespace

int any_col_fft(double2* a, int n, int ncols, long es, long pstride, int nplanes, int dir, hipStream_t st)
{
    TARA_CHECK(n >= 1 && ncols >= 1 && nplanes >= 1, TARA_ERR_ARG, "any_col_fft: bad shape");
    if (n == 1) return TARA_OK;
    const long ntrans = (long)ncols * nplanes;
    DevBuf b0, b1;
    int rc;
    if ((rc = b0.alloc((size_t)ntrans * n * sizeof(double2))) || (rc = b1.alloc((size_t)ntrans * n * sizeof(double2))))
        return rc;
    hipLaunchKernelGGL(k_any_gather, dim3(blocks_for(ntrans * n)), dim3(256), 0, st, a, b0.as<double2>(), n, ncols, es,
                       pstride, ntrans);
    TARA_HIP(hipGetLastError());
    double2* res = nullptr;
    if ((rc = run_passes(b0.as<double2>(), b1.as<double2>(), n, ntrans, ntrans, 1, 1, dir, st, &res))) return rc;
    hipLaunchKernelGGL(k_any_scatter, dim3(blocks_for(ntrans * n)), dim3(256), 0, st, res, a, n, ncols, es, pstride,
                       ntrans);
    TARA_HIP(hipGetLastError());
    TARA_HIP(hipStreamSynchronize(st));
    return TARA_OK;
}

int any_row_r2c(const double* x, double2* X, int nx, int ld, int nrows, hipStream_t st)
{
    TARA_CHECK(nx >= 1 && nrows >= 1 && ld >= nx / 2 + 1, TARA_ERR_ARG, "any_row_r2c: bad shape");
    const long total = (long)nrows * nx;
    DevBuf b0, b1;
    int rc;
    if ((rc = b0.alloc((size_t)total * sizeof(double2))) || (rc = b1.alloc((size_t)total * sizeof(double2)))) return rc;
    hipLaunchKernelGGL(k_any_real_in, dim3(blocks_for(total)), dim3(256), 0, st, x, b0.as<double2>(), total);
    TARA_HIP(hipGetLastError());
    double2* res = nullptr;
    if ((rc = run_passes(b0.as<double2>(), b1.as<double2>(), nx, nrows, 1, nx, 0, -1, st, &res))) return rc;
    hipLaunchKernelGGL(k_any_half_out, dim3(blocks_for((long)nrows * ld)), dim3(256), 0, st, res, X, nx, ld,
                       (long)nrows);
    TARA_HIP(hipGetLastError());
    TARA_HIP(hipStreamSynchronize(st));
    return TARA_OK;
}

int any_row_c2r(const double2* X, double* x, int nx, int ld, int nrows, double s, hipStream_t st)
{
    TARA_CHECK(nx >= 1 && nrows >= 1 && ld >= nx / 2 + 1, TARA_ERR_ARG, "any_row_c2r: bad shape");
    const long total = (long)nrows * nx;
    DevBuf b0, b1;
    int rc;
    if ((rc = b0.alloc((size_t)total * sizeof(double2))) || (rc = b1.alloc((size_t)total * sizeof(double2)))) return rc;
    hipLaunchKernelGGL(k_any_herm_in, dim3(blocks_for(total)), dim3(256), 0, st, X, b0.as<double2>(), nx, ld,
                       (long)nrows);
    TARA_HIP(hipGetLastError());
    double2* res = nullptr;
    if ((rc = run_passes(b0.as<double2>(), b1.as<double2>(), nx, nrows, 1, nx, 0, +1, st, &res))) return rc;
    hipLaunchKernelGGL(k_any_real_out, dim3(blocks_for(total)), dim3(256), 0, st, res, x, s, total);
    TARA_HIP(hipGetLastError());
    TARA_HIP(hipStreamSynchronize(st));
    return TARA_OK;
}

}  // namespace tara
