// Plain multi-dimensional r2c / c2r transforms with FFTW's semantics
// (lib/fftw/api/f77funcs.h:206-218, 300-312; ordering of
// lib/fftw/rdft/rank-geq2-rdft2.c:40-68): r2c = x pass (real -> half-length
// complex packing) then complex passes over y (and z); c2r = complex passes
// over z and y first, then the x pass, which reads only the real part of the
// kx = 0 and kx = Nx/2 entries.  Unnormalised in both directions.
//
// These are the building blocks the fused solver passes are made of, exposed
// on their own for parity testing and for the driver-side transforms.
#include <cstring>
#include <memory>

#include "common.h"
#include "fft_api.h"
#include "passes.h"

namespace tara {

// In-place complex FFT along one strided axis.  Tile = TC consecutive
// contiguous columns (i0 .. i0+TC-1) of plane blockIdx.y; element n of column c
// at a[plane*pstride + n*es + i0 + c].
template <int N, int DIR>
__global__ __launch_bounds__(ColShape<N>::NT) void k_col_fft(double2* __restrict__ a, long es, long pstride,
                                                              const double2* __restrict__ tw)
{
    using S = ColShape<N>;
    constexpr int TC = S::TC, NT = S::NT, PER = S::PER;
    __shared__ double2 lds[TC * N];
    double2* base = a + (size_t)blockIdx.y * pstride + (size_t)blockIdx.x * TC;
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int e = threadIdx.x + m * NT, j = e / TC, c = e % TC;
        lds[e] = base[j * es + c];
    }
    __syncthreads();
    lds_fft<N, DIR, NT, TC, true>(lds, TC, 1, tw);
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int e = threadIdx.x + m * NT, j = e / TC, c = e % TC;
        base[j * es + c] = lds[e];
    }
}

// x pass, real rows (pitch nx) -> half spectra rows (pitch ld); padding zeroed.
template <int M>
__global__ __launch_bounds__(RowShape<M>::NT) void k_row_r2c(const double* __restrict__ x, double2* __restrict__ X,
                                                              int ld, int nh, const double2* __restrict__ tw,
                                                              const double2* __restrict__ rctw)
{
    using S = RowShape<M>;
    constexpr int RB = S::RB, NT = S::NT, PER = S::PER, P = S::PITCH;
    __shared__ double2 lds[RB * P];
    const size_t row0 = (size_t)blockIdx.x * RB;
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int e = threadIdx.x + m * NT, r = e / M, k = e % M;
        const double2* xr = reinterpret_cast<const double2*>(x + (row0 + r) * (size_t)(2 * M));
        lds[r * P + k] = xr[k];
    }
    __syncthreads();
    lds_fft<M, -1, NT, RB, false>(lds, 1, P, tw);
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int e = threadIdx.x + m * NT, r = e / M, k = e % M;
        double2 xm;
        double2* dst = X + (row0 + r) * ld;
        dst[k] = r2c_unpack(lds[r * P + k], lds[r * P + (k == 0 ? 0 : M - k)], k, rctw, &xm);
        if (k == 0) dst[M] = xm;
        if (k < ld - nh) dst[nh + k] = make_double2(0.0, 0.0);
    }
}

// x pass, half spectra rows (pitch ld) -> real rows (pitch nx), scaled by s.
template <int M>
__global__ __launch_bounds__(RowShape<M>::NT) void k_row_c2r(const double2* __restrict__ X, double* __restrict__ x,
                                                              int ld, double s, const double2* __restrict__ tw,
                                                              const double2* __restrict__ rctw)
{
    using S = RowShape<M>;
    constexpr int RB = S::RB, NT = S::NT, PER = S::PER, P = S::PITCH;
    __shared__ double2 lds[RB * P];
    const size_t row0 = (size_t)blockIdx.x * RB;
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int e = threadIdx.x + m * NT, r = e / M, k = e % M;
        const double2* src = X + (row0 + r) * ld;
        lds[r * P + k] = c2r_pack(src[k], src[M - k], k, rctw);
    }
    __syncthreads();
    lds_fft<M, +1, NT, RB, false>(lds, 1, P, tw);
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int e = threadIdx.x + m * NT, r = e / M, k = e % M;
        double2* xr = reinterpret_cast<double2*>(x + (row0 + r) * (size_t)(2 * M));
        xr[k] = cscale(lds[r * P + k], s);
    }
}

#define TARA_LEN_SWITCH(n, FN, ...)                 \
    switch (n) {                                    \
        case 8: return FN<8>(__VA_ARGS__);          \
        case 16: return FN<16>(__VA_ARGS__);        \
        case 32: return FN<32>(__VA_ARGS__);        \
        case 64: return FN<64>(__VA_ARGS__);        \
        case 128: return FN<128>(__VA_ARGS__);      \
        case 256: return FN<256>(__VA_ARGS__);      \
        case 512: return FN<512>(__VA_ARGS__);      \
        case 1024: return FN<1024>(__VA_ARGS__);    \
        case 2048: return FN<2048>(__VA_ARGS__);    \
        case 4096: return FN<4096>(__VA_ARGS__);    \
        default: set_error("unsupported transform length %d", n); return TARA_ERR_UNSUPPORTED; \
    }

template <int N>
static int col_fft_n(double2* a, int ld, long es, long pstride, int nplanes, int dir, hipStream_t st)
{
    using S = ColShape<N>;
    dim3 g(ld / S::TC, nplanes), b(S::NT);
    const double2* tw = twiddles(N);
    TARA_CHECK(tw, TARA_ERR_ALLOC, "twiddle table allocation failed");
    if (dir < 0)
        hipLaunchKernelGGL((k_col_fft<N, -1>), g, b, 0, st, a, es, pstride, tw);
    else
        hipLaunchKernelGGL((k_col_fft<N, +1>), g, b, 0, st, a, es, pstride, tw);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

// Lengths and batch shapes the register-resident kernels take; anything else
// goes to the mixed-radix transforms of fft_any.hip.
static bool fast_col(int n, int ld)
{
    const int tc = n >= 4096 ? 2 : n >= 2048 ? 4 : 8;
    return is_pow2(n) && n >= 8 && n <= 4096 && ld % tc == 0;
}

static bool fast_row(int nx, int nrows)
{
    const int m = nx / 2, rb = m >= 2048 ? 1 : m >= 1024 ? 2 : m >= 512 ? 4 : 8;
    return nx % 2 == 0 && is_pow2(m) && m >= 8 && m <= 4096 && nrows % rb == 0;
}

int col_fft(double2* a, int n, int ld, long es, long pstride, int nplanes, int dir, hipStream_t st)
{
    if (!fast_col(n, ld)) return any_col_fft(a, n, ld, es, pstride, nplanes, dir, st);
    TARA_LEN_SWITCH(n, col_fft_n, a, ld, es, pstride, nplanes, dir, st)
}

template <int M>
static int row_r2c_n(const double* x, double2* X, int ld, int nrows, hipStream_t st)
{
    using S = RowShape<M>;
    TARA_CHECK(nrows % S::RB == 0, TARA_ERR_UNSUPPORTED, "row count %d not a multiple of %d", nrows, S::RB);
    const double2 *tw = twiddles(M), *rc = rc_twiddles(2 * M);
    TARA_CHECK(tw && rc, TARA_ERR_ALLOC, "twiddle table allocation failed");
    hipLaunchKernelGGL(k_row_r2c<M>, dim3(nrows / S::RB), dim3(S::NT), 0, st, x, X, ld, M + 1, tw, rc);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

int row_r2c(const double* x, double2* X, int nx, int ld, int nrows, hipStream_t st)
{
    if (!fast_row(nx, nrows)) return any_row_r2c(x, X, nx, ld, nrows, st);
    TARA_LEN_SWITCH(nx / 2, row_r2c_n, x, X, ld, nrows, st)
}

template <int M>
static int row_c2r_n(const double2* X, double* x, int ld, int nrows, double s, hipStream_t st)
{
    using S = RowShape<M>;
    TARA_CHECK(nrows % S::RB == 0, TARA_ERR_UNSUPPORTED, "row count %d not a multiple of %d", nrows, S::RB);
    const double2 *tw = twiddles(M), *rc = rc_twiddles(2 * M);
    TARA_CHECK(tw && rc, TARA_ERR_ALLOC, "twiddle table allocation failed");
    hipLaunchKernelGGL(k_row_c2r<M>, dim3(nrows / S::RB), dim3(S::NT), 0, st, X, x, ld, s, tw, rc);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

int row_c2r(const double2* X, double* x, int nx, int ld, int nrows, double s, hipStream_t st)
{
    if (!fast_row(nx, nrows)) return any_row_c2r(X, x, nx, ld, nrows, s, st);
    TARA_LEN_SWITCH(nx / 2, row_c2r_n, X, x, ld, nrows, s, st)
}

int spec_ld(int nx) { return (nx / 2 + 1 + 7) / 8 * 8; }

}  // namespace tara

using namespace tara;

namespace {
int check_dims(int nx, int ny, int nz)
{
    TARA_CHECK(nx >= 1 && ny >= 1 && nz >= 1 && (double)nx * ny * nz <= (double)(1L << 31), TARA_ERR_UNSUPPORTED,
               "transform size %dx%dx%d unsupported (each >= 1, at most 2^31 points)", nx, ny, nz);
    return TARA_OK;
}

// Host-array transform: copy in (host layout pitch nh -> device pitch ld), run, copy out.
// A 1-D transform runs as row 0 of a block of 8 rows (the row kernels' unit).
// Any sizes: lengths the fast kernels do not take run on the mixed-radix path.
// Device work space of the host-array transforms, per device, kept between
// calls and grown when a larger transform comes (the kept mains call the same
// sizes every step).
struct HostWs {
    DevBuf spec, real;
};
std::mutex g_ws_mu;
std::map<int, std::unique_ptr<HostWs>> g_ws;

int fft_host(int nx, int ny, int nz, const double* in, double* out, bool forward)
{
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    const int nh = nx / 2 + 1, ld = spec_ld(nx);
    const size_t rows = (size_t)ny * nz, drows = rows < 8 ? 8 : rows;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto& ws = g_ws[current_device()];
    if (!ws) ws = std::make_unique<HostWs>();
    DevBuf& spec = ws->spec;
    DevBuf& real = ws->real;
    if (spec.bytes < drows * ld * sizeof(double2) && (rc = spec.alloc(drows * ld * sizeof(double2)))) return rc;
    if (real.bytes < drows * nx * sizeof(double) && (rc = real.alloc(drows * nx * sizeof(double)))) return rc;
    hipStream_t st = nullptr;
    double2* S = spec.as<double2>();
    if (forward) {
        if (drows > rows) TARA_HIP(hipMemset(real.p, 0, drows * nx * sizeof(double)));
        TARA_HIP(hipMemcpy(real.p, in, rows * nx * sizeof(double), hipMemcpyHostToDevice));
        if ((rc = row_r2c(real.as<double>(), S, nx, ld, (int)drows, st))) return rc;
        if (ny > 1 && (rc = col_fft(S, ny, ld, ld, (long)ld * ny, nz, -1, st))) return rc;
        if (nz > 1 && (rc = col_fft(S, nz, ld, (long)ld * ny, ld, ny, -1, st))) return rc;
        TARA_HIP(hipMemcpy2D(out, nh * sizeof(double2), S, ld * sizeof(double2), nh * sizeof(double2), rows,
                             hipMemcpyDeviceToHost));
    } else {
        TARA_HIP(hipMemset(S, 0, drows * ld * sizeof(double2)));
        TARA_HIP(hipMemcpy2D(S, ld * sizeof(double2), in, nh * sizeof(double2), nh * sizeof(double2), rows,
                             hipMemcpyHostToDevice));
        if (nz > 1 && (rc = col_fft(S, nz, ld, (long)ld * ny, ld, ny, +1, st))) return rc;
        if (ny > 1 && (rc = col_fft(S, ny, ld, ld, (long)ld * ny, nz, +1, st))) return rc;
        if ((rc = row_c2r(S, real.as<double>(), nx, ld, (int)drows, 1.0, st))) return rc;
        TARA_HIP(hipMemcpy(out, real.p, rows * nx * sizeof(double), hipMemcpyDeviceToHost));
    }
    TARA_HIP(hipDeviceSynchronize());
    return TARA_OK;
}
}  // namespace

extern "C" int tara_fft_r2c_1d(int nx, const double* in, double* out)
{
    TARA_CHECK(in && out, TARA_ERR_ARG, "null pointer");
    return fft_host(nx, 1, 1, in, out, true);
}
extern "C" int tara_fft_c2r_1d(int nx, const double* in, double* out)
{
    TARA_CHECK(in && out, TARA_ERR_ARG, "null pointer");
    return fft_host(nx, 1, 1, in, out, false);
}
extern "C" int tara_fft_r2c_2d(int nx, int ny, const double* in, double* out)
{
    TARA_CHECK(in && out, TARA_ERR_ARG, "null pointer");
    return fft_host(nx, ny, 1, in, out, true);
}
extern "C" int tara_fft_c2r_2d(int nx, int ny, const double* in, double* out)
{
    TARA_CHECK(in && out, TARA_ERR_ARG, "null pointer");
    return fft_host(nx, ny, 1, in, out, false);
}
extern "C" int tara_fft_r2c_3d(int nx, int ny, int nz, const double* in, double* out)
{
    TARA_CHECK(in && out, TARA_ERR_ARG, "null pointer");
    return fft_host(nx, ny, nz, in, out, true);
}
extern "C" int tara_fft_c2r_3d(int nx, int ny, int nz, const double* in, double* out)
{
    TARA_CHECK(in && out, TARA_ERR_ARG, "null pointer");
    return fft_host(nx, ny, nz, in, out, false);
}
