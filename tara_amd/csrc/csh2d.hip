// 2D incompressible hydro in vorticity-streamfunction form on gfx950:
// the right-hand side of src/derive2csh.f95 and the AB2 / RK4 updates
// (src/ab2csh.f95; classic RK4), as three fused HBM sweeps per RHS.
//
//   pass 1  k_csh_yinv : per column tile (TC kx columns x Ny) read omega_hat,
//                        form the four spectra of derive2csh.f95:29-71
//                        (i ky psi, -i kx psi, i kx omega, i ky omega; psi = omega/k^2,
//                        psi(0,0) = 0) and run the backward y FFTs -> G0..G3.
//   pass 2  k_csh_x    : per row, four x c2r (half-length packing), normalise by
//                        Nx*Ny and form ux*d_x omega + uy*d_y omega
//                        (derive2csh.f95:89-98), one x r2c -> H.
//                        (The reference transforms the two products separately
//                        (:100-106); the sum is transformed once, which is the
//                        same linear map.)
//   pass 3  k_csh_yfwd : per column tile, forward y FFT of H, circular dealias
//                        sqrt(kx^2+ky^2) >= (Nx+Ny)/6 on the nonlinear part only
//                        (:108-125), add nu k^2 omega (:127-139), negate (:141-145),
//                        then the fused update (RHS out / AB2 / RK4 stage).
//
// Wavenumbers: derive2csh uses pi = 3.1428 (a default-real literal) and
// Lx = Ly = 2 pi (:16-21), so kx = i-1 and ky = j-1 or j-1-Ny are exact integers;
// the host builds the tables with that exact expression (csh_tables).
#include "common.h"
#include "csh2d.h"
#include "passes.h"

namespace tara {

// --------------------------------------------------------------------- pass 1
template <int NY>
__global__ __launch_bounds__(ColShape<NY>::NT) void k_csh_yinv(const double2* __restrict__ w, int ld,
                                                                const double* __restrict__ kx,
                                                                const double* __restrict__ ky,
                                                                double2* __restrict__ g0, double2* __restrict__ g1,
                                                                double2* __restrict__ g2, double2* __restrict__ g3,
                                                                const double2* __restrict__ tw)
{
    using S = ColShape<NY>;
    constexpr int TC = S::TC, NT = S::NT, PER = S::PER;
    __shared__ double2 lds[TC * NY];
    const int i0 = blockIdx.x * TC;
    double2 wv[PER];
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int e = threadIdx.x + m * NT, j = e / TC, c = e % TC;
        wv[m] = w[(size_t)j * ld + i0 + c];
    }
    // i kx omega and i ky omega first, then omega -> psi in place, then the
    // velocity spectra +i ky psi and -i kx psi.
    const double kxv = kx[i0 + threadIdx.x % TC];  // NT % TC == 0: one column per thread
#pragma unroll 1
    for (int f = 0; f < 4; ++f) {
        if (f == 2) {
#pragma unroll
            for (int m = 0; m < PER; ++m) {
                const int j = (threadIdx.x + m * NT) / TC;
                const double kyv = ky[j];
                const double k2 = kxv * kxv + kyv * kyv;
                wv[m] = (k2 != 0.0) ? make_double2(wv[m].x / k2, wv[m].y / k2) : make_double2(0.0, 0.0);
            }
        }
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int e = threadIdx.x + m * NT, j = e / TC;
            const double kyv = ky[j];
            const double2 z = wv[m];
            double2 v;
            if (f == 0)
                v = make_double2(-(kxv * z.y), kxv * z.x);      // i kx omega
            else if (f == 1)
                v = make_double2(-(kyv * z.y), kyv * z.x);      // i ky omega
            else if (f == 2)
                v = make_double2(-(kyv * z.y), kyv * z.x);      // +i ky psi
            else
                v = make_double2(kxv * z.y, -(kxv * z.x));      // -i kx psi
            lds[e] = v;
        }
        __syncthreads();
        lds_fft<NY, +1, NT, TC, true>(lds, TC, 1, tw);
        // G0 = ux, G1 = uy, G2 = d omega/dx, G3 = d omega/dy
        double2* o = f == 0 ? g2 : f == 1 ? g3 : f == 2 ? g0 : g1;
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int e = threadIdx.x + m * NT, j = e / TC, c = e % TC;
            o[(size_t)j * ld + i0 + c] = lds[e];
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------------- pass 2
template <int M>
__global__ __launch_bounds__(RowShape<M>::NT) void k_csh_x(const double2* __restrict__ g0,
                                                            const double2* __restrict__ g1,
                                                            const double2* __restrict__ g2,
                                                            const double2* __restrict__ g3, int ld, int nh,
                                                            double scale, double2* __restrict__ h,
                                                            const double2* __restrict__ tw,
                                                            const double2* __restrict__ rctw)
{
    using S = RowShape<M>;
    constexpr int RB = S::RB, NT = S::NT, PER = S::PER, P = S::PITCH;
    __shared__ double2 lds[RB * P];
    const int row0 = blockIdx.x * RB;
    const double2* gin[4] = {g0, g1, g2, g3};
    double2 z[4][PER];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int e = threadIdx.x + m * NT, r = e / M, k = e % M;
            const double2* src = gin[f] + (size_t)(row0 + r) * ld;
            lds[r * P + k] = c2r_pack(src[k], src[M - k], k, rctw);
        }
        __syncthreads();
        lds_fft<M, +1, NT, RB, false>(lds, 1, P, tw);
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int e = threadIdx.x + m * NT, r = e / M, k = e % M;
            z[f][m] = lds[r * P + k];
        }
        __syncthreads();
    }
    // real-space products at x = 2k (real parts) and x = 2k+1 (imaginary parts)
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int e = threadIdx.x + m * NT, r = e / M, k = e % M;
        const double ev = (z[0][m].x * z[2][m].x + z[1][m].x * z[3][m].x) * scale;
        const double od = (z[0][m].y * z[2][m].y + z[1][m].y * z[3][m].y) * scale;
        lds[r * P + k] = make_double2(ev, od);
    }
    __syncthreads();
    lds_fft<M, -1, NT, RB, false>(lds, 1, P, tw);
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int e = threadIdx.x + m * NT, r = e / M, k = e % M;
        const double2 zk = lds[r * P + k];
        const double2 zmk = lds[r * P + (k == 0 ? 0 : M - k)];
        double2 xm;
        double2* dst = h + (size_t)(row0 + r) * ld;
        dst[k] = r2c_unpack(zk, zmk, k, rctw, &xm);
        if (k == 0) dst[M] = xm;
        if (k < ld - nh) dst[nh + k] = make_double2(0.0, 0.0);
    }
}

// --------------------------------------------------------------------- pass 3
struct YfwdArgs {
    const double2* h;      // nonlinear term, x-transformed
    const double2* ws;     // stage input omega_s (viscous term)
    double2* base;         // AB2: state (in/out); RK4: omega_n (stage 4: out)
    double2* aux;          // AB2: history f_{n-1} (in/out); RK4: accumulator
    double2* next;         // RK4: next stage input (out); RHS mode: f (out)
    const double* kx;
    const double* ky;
    const int* imin;       // dealias: modes with i >= imin[j] drop the nonlinear term
    int ld;
    double nu, dt;
};

template <int NY, int MODE>
__global__ __launch_bounds__(ColShape<NY>::NT) void k_csh_yfwd(YfwdArgs a, const double2* __restrict__ tw)
{
    using S = ColShape<NY>;
    constexpr int TC = S::TC, NT = S::NT, PER = S::PER;
    __shared__ double2 lds[TC * NY];
    const int i0 = blockIdx.x * TC;
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int e = threadIdx.x + m * NT, j = e / TC, c = e % TC;
        lds[e] = a.h[(size_t)j * a.ld + i0 + c];
    }
    __syncthreads();
    lds_fft<NY, -1, NT, TC, true>(lds, TC, 1, tw);
#pragma unroll
    for (int m = 0; m < PER; ++m) {
        const int e = threadIdx.x + m * NT, j = e / TC, c = e % TC;
        const int i = i0 + c;
        const size_t q = (size_t)j * a.ld + i;
        double2 nl = lds[e];
        if (i >= a.imin[j]) nl = make_double2(0.0, 0.0);
        const double kxv = a.kx[i], kyv = a.ky[j];
        const double vk = a.nu * (kxv * kxv + kyv * kyv);
        const double2 ws = a.ws[q];
        const double2 f = make_double2(-(nl.x + vk * ws.x), -(nl.y + vk * ws.y));
        if (MODE == CSH_RHS) {
            a.next[q] = f;
        } else if (MODE == CSH_AB2) {
            const double2 fo = a.aux[q];
            a.base[q] = make_double2(ws.x + (1.5 * f.x - 0.5 * fo.x) * a.dt, ws.y + (1.5 * f.y - 0.5 * fo.y) * a.dt);
            a.aux[q] = f;
        } else if (MODE == CSH_RK4_1) {
            const double c6 = a.dt / 6.0, c2 = 0.5 * a.dt;
            a.aux[q] = make_double2(ws.x + c6 * f.x, ws.y + c6 * f.y);
            a.next[q] = make_double2(ws.x + c2 * f.x, ws.y + c2 * f.y);
        } else if (MODE == CSH_RK4_2 || MODE == CSH_RK4_3) {
            const double c3 = a.dt / 3.0, cn = (MODE == CSH_RK4_2) ? 0.5 * a.dt : a.dt;
            const double2 w0 = a.base[q], acc = a.aux[q];
            a.aux[q] = make_double2(acc.x + c3 * f.x, acc.y + c3 * f.y);
            a.next[q] = make_double2(w0.x + cn * f.x, w0.y + cn * f.y);
        } else {  // CSH_RK4_4
            const double c6 = a.dt / 6.0;
            const double2 acc = a.aux[q];
            a.base[q] = make_double2(acc.x + c6 * f.x, acc.y + c6 * f.y);
        }
    }
}

// ------------------------------------------------------------ pointwise AB2
__global__ void k_ab2(const double2* __restrict__ w, double2* __restrict__ wn, const double2* __restrict__ fo,
                      const double2* __restrict__ fn, double dt, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const double2 a = w[q], o = fo[q], f = fn[q];
        wn[q] = make_double2(a.x + (1.5 * f.x - 0.5 * o.x) * dt, a.y + (1.5 * f.y - 0.5 * o.y) * dt);
    }
}

// ------------------------------------------------------------------ dispatch
template <int NY>
static int launch_yinv(const CshPlan& p, const double2* w, hipStream_t st)
{
    using S = ColShape<NY>;
    hipLaunchKernelGGL(k_csh_yinv<NY>, dim3(p.ld / S::TC), dim3(S::NT), 0, st, w, p.ld, p.kx.as<double>(),
                       p.ky.as<double>(), p.g[0].as<double2>(), p.g[1].as<double2>(), p.g[2].as<double2>(),
                       p.g[3].as<double2>(), p.tw_y);
    return TARA_OK;
}

template <int M>
static int launch_x(const CshPlan& p, hipStream_t st)
{
    using S = RowShape<M>;
    const double scale = 1.0 / ((double)p.nx * (double)p.ny * (double)p.nx * (double)p.ny);
    hipLaunchKernelGGL(k_csh_x<M>, dim3(p.ny / S::RB), dim3(S::NT), 0, st, p.g[0].as<double2>(),
                       p.g[1].as<double2>(), p.g[2].as<double2>(), p.g[3].as<double2>(), p.ld, p.nh, scale,
                       p.h.as<double2>(), p.tw_x, p.rctw);
    return TARA_OK;
}

template <int NY>
static int launch_yfwd(const CshPlan& p, int mode, const YfwdArgs& a, hipStream_t st)
{
    using S = ColShape<NY>;
    dim3 g(p.ld / S::TC), b(S::NT);
    switch (mode) {
        case CSH_RHS: hipLaunchKernelGGL((k_csh_yfwd<NY, CSH_RHS>), g, b, 0, st, a, p.tw_y); break;
        case CSH_AB2: hipLaunchKernelGGL((k_csh_yfwd<NY, CSH_AB2>), g, b, 0, st, a, p.tw_y); break;
        case CSH_RK4_1: hipLaunchKernelGGL((k_csh_yfwd<NY, CSH_RK4_1>), g, b, 0, st, a, p.tw_y); break;
        case CSH_RK4_2: hipLaunchKernelGGL((k_csh_yfwd<NY, CSH_RK4_2>), g, b, 0, st, a, p.tw_y); break;
        case CSH_RK4_3: hipLaunchKernelGGL((k_csh_yfwd<NY, CSH_RK4_3>), g, b, 0, st, a, p.tw_y); break;
        case CSH_RK4_4: hipLaunchKernelGGL((k_csh_yfwd<NY, CSH_RK4_4>), g, b, 0, st, a, p.tw_y); break;
        default: set_error("bad mode %d", mode); return TARA_ERR_ARG;
    }
    return TARA_OK;
}

#define TARA_SIZE_SWITCH(n, FN, ...)                \
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

static int yinv(const CshPlan& p, const double2* w, hipStream_t st) { TARA_SIZE_SWITCH(p.ny, launch_yinv, p, w, st) }
static int xpass(const CshPlan& p, hipStream_t st) { TARA_SIZE_SWITCH(p.nx / 2, launch_x, p, st) }
static int yfwd(const CshPlan& p, int mode, const YfwdArgs& a, hipStream_t st)
{
    TARA_SIZE_SWITCH(p.ny, launch_yfwd, p, mode, a, st)
}

// ------------------------------------------------------------------ host side
int csh_plan_init(CshPlan& p, int nx, int ny)
{
    TARA_CHECK(is_pow2(nx) && is_pow2(ny) && nx >= 32 && ny >= 32 && nx <= 8192 && ny <= 4096, TARA_ERR_UNSUPPORTED,
               "derive2csh: grid %dx%d unsupported (powers of two, nx 32..8192, ny 32..4096)", nx, ny);
    p.nx = nx;
    p.ny = ny;
    p.nh = nx / 2 + 1;
    p.ld = (p.nh + 7) / 8 * 8;
    const size_t ns = (size_t)p.ld * ny;
    int rc;
    for (int f = 0; f < 4; ++f)
        if ((rc = p.g[f].alloc(ns * sizeof(double2)))) return rc;
    if ((rc = p.h.alloc(ns * sizeof(double2)))) return rc;
    TARA_HIP(hipMemset(p.h.p, 0, ns * sizeof(double2)));
    // wavenumber and dealias tables, with derive2csh's own expressions
    std::vector<double> kx(p.ld, 0.0), ky(ny);
    std::vector<int> imin(ny);
    const double pi_s = (double)3.1428f;  // real(8), parameter :: pi = 3.1428 (default-real literal)
    const double L = 2.0 * pi_s;
    for (int i = 0; i < p.nh; ++i) kx[i] = 2.0 * pi_s * (double)i / L;
    for (int j = 0; j < ny; ++j) ky[j] = (j < ny / 2) ? 2.0 * pi_s * (double)j / L : 2.0 * pi_s * ((double)j - ny) / L;
    const double cut = ((double)(nx + ny) / 2.0) / 3.0;
    for (int j = 0; j < ny; ++j) {
        int m = p.nh;
        for (int i = 0; i < p.nh; ++i)
            if (std::sqrt(kx[i] * kx[i] + ky[j] * ky[j]) >= cut) { m = i; break; }
        // monotone in i (kx increasing): everything from m on is cut
        imin[j] = m;
    }
    if ((rc = p.kx.alloc(p.ld * sizeof(double)))) return rc;
    if ((rc = p.ky.alloc(ny * sizeof(double)))) return rc;
    if ((rc = p.imin.alloc(ny * sizeof(int)))) return rc;
    TARA_HIP(hipMemcpy(p.kx.p, kx.data(), p.ld * sizeof(double), hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(p.ky.p, ky.data(), ny * sizeof(double), hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(p.imin.p, imin.data(), ny * sizeof(int), hipMemcpyHostToDevice));
    p.tw_y = twiddles(ny);
    p.tw_x = twiddles(nx / 2);
    p.rctw = rc_twiddles(nx);
    TARA_CHECK(p.tw_y && p.tw_x && p.rctw, TARA_ERR_ALLOC, "twiddle table allocation failed");
    return TARA_OK;
}

int csh_rhs_stage(const CshPlan& p, const double2* ws, int mode, double nu, double dt, double2* base, double2* aux,
                  double2* next, hipStream_t st, EventTimer* tm)
{
    int rc;
    if (tm) tm->begin(0, st);
    if ((rc = yinv(p, ws, st))) return rc;
    if (tm) tm->end(0, st);
    if (tm) tm->begin(1, st);
    if ((rc = xpass(p, st))) return rc;
    if (tm) tm->end(1, st);
    YfwdArgs a;
    a.h = p.h.as<double2>();
    a.ws = ws;
    a.base = base;
    a.aux = aux;
    a.next = next;
    a.kx = p.kx.as<double>();
    a.ky = p.ky.as<double>();
    a.imin = p.imin.as<int>();
    a.ld = p.ld;
    a.nu = nu;
    a.dt = dt;
    if (tm) tm->begin(2, st);
    if ((rc = yfwd(p, mode, a, st))) return rc;
    if (tm) tm->end(2, st);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

int launch_ab2(const double2* w, double2* wn, const double2* fo, const double2* fn, double dt, size_t n,
               hipStream_t st)
{
    const int bs = 256;
    size_t nb = (n + bs - 1) / bs;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_ab2, dim3((unsigned)nb), dim3(bs), 0, st, w, wn, fo, fn, dt, n);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

}  // namespace tara
