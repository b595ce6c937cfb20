// 2D compressible families on gfx950: the right-hand sides of
// src/derive2coh.f95 (isothermal hydro: n, ux, uy) and src/derive2com.f95
// (compressible MHD with forcing: rho, rho ux, rho uy, E, Bx, By), their AB2
// updates (src/ab2coh.f95, src/ab2com.f95) and the Euler-type steppers of the
// coh family.
//
// Device layout: kx-fastest, row padded to ld (multiple of 8), element (i, j)
// at j*ld + i.  Sweeps per right-hand side:
//   coh: 1 k_f2_prep   per tile of 8 kx columns: spectral multiply (1, i ky,
//                      ky^2) of n, ux, uy and the backward y FFT -> A (8 fields)
//        2 k_coh_x     per x row: the 13 c2r of derive2coh.f95:92-129 as 7
//                      paired complex FFTs (x derivatives i kx, kx^2 are applied
//                      here, where kx is constant along the completed y
//                      transform), the real-space terms of :131-167 combined
//                      into three sums, 2 paired forward FFTs -> B (3 fields)
//        3 k_coh_yfwd  forward y FFT, circular dealias on every term
//                      (:201-242), negate (:246-258), RHS out or AB2 in place.
//      The reference transforms the nine real-space terms separately and sums
//      the masked spectra (:246-252); the sums are formed before one transform
//      each, which is the same linear map.
//   com: 1 k_f2_prep   backward y FFT of the 6 state fields -> A
//        2 k_com_x1    rho, rho u -> u = rho u/rho and the forcing
//                      (derive2com.f95:88-117), forward x -> C (4 fields)
//        3 k_com_mid   forward y of u, F -> K (ux_k, uy_k, Fx_k, Fy_k), and the
//                      backward y of i kx ux_k, i ky uy_k -> D (:124-168)
//        4 k_com_x2    8 c2r (state + the two derivatives), the fluxes of
//                      :180-214, 7 distinct forward transforms -> B
//        5 k_com_yfwd  forward y, i k . flux, k^2 terms, dealias, forcing, assembly
//                      (:255-420), RHS out or AB2 in place.
#include <cmath>
#include <vector>

#include "colpass.h"
#include "fam2d.h"

namespace tara {

// spectral multipliers of the prep pass
enum : int { MUL_1 = 0, MUL_IKY = 1, MUL_KY2 = 2, MUL_IKX = 3, MUL_KX2 = 4 };

// (0, 1) k z as the reference forms it: (-(k z.y), k z.x)
__device__ __forceinline__ double2 f2_ik(double k, double2 z) { return make_double2(-(k * z.y), k * z.x); }
__device__ __forceinline__ double2 f2_k2(double k, double2 z)
{
    const double kk = k * k;
    return make_double2(kk * z.x, kk * z.y);
}
__device__ __forceinline__ double2 f2_mul(int kind, double kx, double ky, double2 z)
{
    switch (kind) {
        case MUL_IKY: return f2_ik(ky, z);
        case MUL_KY2: return f2_k2(ky, z);
        case MUL_IKX: return f2_ik(kx, z);
        case MUL_KX2: return f2_k2(kx, z);
        default: return z;
    }
}

// ------------------------------------------------------------ pass 1 (prep)
struct PrepArgs {
    const double2* src[8];
    double2* dst[8];
    int kind[8];
    const double* kx;
    const double* ky;
    int ld;
};

template <int NY>
__global__ __launch_bounds__(CShape<NY>::NT) void k_f2_prep(PrepArgs a, const double2* tw)
{
    using S = CShape<NY>;
    constexpr int E = RF<NY>::E;
    __shared__ double2 lds[8 * S::LP];
    const int c = threadIdx.x % 8, t = threadIdx.x / 8;
    const int i = xcd_remap(blockIdx.x, gridDim.x) * 8 + c;
    const int o = blockIdx.y;
    const double2* src = a.src[o] + i;
    const int kind = a.kind[o];
    const double kxv = a.kx[i];
    double2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int j = rf_first_idx<NY>(t, e);
        v[e] = f2_mul(kind, kxv, a.ky[j], src[(size_t)j * a.ld]);
    }
    rf_fft<NY, +1>(v, t, lds + c * S::LP, tw);
    double2* dst = a.dst[o] + i;
#pragma unroll
    for (int e = 0; e < E; ++e) dst[(size_t)rf_final_idx<NY>(t, e) * a.ld] = v[e];
}

// ------------------------------------------------------------ x-pass helpers
// One x row per group of T = NX/4 lanes inside a wavefront (4 complex elements
// per lane), so the exchanges between FFT passes need only wave-level ordering.
template <int NX>
struct FXShape {
    static constexpr int E = RF<NX, 4>::E;
    static constexpr int T = RF<NX, 4>::T;
    static_assert(64 % T == 0, "x pass rows must not straddle wavefronts");
    static constexpr int RPW = 64 / T;
    static constexpr int RPB = 4 * RPW;  // rows per 256-thread workgroup
    static constexpr int LP = RF<NX, 4>::LDS + 1;
};

// Two real c2r rows in one complex FFT: Y = X_a + i X_b over the full length
// (X[Nx-k] = conj X[k]), with the x multipliers ka / kb applied to the stored
// half spectra first and the imaginary parts at kx = 0 and Nx/2 dropped as
// FFTW's c2r does.  fb may be null (second row zero).  Lane values end at
// rf_final_idx positions: Re = row a, Im = row b (unnormalised).
template <int NX>
__device__ __forceinline__ void f2_c2r_pair(const double2* fa, int ka, const double2* fb, int kb, const double* kx,
                                            int t, double2* xl, const double2* tw, double2 (&v)[RF<NX, 4>::E])
{
    constexpr int M = NX / 2;
#pragma unroll
    for (int e = 0; e < RF<NX, 4>::E; ++e) {
        const int k = rf_first_idx<NX, 4>(t, e);
        const int kk = k <= M ? k : NX - k;
        const double kxv = kx[kk];
        double2 u = f2_mul(ka, kxv, 0.0, fa[kk]);
        double2 w = fb ? f2_mul(kb, kxv, 0.0, fb[kk]) : make_double2(0.0, 0.0);
        if (kk == 0 || kk == M) {
            u.y = 0.0;
            w.y = 0.0;
        }
        if (k > M) {
            u.y = -u.y;
            w.y = -w.y;
        }
        v[e] = make_double2(u.x - w.y, u.y + w.x);
    }
    rf_fft<NX, +1, 4, true>(v, t, xl, tw);
}

// Forward x transform of two real rows held at rf_final_idx positions
// (unnormalised r2c, FFTW's e^{-i}), split by the two-for-one identity
// A[k] = (Z[k] + conj Z[N-k])/2, B[k] = (Z[k] - conj Z[N-k])/(2i).
template <int NX>
__device__ __forceinline__ void f2_r2c_pair(const double (&ra)[RF<NX, 4>::E], const double (&rb)[RF<NX, 4>::E],
                                            double2* oa, double2* ob, bool live, int t, double2* xl,
                                            const double2* tw)
{
    constexpr int E = RF<NX, 4>::E, T = RF<NX, 4>::T, M = NX / 2;
    double2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = make_double2(ra[e], rb[e]);
    rf_final_to_first<NX, 4, true>(v, t, xl);
    rf_fft<NX, -1, 4, true>(v, t, xl, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) xl[rf_pad(rf_final_idx<NX, 4>(t, e))] = v[e];
    rf_sync<true>();
    if (live) {
        for (int k = t; k <= M; k += T) {
            const double2 z = xl[rf_pad(k)], zm = xl[rf_pad(k == 0 ? 0 : NX - k)];
            oa[k] = make_double2(0.5 * (z.x + zm.x), 0.5 * (z.y - zm.y));
            if (ob) ob[k] = make_double2(0.5 * (z.y + zm.y), -0.5 * (z.x - zm.x));
        }
    }
    rf_sync<true>();
}

struct F2XArgs {
    const double2* a;  // input fields, field stride fs
    const double2* d;  // com: derivative fields
    double2* b;        // output fields
    size_t fs;
    const double* kx;
    int ld, nrows, nx;
    double norm;       // 1 / (Nx Ny)
    double p0, p1, p2, p3, p4;  // family constants
};

#define F2X_PROLOGUE(NX)                                                                  \
    using S = FXShape<NX>;                                                                \
    constexpr int E = S::E;                                                               \
    __shared__ double2 lds[S::RPB * S::LP];                                               \
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;                           \
    const int t = lane % S::T;                                                            \
    const int slot = wave * S::RPW + lane / S::T;                                         \
    const int row = xcd_remap(blockIdx.x, gridDim.x) * S::RPB + slot;                     \
    const bool live = row < x.nrows;                                                      \
    const size_t ro = (size_t)(live ? row : x.nrows - 1) * x.ld;                          \
    double2* xl = lds + slot * S::LP;                                                     \
    (void)E;

// ------------------------------------------------------------------ coh x pass
// A: 0 n, 1 ux, 2 uy, 3 i ky n, 4 i ky ux, 5 i ky uy, 6 ky^2 ux, 7 ky^2 uy.
// p0 = mux, p1 = muy, p2 = CS^2.  B: 0 NLn, 1 NLx + visc + CS^2 Px, 2 the y twin.
template <int NX>
__global__ __launch_bounds__(256) void k_coh_x(F2XArgs x, const double2* tw)
{
    F2X_PROLOGUE(NX)
    auto F = [&](int f) { return x.a + (size_t)f * x.fs + ro; };
    double2 v[E];
    double n[E], ux[E], uy[E], dnx[E], duxx[E], duyx[E];
    double rn[E], rx[E], ry[E], vx[E], vy[E];
    f2_c2r_pair<NX>(F(0), MUL_1, F(0), MUL_IKX, x.kx, t, xl, tw, v);  // n, dn/dx
#pragma unroll
    for (int e = 0; e < E; ++e) {
        n[e] = v[e].x * x.norm;
        dnx[e] = v[e].y * x.norm;
    }
    f2_c2r_pair<NX>(F(1), MUL_1, F(1), MUL_IKX, x.kx, t, xl, tw, v);  // ux, dux/dx
#pragma unroll
    for (int e = 0; e < E; ++e) {
        ux[e] = v[e].x * x.norm;
        duxx[e] = v[e].y * x.norm;
    }
    f2_c2r_pair<NX>(F(2), MUL_1, F(2), MUL_IKX, x.kx, t, xl, tw, v);  // uy, duy/dx
#pragma unroll
    for (int e = 0; e < E; ++e) {
        uy[e] = v[e].x * x.norm;
        duyx[e] = v[e].y * x.norm;
        rn[e] = ux[e] * dnx[e];
        rx[e] = ux[e] * duxx[e];
        ry[e] = ux[e] * duyx[e];
    }
    f2_c2r_pair<NX>(F(1), MUL_KX2, F(2), MUL_KX2, x.kx, t, xl, tw, v);  // d2ux/dx2, d2uy/dx2 (kx^2 u)
#pragma unroll
    for (int e = 0; e < E; ++e) {
        vx[e] = x.p0 * ((v[e].x * x.norm) / n[e]);
        vy[e] = x.p0 * ((v[e].y * x.norm) / n[e]);
    }
    f2_c2r_pair<NX>(F(3), MUL_1, F(4), MUL_1, x.kx, t, xl, tw, v);  // dn/dy, dux/dy
    double py[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double dny = v[e].x * x.norm, duxy = v[e].y * x.norm;
        rn[e] = rn[e] + uy[e] * dny;
        rx[e] = rx[e] + uy[e] * duxy;
        py[e] = dny / n[e];
    }
    f2_c2r_pair<NX>(F(5), MUL_1, F(6), MUL_1, x.kx, t, xl, tw, v);  // duy/dy, d2ux/dy2
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double duyy = v[e].x * x.norm, d2uxy = v[e].y * x.norm;
        rn[e] = rn[e] + n[e] * (duxx[e] + duyy);
        ry[e] = ry[e] + uy[e] * duyy;
        vx[e] = vx[e] + x.p1 * (d2uxy / n[e]);
    }
    f2_c2r_pair<NX>(F(7), MUL_1, nullptr, MUL_1, x.kx, t, xl, tw, v);  // d2uy/dy2
#pragma unroll
    for (int e = 0; e < E; ++e) {
        vy[e] = vy[e] + x.p1 * ((v[e].x * x.norm) / n[e]);
        // NLx + (mux Visc_xx + muy Visc_xy) + CS^2 Px, derive2coh.f95:248-249
        rx[e] = (rx[e] + vx[e]) + x.p2 * (dnx[e] / n[e]);
        ry[e] = (ry[e] + vy[e]) + x.p2 * py[e];
    }
    double2* B0 = x.b + ro;
    f2_r2c_pair<NX>(rn, rx, B0, B0 + x.fs, live, t, xl, tw);
    f2_r2c_pair<NX>(ry, ry, B0 + 2 * x.fs, nullptr, live, t, xl, tw);
}

// ------------------------------------------------------------------ com x passes
// First round.  A: 0 rho, 1 rho ux, 2 rho uy.  p0 = A (forcing amplitude),
// p1 = kf, p2 = dx, p3 = dy.  C: 0 ux, 1 uy, 2 Fx, 3 Fy (x-transformed).
template <int NX>
__global__ __launch_bounds__(256) void k_com_x1(F2XArgs x, const double2* tw)
{
    F2X_PROLOGUE(NX)
    auto F = [&](int f) { return x.a + (size_t)f * x.fs + ro; };
    double2 v[E];
    double rho[E], ux[E], uy[E], fx[E], fy[E];
    f2_c2r_pair<NX>(F(0), MUL_1, F(1), MUL_1, x.kx, t, xl, tw, v);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        rho[e] = v[e].x * x.norm;
        ux[e] = (v[e].y * x.norm) / rho[e];
    }
    f2_c2r_pair<NX>(F(2), MUL_1, nullptr, MUL_1, x.kx, t, xl, tw, v);
    const int j = live ? row : 0;
    const double yj = (double)j * x.p3;
    const double sy = sin(x.p1 * yj);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        uy[e] = (v[e].x * x.norm) / rho[e];
        const double xi = (double)rf_final_idx<NX, 4>(t, e) * x.p2;
        // Fx = -rho A sin(kf y), Fy = +rho A sin(kf x)  (derive2com.f95:109-110)
        fx[e] = -(rho[e] * x.p0 * sy);
        fy[e] = rho[e] * x.p0 * sin(x.p1 * xi);
    }
    double2* C0 = x.b + ro;
    f2_r2c_pair<NX>(ux, uy, C0, C0 + x.fs, live, t, xl, tw);
    f2_r2c_pair<NX>(fx, fy, C0 + 2 * x.fs, C0 + 3 * x.fs, live, t, xl, tw);
}

// Second round.  A: 0 rho, 1 rho ux, 2 rho uy, 3 E, 4 Bx, 5 By; D: 0 i kx ux_k,
// 1 i ky uy_k (y-inverted).  p0 = CS^2, p1 = eta.  B: 0 Mom_x_1, 1 Mom_x_2
// (= Mom_y_1), 2 Mom_y_2, 3 Energy_x, 4 Energy_y, 5 E_Visc, 6 Mag_x (Mag_y = -Mag_x).
template <int NX>
__global__ __launch_bounds__(256) void k_com_x2(F2XArgs x, const double2* tw)
{
    F2X_PROLOGUE(NX)
    auto F = [&](int f) { return x.a + (size_t)f * x.fs + ro; };
    const double2* D0 = x.d + ro;
    double2 v[E];
    double rho[E], rux[E], ruy[E], en[E], bx[E], by[E], dudx[E], dvdy[E];
    f2_c2r_pair<NX>(F(0), MUL_1, F(1), MUL_1, x.kx, t, xl, tw, v);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        rho[e] = v[e].x * x.norm;
        rux[e] = v[e].y * x.norm;
    }
    f2_c2r_pair<NX>(F(2), MUL_1, F(3), MUL_1, x.kx, t, xl, tw, v);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        ruy[e] = v[e].x * x.norm;
        en[e] = v[e].y * x.norm;
    }
    f2_c2r_pair<NX>(F(4), MUL_1, F(5), MUL_1, x.kx, t, xl, tw, v);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        bx[e] = v[e].x * x.norm;
        by[e] = v[e].y * x.norm;
    }
    f2_c2r_pair<NX>(D0, MUL_1, D0 + x.fs, MUL_1, x.kx, t, xl, tw, v);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        dudx[e] = v[e].x * x.norm;
        dvdy[e] = v[e].y * x.norm;
    }
    double f0[E], f1[E];
    double2* B0 = x.b + ro;
    // momentum fluxes (derive2com.f95:196-199)
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double ux = rux[e] / rho[e], uy = ruy[e] / rho[e];
        const double b2 = bx[e] * bx[e] + by[e] * by[e];
        const double p = x.p0 * rho[e];
        f0[e] = ((rux[e] * ux + p) + b2 / 2.0) - bx[e] * bx[e];
        f1[e] = rux[e] * uy - bx[e] * by[e];
    }
    f2_r2c_pair<NX>(f0, f1, B0, B0 + x.fs, live, t, xl, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double ux = rux[e] / rho[e], uy = ruy[e] / rho[e];
        const double b2 = bx[e] * bx[e] + by[e] * by[e];
        const double p = x.p0 * rho[e];
        f0[e] = ((ruy[e] * uy + p) + b2 / 2.0) - by[e] * by[e];
        // The reference sets d_Bx_dy and d_By_dx from the already normalised
        // d_ux_dx and d_uy_dy, divided by Nx*Ny once more (derive2com.f95:186-187).
        const double dbxdy = dudx[e] * x.norm, dbydx = dvdy[e] * x.norm;
        const double curl = dbydx - dbxdy;
        const double ex = (((en[e] + p) + b2 / 2.0) * ux - (ux * bx[e]) * bx[e]) - (uy * bx[e]) * by[e];
        f1[e] = ex - (x.p1 * by[e]) * curl;
    }
    f2_r2c_pair<NX>(f0, f1, B0 + 2 * x.fs, B0 + 3 * x.fs, live, t, xl, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double ux = rux[e] / rho[e], uy = ruy[e] / rho[e];
        const double b2 = bx[e] * bx[e] + by[e] * by[e];
        const double p = x.p0 * rho[e];
        const double dbxdy = dudx[e] * x.norm, dbydx = dvdy[e] * x.norm;
        const double curl = dbydx - dbxdy;
        const double ey = (((en[e] + p) + b2 / 2.0) * uy - (ux * bx[e]) * by[e]) - (uy * by[e]) * by[e];
        f0[e] = ey + (x.p1 * bx[e]) * curl;
        const double dv = dudx[e] + dvdy[e];
        f1[e] = dv * dv;
    }
    f2_r2c_pair<NX>(f0, f1, B0 + 4 * x.fs, B0 + 5 * x.fs, live, t, xl, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double ux = rux[e] / rho[e], uy = ruy[e] / rho[e];
        f0[e] = uy * bx[e] - by[e] * ux;  // Mag_x (:208)
    }
    f2_r2c_pair<NX>(f0, f0, B0 + 6 * x.fs, nullptr, live, t, xl, tw);
}

// ------------------------------------------------------- y forward helpers
struct F2YArgs {
    const double2* b;   // x-pass output fields
    const double2* c;   // com: first-round x output (ux, uy, Fx, Fy)
    double2* d;         // com: derivative fields (out of k_com_mid)
    double2* k;         // com: ux_k, uy_k, Fx_k, Fy_k
    size_t fs;
    double2* st[6];     // state
    double2* out[6];    // RHS (F2_RHS) or history f_{n-1} -> f_n (F2_AB2)
    const double* kx;
    const double* ky;
    const int* imin;
    int ld;
    double dt, mu, eta;
    int* bad;           // set when an AB2 update stores a non-finite value
};

#define F2Y_PROLOGUE(NY)                                                      \
    using S = CShape<NY>;                                                     \
    constexpr int E = RF<NY>::E;                                              \
    __shared__ double2 lds[8 * S::LP];                                        \
    const int c = threadIdx.x % 8, t = threadIdx.x / 8;                       \
    const int i = xcd_remap(blockIdx.x, gridDim.x) * 8 + c;                   \
    double2* xl = lds + c * S::LP;                                            \
    const double kxv = a.kx[i];                                               \
    size_t q[E];                                                              \
    double kyv[E];                                                            \
    bool keep[E];                                                             \
    _Pragma("unroll") for (int e = 0; e < E; ++e)                             \
    {                                                                         \
        const int j = rf_final_idx<NY>(t, e);                                 \
        q[e] = (size_t)j * a.ld + i;                                          \
        kyv[e] = a.ky[j];                                                     \
        keep[e] = i < a.imin[j];                                              \
    }

template <int NY>
__device__ __forceinline__ void f2_yfft_fwd(const double2* src, int i, int ld, int t, double2* xl, const double2* tw,
                                            double2 (&v)[RF<NY>::E])
{
#pragma unroll
    for (int e = 0; e < RF<NY>::E; ++e) v[e] = src[(size_t)rf_first_idx<NY>(t, e) * ld + i];
    rf_fft<NY, -1>(v, t, xl, tw);
}

template <int MODE>
__device__ __forceinline__ void f2_finish(double2* st, double2* out, size_t q, double2 f, double dt, int* bad)
{
    if (MODE == F2_RHS) {
        out[q] = f;
    } else {
        const double2 fo = out[q], w = st[q];
        const double2 nw = make_double2(w.x + (1.5 * f.x - 0.5 * fo.x) * dt, w.y + (1.5 * f.y - 0.5 * fo.y) * dt);
        if (!(isfinite(nw.x) && isfinite(nw.y))) *bad = 1;
        st[q] = nw;
        out[q] = f;
    }
}

// ------------------------------------------------------------ coh y forward
template <int NY, int MODE>
__global__ __launch_bounds__(CShape<NY>::NT) void k_coh_yfwd(F2YArgs a, const double2* tw)
{
    F2Y_PROLOGUE(NY)
    (void)kxv;
    double2 v[E];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        f2_yfft_fwd<NY>(a.b + (size_t)f * a.fs, i, a.ld, t, xl, tw, v);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const double2 r = keep[e] ? v[e] : make_double2(0.0, 0.0);
            f2_finish<MODE>(a.st[f], a.out[f], q[e], make_double2(-r.x, -r.y), a.dt, a.bad);
        }
    }
}

// ------------------------------------------------------------ com middle pass
// Forward y of ux, uy, Fx, Fy (-> K, the reference's ux_k, uy_k, Fx_k, Fy_k),
// then the backward y of i kx ux_k and i ky uy_k (-> D).
template <int NY>
__global__ __launch_bounds__(CShape<NY>::NT) void k_com_mid(F2YArgs a, const double2* tw, const double2* tw_inv)
{
    F2Y_PROLOGUE(NY)
    (void)keep;
    double2 v[E];
#pragma unroll 1
    for (int f = 0; f < 4; ++f) {
        f2_yfft_fwd<NY>(a.c + (size_t)f * a.fs, i, a.ld, t, xl, tw, v);
        double2* kf = a.k + (size_t)f * a.fs;
#pragma unroll
        for (int e = 0; e < E; ++e) kf[q[e]] = v[e];
        if (f < 2) {
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = f2_ik(f == 0 ? kxv : kyv[e], v[e]);
            rf_final_to_first<NY>(v, t, xl);
            rf_fft<NY, +1>(v, t, xl, tw_inv);
            double2* df = a.d + (size_t)f * a.fs;
#pragma unroll
            for (int e = 0; e < E; ++e) df[(size_t)rf_final_idx<NY>(t, e) * a.ld + i] = v[e];
        }
    }
}

// ------------------------------------------------------------ com y forward
template <int NY, int MODE>
__global__ __launch_bounds__(CShape<NY>::NT) void k_com_yfwd(F2YArgs a, const double2* tw)
{
    F2Y_PROLOGUE(NY)
    double2 v[E], acc[E], acc2[E];
    auto msk = [&](int e, double2 z) { return keep[e] ? z : make_double2(0.0, 0.0); };
    // density: -(i kx rho ux + i ky rho uy), masked (derive2com.f95:267, 339, 402)
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double2 s1 = msk(e, f2_ik(kxv, a.st[1][q[e]])), s2 = msk(e, f2_ik(kyv[e], a.st[2][q[e]]));
        f2_finish<MODE>(a.st[0], a.out[0], q[e], make_double2(-(s1.x + s2.x), -(s1.y + s2.y)), a.dt, a.bad);
    }
    // rho ux: -(i kx Mom_x_1 + i ky Mom_x_2) - mu (kx^2 ux_k + ky^2 ux_k) + Fx_k
    f2_yfft_fwd<NY>(a.b, i, a.ld, t, xl, tw, v);  // Mom_x_1
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = msk(e, f2_ik(kxv, v[e]));
    f2_yfft_fwd<NY>(a.b + a.fs, i, a.ld, t, xl, tw, v);  // Mom_x_2 = Mom_y_1
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double2 d = msk(e, f2_ik(kyv[e], v[e]));
        acc2[e] = msk(e, f2_ik(kxv, v[e]));  // i kx Mom_y_1
        const double2 u = a.k[q[e]], fx = a.k[2 * a.fs + q[e]];
        const double2 k1 = msk(e, f2_k2(kxv, u)), k2 = msk(e, f2_k2(kyv[e], u));
        double2 r = make_double2(-(acc[e].x + d.x), -(acc[e].y + d.y));
        r = make_double2((r.x - a.mu * (k1.x + k2.x)) + fx.x, (r.y - a.mu * (k1.y + k2.y)) + fx.y);
        f2_finish<MODE>(a.st[1], a.out[1], q[e], r, a.dt, a.bad);
    }
    // rho uy: -(i kx Mom_y_1 + i ky Mom_y_2) - mu (kx^2 uy_k + ky^2 uy_k) + Fy_k
    f2_yfft_fwd<NY>(a.b + 2 * a.fs, i, a.ld, t, xl, tw, v);  // Mom_y_2
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double2 d = msk(e, f2_ik(kyv[e], v[e]));
        const double2 u = a.k[a.fs + q[e]], fy = a.k[3 * a.fs + q[e]];
        const double2 k1 = msk(e, f2_k2(kxv, u)), k2 = msk(e, f2_k2(kyv[e], u));
        double2 r = make_double2(-(acc2[e].x + d.x), -(acc2[e].y + d.y));
        r = make_double2((r.x - a.mu * (k1.x + k2.x)) + fy.x, (r.y - a.mu * (k1.y + k2.y)) + fy.y);
        f2_finish<MODE>(a.st[2], a.out[2], q[e], r, a.dt, a.bad);
    }
    // E: -(i kx Energy_x + i ky Energy_y) + mu E_Visc
    f2_yfft_fwd<NY>(a.b + 3 * a.fs, i, a.ld, t, xl, tw, v);
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = msk(e, f2_ik(kxv, v[e]));
    f2_yfft_fwd<NY>(a.b + 4 * a.fs, i, a.ld, t, xl, tw, v);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double2 d = msk(e, f2_ik(kyv[e], v[e]));
        acc[e] = make_double2(-(acc[e].x + d.x), -(acc[e].y + d.y));
    }
    f2_yfft_fwd<NY>(a.b + 5 * a.fs, i, a.ld, t, xl, tw, v);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double2 ev = msk(e, v[e]);
        f2_finish<MODE>(a.st[3], a.out[3], q[e], make_double2(acc[e].x + a.mu * ev.x, acc[e].y + a.mu * ev.y), a.dt, a.bad);
    }
    // Bx: -i ky Mag_x - eta (kx^2 + ky^2) Bx;  By: -i kx Mag_y - eta (...) By, Mag_y = -Mag_x
    f2_yfft_fwd<NY>(a.b + 6 * a.fs, i, a.ld, t, xl, tw, v);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double2 bxk = a.st[4][q[e]], byk = a.st[5][q[e]];
        const double2 mx = msk(e, f2_ik(kyv[e], v[e]));
        const double2 my = msk(e, f2_ik(kxv, make_double2(-v[e].x, -v[e].y)));
        const double2 b1 = msk(e, f2_k2(kxv, bxk)), b2 = msk(e, f2_k2(kyv[e], bxk));
        const double2 c1 = msk(e, f2_k2(kxv, byk)), c2 = msk(e, f2_k2(kyv[e], byk));
        const double2 rbx = make_double2(-mx.x - a.eta * (b1.x + b2.x), -mx.y - a.eta * (b1.y + b2.y));
        const double2 rby = make_double2(-my.x - a.eta * (c1.x + c2.x), -my.y - a.eta * (c1.y + c2.y));
        f2_finish<MODE>(a.st[4], a.out[4], q[e], rbx, a.dt, a.bad);
        f2_finish<MODE>(a.st[5], a.out[5], q[e], rby, a.dt, a.bad);
    }
}

// ------------------------------------------------------------ pointwise steppers
__global__ void k_f2_point(int kind, const double2* w, double2* __restrict__ wn, const double2* fo,
                           const double2* fn, double dt, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const double2 a = w[q], f = fn[q];
        double2 r;
        if (kind == F2P_AB2) {
            const double2 o = fo[q];
            r = make_double2(a.x + (1.5 * f.x - 0.5 * o.x) * dt, a.y + (1.5 * f.y - 0.5 * o.y) * dt);
        } else if (kind == F2P_EULER) {
            r = make_double2(a.x + f.x * dt, a.y + f.y * dt);
        } else if (kind == F2P_RK4_DEGEN) {
            // all four stage slopes are the input slope (src/rk4nk2coh.f95:22-56)
            const double c = dt / 6.0;
            r = make_double2(a.x + c * (((f.x + 2.0 * f.x) + 2.0 * f.x) + f.x),
                             a.y + c * (((f.y + 2.0 * f.y) + 2.0 * f.y) + f.y));
        } else {
            // predictor and corrector slopes are both the input slope (src/pcnk2coh.f95:20-33)
            r = make_double2(a.x + ((f.x + f.x) * dt) / 2.0, a.y + ((f.y + f.y) * dt) / 2.0);
        }
        wn[q] = r;
    }
}

// ------------------------------------------------------------------ dispatch
#define TARA_F2_SWITCH(n, FN, ...)                  \
    switch (n) {                                    \
        case 16: return FN<16>(__VA_ARGS__);        \
        case 32: return FN<32>(__VA_ARGS__);        \
        case 64: return FN<64>(__VA_ARGS__);        \
        case 128: return FN<128>(__VA_ARGS__);      \
        case 256: return FN<256>(__VA_ARGS__);      \
        default: set_error("2D compressible: unsupported length %d (16..256)", n); return TARA_ERR_UNSUPPORTED; \
    }

template <int NY>
static int launch_prep(const F2Plan& p, const PrepArgs& a, int nfields, hipStream_t st)
{
    hipLaunchKernelGGL(k_f2_prep<NY>, dim3(p.ld / 8, nfields), dim3(CShape<NY>::NT), 0, st, a, p.tw_y);
    return TARA_OK;
}
static int prep(const F2Plan& p, const PrepArgs& a, int nfields, hipStream_t st)
{
    TARA_F2_SWITCH(p.ny, launch_prep, p, a, nfields, st)
}

template <int NX>
static int launch_x(const F2Plan& p, int which, const F2XArgs& x, hipStream_t st)
{
    using S = FXShape<NX>;
    dim3 g((p.ny + S::RPB - 1) / S::RPB), b(256);
    if (which == 0)
        hipLaunchKernelGGL(k_coh_x<NX>, g, b, 0, st, x, p.tw_x);
    else if (which == 1)
        hipLaunchKernelGGL(k_com_x1<NX>, g, b, 0, st, x, p.tw_x);
    else
        hipLaunchKernelGGL(k_com_x2<NX>, g, b, 0, st, x, p.tw_x);
    return TARA_OK;
}
static int xpass(const F2Plan& p, int which, const F2XArgs& x, hipStream_t st)
{
    TARA_F2_SWITCH(p.nx, launch_x, p, which, x, st)
}

template <int NY>
static int launch_y(const F2Plan& p, int which, int mode, const F2YArgs& a, hipStream_t st)
{
    dim3 g(p.ld / 8), b(CShape<NY>::NT);
    if (which == 0) {
        if (mode == F2_RHS)
            hipLaunchKernelGGL((k_coh_yfwd<NY, F2_RHS>), g, b, 0, st, a, p.tw_y);
        else
            hipLaunchKernelGGL((k_coh_yfwd<NY, F2_AB2>), g, b, 0, st, a, p.tw_y);
    } else if (which == 1) {
        hipLaunchKernelGGL(k_com_mid<NY>, g, b, 0, st, a, p.tw_y, p.tw_y);
    } else {
        if (mode == F2_RHS)
            hipLaunchKernelGGL((k_com_yfwd<NY, F2_RHS>), g, b, 0, st, a, p.tw_y);
        else
            hipLaunchKernelGGL((k_com_yfwd<NY, F2_AB2>), g, b, 0, st, a, p.tw_y);
    }
    return TARA_OK;
}
static int ypass(const F2Plan& p, int which, int mode, const F2YArgs& a, hipStream_t st)
{
    TARA_F2_SWITCH(p.ny, launch_y, p, which, mode, a, st)
}

// ------------------------------------------------------------------ host side
int f2_plan_init(F2Plan& p, int family, int nx, int ny, double pi, double lx, double ly)
{
    TARA_CHECK(is_pow2(nx) && is_pow2(ny) && nx >= 16 && ny >= 16 && nx <= 256 && ny <= 256, TARA_ERR_UNSUPPORTED,
               "2D compressible: grid %dx%d unsupported (powers of two, 16..256)", nx, ny);
    TARA_CHECK(lx > 0 && ly > 0 && pi > 0, TARA_ERR_ARG, "2D compressible: pi and box lengths must be positive");
    p.nx = nx;
    p.ny = ny;
    p.nh = nx / 2 + 1;
    p.ld = (p.nh + 7) / 8 * 8;
    p.fsz = (size_t)p.ld * ny;
    p.pi = pi;
    p.lx = lx;
    p.ly = ly;
    const size_t fb = p.fsz * sizeof(double2);
    int rc;
    const int na = family == 0 ? 8 : 6, nb = family == 0 ? 3 : 7;
    if ((rc = p.a.alloc(na * fb)) || (rc = p.b.alloc(nb * fb))) return rc;
    TARA_HIP(hipMemset(p.a.p, 0, p.a.bytes));
    TARA_HIP(hipMemset(p.b.p, 0, p.b.bytes));
    if (family == 1) {
        if ((rc = p.c.alloc(4 * fb)) || (rc = p.d.alloc(2 * fb)) || (rc = p.k.alloc(4 * fb))) return rc;
        TARA_HIP(hipMemset(p.c.p, 0, p.c.bytes));
        TARA_HIP(hipMemset(p.d.p, 0, p.d.bytes));
        TARA_HIP(hipMemset(p.k.p, 0, p.k.bytes));
    }
    // kx = 2 pi (i-1)/Lx, ky = 2 pi (j-1)/Ly or 2 pi ((j-1)-Ny)/Ly (derive2coh.f95:59-60, 74-75)
    std::vector<double> kx(p.ld, 0.0), ky(ny);
    for (int i = 0; i < p.nh; ++i) kx[i] = 2.0 * pi * (double)i / lx;
    for (int j = 0; j < ny; ++j) ky[j] = 2.0 * pi * (double)(j < ny / 2 ? j : j - ny) / ly;
    // circular dealias: sqrt(kx^2+ky^2) >= (Nx+Ny)/2/3 (derive2coh.f95:208)
    const double cut = ((double)(nx + ny) / 2.0) / 3.0 + 0;
    std::vector<int> imin(ny);
    for (int j = 0; j < ny; ++j) {
        int m = p.nh;
        for (int i = 0; i < p.nh; ++i)
            if (std::sqrt(kx[i] * kx[i] + ky[j] * ky[j]) >= cut) {
                m = i;
                break;
            }
        imin[j] = m;  // kx increases with i: every i >= m is cut
    }
    if ((rc = p.kx.alloc(p.ld * sizeof(double))) || (rc = p.ky.alloc(ny * sizeof(double))) ||
        (rc = p.imin.alloc(ny * sizeof(int))) || (rc = p.bad.alloc(sizeof(int))))
        return rc;
    TARA_HIP(hipMemset(p.bad.p, 0, sizeof(int)));
    TARA_HIP(hipMemcpy(p.kx.p, kx.data(), p.ld * sizeof(double), hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(p.ky.p, ky.data(), ny * sizeof(double), hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(p.imin.p, imin.data(), ny * sizeof(int), hipMemcpyHostToDevice));
    p.tw_y = twiddles(ny);
    p.tw_x = twiddles(nx, 4);
    TARA_CHECK(p.tw_y && p.tw_x, TARA_ERR_ALLOC, "twiddle table allocation failed");
    return TARA_OK;
}

int coh_rhs_stage(const F2Plan& p, const CohPhys& ph, double2* const st[3], int mode, double dt,
                  double2* const out[3], hipStream_t s, EventTimer* tm)
{
    int rc;
    double2* A = p.a.as<double2>();
    PrepArgs pa{};
    const int src[8] = {0, 1, 2, 0, 1, 2, 1, 2};
    const int kind[8] = {MUL_1, MUL_1, MUL_1, MUL_IKY, MUL_IKY, MUL_IKY, MUL_KY2, MUL_KY2};
    for (int o = 0; o < 8; ++o) {
        pa.src[o] = st[src[o]];
        pa.dst[o] = A + o * p.fsz;
        pa.kind[o] = kind[o];
    }
    pa.kx = p.kx.as<double>();
    pa.ky = p.ky.as<double>();
    pa.ld = p.ld;
    if (tm) tm->begin(0, s);
    if ((rc = prep(p, pa, 8, s))) return rc;
    if (tm) tm->end(0, s);
    F2XArgs x{};
    x.a = A;
    x.b = p.b.as<double2>();
    x.fs = p.fsz;
    x.kx = p.kx.as<double>();
    x.ld = p.ld;
    x.nrows = p.ny;
    x.nx = p.nx;
    x.norm = 1.0 / ((double)p.nx * (double)p.ny);
    x.p0 = ph.mux;
    x.p1 = ph.muy;
    x.p2 = ph.cs * ph.cs;
    if (tm) tm->begin(1, s);
    if ((rc = xpass(p, 0, x, s))) return rc;
    if (tm) tm->end(1, s);
    F2YArgs y{};
    y.b = p.b.as<double2>();
    y.fs = p.fsz;
    for (int f = 0; f < 3; ++f) {
        y.st[f] = st[f];
        y.out[f] = out[f];
    }
    y.kx = p.kx.as<double>();
    y.ky = p.ky.as<double>();
    y.imin = p.imin.as<int>();
    y.ld = p.ld;
    y.dt = dt;
    y.bad = p.bad.as<int>();
    if (tm) tm->begin(2, s);
    if ((rc = ypass(p, 0, mode, y, s))) return rc;
    if (tm) tm->end(2, s);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

int com_rhs_stage(const F2Plan& p, const ComPhys& ph, double2* const st[6], int mode, double dt,
                  double2* const out[6], hipStream_t s, EventTimer* tm)
{
    int rc;
    double2* A = p.a.as<double2>();
    PrepArgs pa{};
    for (int o = 0; o < 6; ++o) {
        pa.src[o] = st[o];
        pa.dst[o] = A + o * p.fsz;
        pa.kind[o] = MUL_1;
    }
    pa.kx = p.kx.as<double>();
    pa.ky = p.ky.as<double>();
    pa.ld = p.ld;
    if (tm) tm->begin(0, s);
    if ((rc = prep(p, pa, 6, s))) return rc;
    if (tm) tm->end(0, s);
    F2XArgs x{};
    x.a = A;
    x.b = p.c.as<double2>();
    x.fs = p.fsz;
    x.kx = p.kx.as<double>();
    x.ld = p.ld;
    x.nrows = p.ny;
    x.nx = p.nx;
    x.norm = 1.0 / ((double)p.nx * (double)p.ny);
    x.p0 = ph.amp;
    x.p1 = ph.kf;
    x.p2 = p.lx / (double)p.nx;  // x(i) = (i-1) dx as src/main2com.f95:231 builds it
    x.p3 = p.ly / (double)p.ny;
    if (tm) tm->begin(1, s);
    if ((rc = xpass(p, 1, x, s))) return rc;
    F2YArgs y{};
    y.b = p.b.as<double2>();
    y.c = p.c.as<double2>();
    y.d = p.d.as<double2>();
    y.k = p.k.as<double2>();
    y.fs = p.fsz;
    for (int f = 0; f < 6; ++f) {
        y.st[f] = st[f];
        y.out[f] = out[f];
    }
    y.kx = p.kx.as<double>();
    y.ky = p.ky.as<double>();
    y.imin = p.imin.as<int>();
    y.ld = p.ld;
    y.dt = dt;
    y.mu = ph.mu;
    y.eta = ph.eta;
    y.bad = p.bad.as<int>();
    if ((rc = ypass(p, 1, mode, y, s))) return rc;
    x.b = p.b.as<double2>();
    x.d = p.d.as<double2>();
    x.p0 = ph.cs * ph.cs;
    x.p1 = ph.eta;
    if ((rc = xpass(p, 2, x, s))) return rc;
    if (tm) tm->end(1, s);
    if (tm) tm->begin(2, s);
    if ((rc = ypass(p, 2, mode, y, s))) return rc;
    if (tm) tm->end(2, s);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

int f2_pointwise(int kind, const double2* w, double2* wn, const double2* fo, const double2* fn, double dt, size_t n,
                 hipStream_t st)
{
    const int bs = 256;
    size_t nb = (n + bs - 1) / bs;
    if (nb > 65536) nb = 65536;
    if (nb == 0) return TARA_OK;
    hipLaunchKernelGGL(k_f2_point, dim3((unsigned)nb), dim3(bs), 0, st, kind, w, wn, fo, fn, dt, n);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

}  // namespace tara
