// 3D compressible isothermal MHD on gfx950: the right-hand side of
// src/derive3com.f95 and the AB2 update of src/ab3com.f95.
//
// Only the live dataflow of derive3com is computed.  ab3com freezes the
// energy (src/ab3com.f95:51) and the pressure is P = CS^2 rho
// (src/derive3com.f95:395), so E, curl B, the energy fluxes and the
// (commented-out) forcing never reach rho, rho u or B; they are not formed.
//
// Five sweeps per right-hand side (device layout: kx-fastest, row length
// padded to ld = Nh rounded up to 8; element (i, j, k) at (k*Ny + j)*ld + i):
//   1 k_col_fft (z, backward)  state (7 fields) -> A, tiles of 8 kx columns
//   2 k_col_fft (y, backward)  A in place
//   3 k_mhd_x                  per x row: 7 c2r as 4 paired complex FFTs,
//                              u = rho u / rho, P, B^2 and the 14 live fluxes
//                              (derive3com.f95:176-186, 395-436), 7 paired
//                              forward FFTs -> B (14 fields)
//   4 k_col_fft (y, forward)   B in place
//   5 k_mhd_zfwd               z forward of the 14 fields, i k . flux, k^2
//                              diffusion, spherical dealias on all terms
//                              (derive3com.f95:574-843), assembly
//                              (:1018-1037), then RHS out or AB2 in place.
// The diffusion coefficients are parameters (the reference hard-wires 1/450,
// derive3com.f95:1023-1037).
#include <cmath>

#include "mhd3d.h"
#include "colpass.h"

namespace tara {

// ------------------------------------------------------------------ x pass
// One x row per wavefront group of T = NX/4 lanes (4 complex elements per
// lane), so the exchanges between FFT passes need only wave-level ordering.
template <int NX>
struct MXShape {
    static constexpr int E = RF<NX, 4>::E;
    static constexpr int T = RF<NX, 4>::T;
    static_assert(64 % T == 0, "x pass rows must not straddle wavefronts");
    static constexpr int RPW = 64 / T;      // rows per wavefront
    static constexpr int RPB = 4 * RPW;     // rows per 256-thread workgroup
    static constexpr int LP = RF<NX, 4>::LDS + 1;
};

struct XArgs {
    const double2* a;  // 7 inverse-transformed fields, field stride fs
    double2* b;        // 14 flux fields
    size_t fs;
    int ld;
    double norm;       // 1 / (Nx Ny Nz)
    double cs2;
};

// At most 168 VGPRs (three waves per SIMD): faster than two waves at 178 VGPRs
// and than four with spills.
template <int NX>
__global__ __launch_bounds__(256, 3) void k_mhd_x(XArgs x, const double2* tw)
{
    using S = MXShape<NX>;
    constexpr int E = S::E, T = S::T, M = NX / 2;
    __shared__ double2 lds[S::RPB * S::LP];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int rw = lane / T, t = lane % T;
    const int slot = wave * S::RPW + rw;
    const size_t row = (size_t)xcd_remap(blockIdx.x, gridDim.x) * S::RPB + slot;  // = k*Ny + j
    const size_t ro = row * x.ld;
    double2* xl = lds + slot * S::LP;

    // backward: pairs (rho, rho ux), (rho uy, rho uz), (Bx, By), (Bz, -)
    double pr[7][E];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const double2* fa = x.a + (size_t)(2 * p) * x.fs + ro;
        const double2* fb = x.a + (size_t)(2 * p + 1) * x.fs + ro;
        double2 v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int k = rf_first_idx<NX, 4>(t, e);
            const int kk = k <= M ? k : NX - k;
            double2 u = fa[kk];
            double2 w = (p < 3) ? fb[kk] : make_double2(0.0, 0.0);
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
#pragma unroll
        for (int e = 0; e < E; ++e) {
            pr[2 * p][e] = v[e].x * x.norm;
            if (p < 3) pr[2 * p + 1][e] = v[e].y * x.norm;
        }
    }
    // real-space fluxes at this lane's positions (the reference's expressions)
    double ux[E], uy[E], uz[E], pp[E], hb2[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double rho = pr[0][e];
        ux[e] = pr[1][e] / rho;
        uy[e] = pr[2][e] / rho;
        uz[e] = pr[3][e] / rho;
        pp[e] = x.cs2 * rho;
        hb2[e] = (pr[4][e] * pr[4][e] + pr[5][e] * pr[5][e] + pr[6][e] * pr[6][e]) / 2.0;
    }
    // forward pairs: (ux, uy) (uz, MX1) (MX2, MX3) (MY2, MY3) (MZ1, MZ2) (MZ3, MG1) (MG2, MG3)
#pragma unroll 1
    for (int p = 0; p < 7; ++p) {
        double2 v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const double rux = pr[1][e], ruy = pr[2][e], ruz = pr[3][e];
            const double bx = pr[4][e], by = pr[5][e], bz = pr[6][e];
            double f0, f1;
            switch (p) {
                case 0: f0 = ux[e]; f1 = uy[e]; break;
                case 1: f0 = uz[e]; f1 = rux * ux[e] + pp[e] + hb2[e] - bx * bx; break;
                case 2: f0 = rux * uy[e] - bx * by; f1 = rux * uz[e] - bx * bz; break;
                case 3: f0 = ruy * uy[e] + pp[e] + hb2[e] - by * by; f1 = ruy * uz[e] - by * bz; break;
                case 4: f0 = ruz * ux[e] - bz * bx; f1 = ruz * uy[e] - bz * by; break;
                case 5: f0 = ruz * uz[e] + pp[e] + hb2[e] - bz * bz; f1 = ux[e] * by - uy[e] * bx; break;
                default: f0 = ux[e] * bz - uz[e] * bx; f1 = uy[e] * bz - uz[e] * by; break;
            }
            v[e] = make_double2(f0, f1);
        }
        rf_final_to_first<NX, 4, true>(v, t, xl);
        rf_fft<NX, -1, 4, true>(v, t, xl, tw);
#pragma unroll
        for (int e = 0; e < E; ++e) xl[rf_pad(rf_final_idx<NX, 4>(t, e))] = v[e];
        rf_sync<true>();
        double2* oa = x.b + (size_t)(2 * p) * x.fs + ro;
        double2* ob = x.b + (size_t)(2 * p + 1) * x.fs + ro;
        for (int k = t; k <= M; k += T) {
            const double2 z = xl[rf_pad(k)], zm = xl[rf_pad(k == 0 ? 0 : NX - k)];
            oa[k] = make_double2(0.5 * (z.x + zm.x), 0.5 * (z.y - zm.y));
            ob[k] = make_double2(0.5 * (z.y + zm.y), -0.5 * (z.x - zm.x));
        }
        rf_sync<true>();
    }
}

// ----------------------------------------------------- z forward + assembly
struct ZArgs {
    const double2* b;  // 14 flux fields after the y pass
    size_t fs;
    double2* st[7];    // state (rho, rho ux, rho uy, rho uz, Bx, By, Bz)
    double2* out[7];   // RHS (MHD_RHS) or history f_{n-1} -> f_n (MHD_AB2)
    const double* kx;
    const double* ky;
    const double* kz;
    const int* imin;   // per (k, j) row: modes with i >= imin are dealiased away
    int ny, ld;
    double mu, eta, dt;
    int* bad;          // AB2: set to 1 when a new state value is not finite
};

// The z pass holds 14 transforms' worth of accumulators; at 4 elements per
// thread and a bound of 4 waves per SIMD (128 VGPRs, two 512-thread workgroups
// per CU) it ran 1.48 ms at 256^3, against 1.85 ms at 8 elements per thread
// (256 VGPRs, one 256-thread workgroup per CU) and 1.57 ms unbounded.  The AB2
// form's opaque slot index took it to 1.44 ms (spill 84 -> 28 bytes per lane),
// forming the density term from the momentum loads of the updates to 1.38 ms
// (110 VGPRs, no spill, 3 fields fewer read).
constexpr int kZfwdEM = 4;
template <int NZ, int MODE, int EM>
__global__ __launch_bounds__((CShape<NZ, EM>::NT), 4) void k_mhd_zfwd(ZArgs a, const double2* tw)
{
    using S = CShape<NZ, EM>;
    constexpr int E = RF<NZ, EM>::E;
    __shared__ double2 lds[8 * S::LP];
    const int c = threadIdx.x % 8, t = threadIdx.x / 8;
    const int tpp = a.ld / 8;
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const int j = tile / tpp, i = (tile % tpp) * 8 + c;
    const size_t base = (size_t)j * a.ld + i, es = (size_t)a.ny * a.ld;
    double2* xl = lds + c * S::LP;
    const double kxv = a.kx[i], kyv = a.ky[j];
    bool keep[E];
    double kzv[E], k2[E];
    size_t q[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = rf_final_idx<NZ, EM>(t, e);
        q[e] = base + (size_t)k * es;
        kzv[e] = a.kz[k];
        k2[e] = (kxv * kxv + kyv * kyv) + kzv[e] * kzv[e];
        keep[e] = i < a.imin[(size_t)k * a.ny + j];
    }
    auto zfft = [&](int f, double2 (&v)[E]) {
        const double2* src = a.b + (size_t)f * a.fs + base;
        // AB2: the slot index is made opaque per transform so that the index
        // arithmetic is redone here rather than held across all 14 transforms
        // (110 VGPRs against 124); the RHS form spills more that way and keeps
        // it.
        const int tt = MODE == MHD_AB2 ? rf_opaque(t) : t;
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = src[(size_t)rf_first_idx<NZ, EM>(tt, e) * es];
        rf_fft<NZ, -1, EM>(v, tt, xl, tw);
    };
    // i k z = (-k z.y, k z.x)
    auto ik = [](double kk, double2 z) { return make_double2(-(kk * z.y), kk * z.x); };
    // AB2: dens accumulates the density term -(i kx rho ux + i ky rho uy +
    // i kz rho uz) from the momentum state that the updates of fields 1..3 load
    // anyway, so the momentum spectra are read once (not again for the density
    // term before the 14 transforms); density is finished after field 3.
    double2 dens[E];
    bool finite = true;  // AB2: every new state value of this lane is finite
    // wpre: the state of field fld at this lane's positions when the caller
    // has already loaded it (AB2), else null.
    auto finish = [&](int fld, const double2 (&acc)[E], const double2* wpre = nullptr) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const double2 f = keep[e] ? acc[e] : make_double2(0.0, 0.0);
            if (MODE == MHD_RHS) {
                a.out[fld][q[e]] = f;
            } else {
                const double2 fo = a.out[fld][q[e]], w = wpre ? wpre[e] : a.st[fld][q[e]];
                if (fld == 1) dens[e] = ik(kxv, w);
                if (fld == 2) { const double2 s2 = ik(kyv, w); dens[e] = make_double2(dens[e].x + s2.x, dens[e].y + s2.y); }
                if (fld == 3) { const double2 s3 = ik(kzv[e], w); dens[e] = make_double2(dens[e].x + s3.x, dens[e].y + s3.y); }
                const double2 wn = make_double2(w.x + (1.5 * f.x - 0.5 * fo.x) * a.dt, w.y + (1.5 * f.y - 0.5 * fo.y) * a.dt);
                a.st[fld][q[e]] = wn;
                a.out[fld][q[e]] = f;
                finite = finite && isfinite(wn.x) && isfinite(wn.y);
            }
        }
    };
    double2 v[E], acc[E], acc2[E];
    if (MODE == MHD_RHS) {
        // density: -(i kx rho ux + i ky rho uy + i kz rho uz) on the state spectra
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const double2 s1 = ik(kxv, a.st[1][q[e]]), s2 = ik(kyv, a.st[2][q[e]]), s3 = ik(kzv[e], a.st[3][q[e]]);
            acc[e] = make_double2(-((s1.x + s2.x) + s3.x), -((s1.y + s2.y) + s3.y));
        }
        finish(0, acc);
    }
    // rho ux: -(i k . (MX1, MX2, MX3)) - mu k^2 ux; MX2 also feeds rho uy (Mom_y_1 == Mom_x_2)
    zfft(0, v);
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = make_double2(-(a.mu * k2[e] * v[e].x), -(a.mu * k2[e] * v[e].y));
    zfft(3, v);
#pragma unroll
    for (int e = 0; e < E; ++e) { const double2 d = ik(kxv, v[e]); acc[e].x -= d.x; acc[e].y -= d.y; }
    zfft(4, v);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double2 d = ik(kyv, v[e]), g = ik(kxv, v[e]);
        acc[e].x -= d.x;
        acc[e].y -= d.y;
        acc2[e] = make_double2(-g.x, -g.y);
    }
    zfft(5, v);
#pragma unroll
    for (int e = 0; e < E; ++e) { const double2 d = ik(kzv[e], v[e]); acc[e].x -= d.x; acc[e].y -= d.y; }
    finish(1, acc);
    // rho uy: acc2 holds -i kx MY1; add -i ky MY2 - i kz MY3 - mu k^2 uy
    zfft(1, v);
#pragma unroll
    for (int e = 0; e < E; ++e) { acc2[e].x -= a.mu * k2[e] * v[e].x; acc2[e].y -= a.mu * k2[e] * v[e].y; }
    zfft(6, v);
#pragma unroll
    for (int e = 0; e < E; ++e) { const double2 d = ik(kyv, v[e]); acc2[e].x -= d.x; acc2[e].y -= d.y; }
    zfft(7, v);
#pragma unroll
    for (int e = 0; e < E; ++e) { const double2 d = ik(kzv[e], v[e]); acc2[e].x -= d.x; acc2[e].y -= d.y; }
    finish(2, acc2);
    // rho uz
    zfft(2, v);
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = make_double2(-(a.mu * k2[e] * v[e].x), -(a.mu * k2[e] * v[e].y));
    zfft(8, v);
#pragma unroll
    for (int e = 0; e < E; ++e) { const double2 d = ik(kxv, v[e]); acc[e].x -= d.x; acc[e].y -= d.y; }
    zfft(9, v);
#pragma unroll
    for (int e = 0; e < E; ++e) { const double2 d = ik(kyv, v[e]); acc[e].x -= d.x; acc[e].y -= d.y; }
    zfft(10, v);
#pragma unroll
    for (int e = 0; e < E; ++e) { const double2 d = ik(kzv[e], v[e]); acc[e].x -= d.x; acc[e].y -= d.y; }
    finish(3, acc);
    if (MODE == MHD_AB2) {
#pragma unroll
        for (int e = 0; e < E; ++e) acc[e] = make_double2(-dens[e].x, -dens[e].y);
        finish(0, acc);
    }
    // induction: dBx = i ky MG1 + i kz MG2, dBy = -i kx MG1 + i kz MG3,
    // dBz = -i kx MG2 - i ky MG3, each minus eta k^2 B (src/derive3com.f95:1030-1037)
    zfft(11, v);  // MG1
#pragma unroll
    for (int e = 0; e < E; ++e) {
        acc[e] = ik(kyv, v[e]);
        const double2 g = ik(kxv, v[e]);
        acc2[e] = make_double2(-g.x, -g.y);
    }
    zfft(12, v);  // MG2
    double2 acc3[E], bs[E];  // bs: the B component just loaded, reused by its update
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double2 d = ik(kzv[e], v[e]), g = ik(kxv, v[e]);
        bs[e] = a.st[4][q[e]];
        acc[e] = make_double2((acc[e].x + d.x) - a.eta * k2[e] * bs[e].x, (acc[e].y + d.y) - a.eta * k2[e] * bs[e].y);
        acc3[e] = make_double2(-g.x, -g.y);
    }
    finish(4, acc, bs);
    zfft(13, v);  // MG3
    double2 bz[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const double2 d = ik(kzv[e], v[e]), g = ik(kyv, v[e]);
        bs[e] = a.st[5][q[e]];
        bz[e] = a.st[6][q[e]];
        acc2[e] = make_double2((acc2[e].x + d.x) - a.eta * k2[e] * bs[e].x, (acc2[e].y + d.y) - a.eta * k2[e] * bs[e].y);
        acc3[e] = make_double2((acc3[e].x - g.x) - a.eta * k2[e] * bz[e].x, (acc3[e].y - g.y) - a.eta * k2[e] * bz[e].y);
    }
    finish(5, acc2, bs);
    finish(6, acc3, bz);
    if (MODE == MHD_AB2 && !finite && a.bad) a.bad[0] = 1;
}

// ------------------------------------------------------------------ dispatch
#define TARA_MHD_SWITCH(n, FN, ...)                 \
    switch (n) {                                    \
        case 16: return FN<16>(__VA_ARGS__);        \
        case 32: return FN<32>(__VA_ARGS__);        \
        case 64: return FN<64>(__VA_ARGS__);        \
        case 128: return FN<128>(__VA_ARGS__);      \
        case 256: return FN<256>(__VA_ARGS__);      \
        default: set_error("3D MHD: unsupported length %d (16..256)", n); return TARA_ERR_UNSUPPORTED; \
    }

template <int N>
static int launch_col(const ColArgs& a, int ntiles, int nfields, int dir, const double2* tw, hipStream_t st)
{
    dim3 g(ntiles, nfields), b(CShape<N>::NT);
    if (dir < 0)
        hipLaunchKernelGGL((k_col_fft<N, -1>), g, b, 0, st, a, tw);
    else
        hipLaunchKernelGGL((k_col_fft<N, +1>), g, b, 0, st, a, tw);
    return TARA_OK;
}
int col_fft_tiles(int n, const ColArgs& a, int ntiles, int nfields, int dir, const double2* tw, hipStream_t st)
{
    TARA_MHD_SWITCH(n, launch_col, a, ntiles, nfields, dir, tw, st)
}
static int col(int n, const ColArgs& a, int ntiles, int nfields, int dir, const double2* tw, hipStream_t st)
{
    return col_fft_tiles(n, a, ntiles, nfields, dir, tw, st);
}

template <int NX>
static int launch_x(const MhdPlan& p, const XArgs& x, hipStream_t st)
{
    using S = MXShape<NX>;
    const int nrows = p.ny * p.nz;
    TARA_CHECK(nrows % S::RPB == 0, TARA_ERR_UNSUPPORTED, "3D MHD: Ny*Nz=%d not a multiple of %d", nrows, S::RPB);
    hipLaunchKernelGGL(k_mhd_x<NX>, dim3(nrows / S::RPB), dim3(256), 0, st, x, p.tw_x);
    return TARA_OK;
}
static int xpass(const MhdPlan& p, const XArgs& x, hipStream_t st) { TARA_MHD_SWITCH(p.nx, launch_x, p, x, st) }

template <int NZ>
static int launch_z(const MhdPlan& p, int mode, const ZArgs& z, hipStream_t st)
{
    constexpr int EM = kZfwdEM;
    dim3 g(p.ny * (p.ld / 8)), b(CShape<NZ, EM>::NT);
    const double2* tw = EM == 4 ? p.tw_z4 : p.tw_z;
    if (mode == MHD_RHS)
        hipLaunchKernelGGL((k_mhd_zfwd<NZ, MHD_RHS, EM>), g, b, 0, st, z, tw);
    else
        hipLaunchKernelGGL((k_mhd_zfwd<NZ, MHD_AB2, EM>), g, b, 0, st, z, tw);
    return TARA_OK;
}
static int zpass(const MhdPlan& p, int mode, const ZArgs& z, hipStream_t st)
{
    TARA_MHD_SWITCH(p.nz, launch_z, p, mode, z, st)
}

// ------------------------------------------------------------------ host side
int mhd_plan_init(MhdPlan& p, int nx, int ny, int nz, double lx, double ly, double lz, double mu, double eta,
                  double cs)
{
    TARA_CHECK(is_pow2(nx) && is_pow2(ny) && is_pow2(nz) && nx >= 16 && ny >= 16 && nz >= 16 && nx <= 256 &&
                   ny <= 256 && nz <= 256,
               TARA_ERR_UNSUPPORTED, "3D MHD: grid %dx%dx%d unsupported (powers of two, 16..256)", nx, ny, nz);
    TARA_CHECK(lx > 0 && ly > 0 && lz > 0, TARA_ERR_ARG, "3D MHD: box lengths must be positive");
    p.nx = nx;
    p.ny = ny;
    p.nz = nz;
    p.nh = nx / 2 + 1;
    p.ld = (p.nh + 7) / 8 * 8;
    p.lx = lx;
    p.ly = ly;
    p.lz = lz;
    p.mu = mu;
    p.eta = eta;
    p.cs = cs;
    p.fsz = (size_t)p.ld * ny * nz;
    int rc;
    if ((rc = p.a.alloc(7 * p.fsz * sizeof(double2))) || (rc = p.b.alloc(14 * p.fsz * sizeof(double2)))) return rc;
    TARA_HIP(hipMemset(p.a.p, 0, p.a.bytes));
    TARA_HIP(hipMemset(p.b.p, 0, p.b.bytes));
    // wavenumbers as derive3com forms them: 2 pi (i-1)/L with true pi
    // (src/derive3com.f95:245-247, 260-262), second halves of y and z negative
    const double pi = 3.14159265358979323846;
    std::vector<double> kx(p.ld, 0.0), ky(ny), kz(nz);
    for (int i = 0; i < p.nh; ++i) kx[i] = 2.0 * pi * (double)i / lx;
    for (int j = 0; j < ny; ++j) ky[j] = 2.0 * pi * (double)(j < ny / 2 ? j : j - ny) / ly;
    for (int k = 0; k < nz; ++k) kz[k] = 2.0 * pi * (double)(k < nz / 2 ? k : k - nz) / lz;
    // spherical dealias of derive3com.f95:802: sqrt(kx^2+ky^2+kz^2) >= (Nx+Ny+Nz)/3/3
    const double cut = ((double)(nx + ny + nz) / 3.0) / 3.0 + 0;
    std::vector<int> imin((size_t)ny * nz);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) {
            int m = p.nh;
            for (int i = 0; i < p.nh; ++i)
                if (std::sqrt(kx[i] * kx[i] + ky[j] * ky[j] + kz[k] * kz[k]) >= cut) {
                    m = i;
                    break;
                }
            imin[(size_t)k * ny + j] = m;  // kx increases with i: every i >= m is cut
        }
    if ((rc = p.kx.alloc(p.ld * sizeof(double))) || (rc = p.ky.alloc(ny * sizeof(double))) ||
        (rc = p.kz.alloc(nz * sizeof(double))) || (rc = p.imin.alloc(imin.size() * sizeof(int))) ||
        (rc = p.bad.alloc(sizeof(int))))
        return rc;
    TARA_HIP(hipMemset(p.bad.p, 0, sizeof(int)));
    TARA_HIP(hipMemcpy(p.kx.p, kx.data(), p.ld * sizeof(double), hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(p.ky.p, ky.data(), ny * sizeof(double), hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(p.kz.p, kz.data(), nz * sizeof(double), hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(p.imin.p, imin.data(), imin.size() * sizeof(int), hipMemcpyHostToDevice));
    p.tw_y = twiddles(ny);
    p.tw_z = twiddles(nz);
    p.tw_z4 = twiddles(nz, 4);
    p.tw_x = twiddles(nx, 4);
    TARA_CHECK(p.tw_y && p.tw_z && p.tw_z4 && p.tw_x, TARA_ERR_ALLOC, "twiddle table allocation failed");
    return TARA_OK;
}

int mhd_rhs_stage(const MhdPlan& p, double2* const st[7], int mode, double dt, double2* const out[7],
                  double2* const hist[7], hipStream_t s, EventTimer* tm)
{
    int rc;
    double2* A = p.a.as<double2>();
    double2* B = p.b.as<double2>();
    const int tpp = p.ld / 8;
    ColArgs c{};
    // 1: z backward, state -> A
    for (int f = 0; f < 7; ++f) {
        c.src[f] = st[f];
        c.dst[f] = A + f * p.fsz;
    }
    c.es = (long)p.ny * p.ld;
    c.pstride = p.ld;
    c.tpp = tpp;
    if (tm) tm->begin(0, s);
    if ((rc = col(p.nz, c, p.ny * tpp, 7, +1, p.tw_z, s))) return rc;
    // 2: y backward in place
    for (int f = 0; f < 7; ++f) c.src[f] = c.dst[f] = A + f * p.fsz;
    c.es = p.ld;
    c.pstride = (long)p.ny * p.ld;
    if ((rc = col(p.ny, c, p.nz * tpp, 7, +1, p.tw_y, s))) return rc;
    if (tm) tm->end(0, s);
    // 3: x pass
    XArgs x{A, B, p.fsz, p.ld, 1.0 / ((double)p.nx * (double)p.ny * (double)p.nz), p.cs * p.cs};
    if (tm) tm->begin(1, s);
    if ((rc = xpass(p, x, s))) return rc;
    if (tm) tm->end(1, s);
    // 4: y forward in place on the 14 flux fields
    for (int f = 0; f < 14; ++f) c.src[f] = c.dst[f] = B + f * p.fsz;
    if (tm) tm->begin(2, s);
    if ((rc = col(p.ny, c, p.nz * tpp, 14, -1, p.tw_y, s))) return rc;
    if (tm) tm->end(2, s);
    // 5: z forward + assembly + update
    ZArgs z{};
    z.b = B;
    z.fs = p.fsz;
    for (int f = 0; f < 7; ++f) {
        z.st[f] = st[f];
        z.out[f] = mode == MHD_RHS ? out[f] : hist[f];
    }
    z.kx = p.kx.as<double>();
    z.ky = p.ky.as<double>();
    z.kz = p.kz.as<double>();
    z.imin = p.imin.as<int>();
    z.ny = p.ny;
    z.ld = p.ld;
    z.mu = p.mu;
    z.eta = p.eta;
    z.dt = dt;
    z.bad = p.bad.as<int>();
    if (tm) tm->begin(3, s);
    if ((rc = zpass(p, mode, z, s))) return rc;
    if (tm) tm->end(3, s);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

}  // namespace tara
