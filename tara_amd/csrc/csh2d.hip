// 2D incompressible hydro in vorticity-streamfunction form on gfx950:
// the right-hand side of src/derive2csh.f95 and the AB2 / RK4 updates
// (src/ab2csh.f95; classic RK4), as three fused HBM sweeps per RHS.
//
// Device layout: the state half spectra are stored ky-major ("transposed"),
// element (kx = i, ky = j) at i*Ny + j, so a kx column (one y transform) is a
// contiguous Ny-element run.  The intermediates are stored in row groups
// interleaved per kx: G0..G3 in 2-row (Nx = 4096) or 4-row groups (blkg), H in
// 8-row groups (blk8, element (i, j) at ((j/8)*Nh + i)*8 + j%8): a y pass reads
// or writes a column as 32- to 128-byte runs, and an x-pass workgroup of fewer
// rows than a group reads pieces of those runs whose other pieces belong to the
// neighbouring workgroups, which the XCD-aware mapping runs next to it.  The
// reference's (Nh, Ny) kx-fastest arrays are converted on upload / download
// (layout.hip).
//
//   pass 1  k_csh_yinv : per kx column read omega_hat, form the four spectra of
//                        derive2csh.f95:29-71 (i kx omega, i ky omega, i ky psi,
//                        -i kx psi; psi = omega/k^2, psi(0,0) = 0) and run the
//                        backward y FFTs -> G0..G3.
//   pass 2  k_csh_x    : per RB rows, four x c2r (half-length packing),
//                        normalise by Nx*Ny and form ux*d_x omega + uy*d_y omega
//                        (derive2csh.f95:89-98), one x r2c -> H.  (The reference
//                        transforms the two products separately (:100-106); the
//                        sum is transformed once, which is the same linear map.)
//   pass 3  k_csh_yfwd : per kx column, forward y FFT of H, circular dealias
//                        sqrt(kx^2+ky^2) >= (Nx+Ny)/6 on the nonlinear part only
//                        (:108-125), add nu k^2 omega (:127-139), negate
//                        (:141-145), then the fused update (RHS / AB2 / RK4 stage).
//
// Wavenumbers: derive2csh uses pi = 3.1428 (a default-real literal) and
// Lx = Ly = 2 pi (:16-21), so kx = i-1 and ky = j-1 or j-1-Ny are exact integers;
// the host builds the tables with that exact expression (csh_plan_init).
#include "common.h"
#include "csh2d.h"
#include "passes.h"
#include "rfft.h"

namespace tara {

// The register-resident transform of the y passes (csh_yem elements per thread).
template <int NY>
using RY = RF<NY, csh_yem(NY)>;

// Columns per workgroup of the y passes: at least 256 threads.
template <int N>
struct YShape {
    static constexpr int T = RY<N>::T;
    static constexpr int CB = T >= 256 ? 1 : 256 / T;
    static constexpr int NT = CB * T;
};
// Offset of element (kx = i, ky = j) in an 8-row-blocked intermediate.
__device__ __forceinline__ size_t blk8(int i, int j, int nh)
{
    return ((size_t)(j >> 3) * nh + i) * 8 + (j & 7);
}

// Offset of element (i, j) in the intermediates G0..G3: groups of 2^gs rows
// interleaved per kx, ld columns per group (CshPlan::gsh, CshPlan::gld).
// Ny = 4096, 2-row groups: a wave of the x pass reading two rows reads
// contiguous 32-byte pairs over consecutive kx, and the inverse y pass writes
// its two interleaved columns' pairs together (YiShape); 0.195 ms inverse y +
// 0.276 ms x at 4096^2, against 0.186 + 0.311 ms with 4-row groups.  Smaller
// Ny, 4-row groups and contiguous columns: at 1024^2 0.0201 ms inverse y +
// 0.0186 ms x, against 0.0236 + 0.0179 ms with 2-row groups and four
// interleaved columns.  (With one column per inverse-y workgroup at 4096^2,
// 4-row groups had measured faster than 8-row (0.20, 0.39 ms), 2-row (0.25,
// 0.31) and plain row-major (0.34, 0.31).)
__device__ __forceinline__ size_t blkg(int i, int j, int ld, int gs)
{
    return ((size_t)(j >> gs) * ld + i) * (1 << gs) + (j & ((1 << gs) - 1));
}

// derive2csh's ky: j-1 for j <= Ny/2, j-1-Ny above (exact integers, checked on the host).
template <int NY>
__device__ __forceinline__ double csh_ky(int j)
{
    return (double)(j < NY / 2 ? j : j - NY);
}

// --------------------------------------------------------------------- pass 1
// Shape of the inverse y pass.  At Ny = 4096 (512 threads per column) two
// columns share a workgroup with their lanes interleaved (column = lane parity):
// a wave's stores then cover both columns of each 2-row group of G0..G3
// together (64 B), where one column per workgroup wrote 32-byte pieces.  The
// second column's LDS buffer is offset by 64 B so the interleaved lanes use
// other banks.  Other lengths, and the screened form (psi kept live, more
// registers), keep the contiguous shape of the y passes.
template <int NY, bool SCR>
struct YiShape {
    static constexpr int T = RY<NY>::T;
    static constexpr bool IL = csh_yinv_interleaved(NY, SCR);
    static constexpr int CB = csh_yinv_cols(NY, SCR);
    static_assert(IL || CB == YShape<NY>::CB, "contiguous shape differs from the y passes'");
    static constexpr int NT = CB * T;
    static constexpr int LDSC = RY<NY>::LDS + (IL ? 4 : 0);  // per-column exchange buffer
    // Ny <= 512: at least 56 KB of LDS per workgroup, which holds the pass to
    // two 256-thread workgroups per CU.  Its registers would allow three, and
    // at three the ensemble step (128-512 grids) measured 6.67 against 6.42 ms.
    static constexpr int LDS_ELEMS = (!IL && NY <= 512 && CB * LDSC < 56 * 1024 / 16) ? 56 * 1024 / 16 : CB * LDSC;
};

// The inverse y pass of one kx column i given omega_hat (wv, at rf_first_idx
// positions): forms the four spectra of derive2csh.f95:29-71 (i kx omega,
// i ky omega, i ky psi, -i kx psi; psi = omega/k^2, psi(0,0) = 0), or of
// derive2css.f95:22-52 when SCR (screened), runs the backward y FFTs and writes
// G0..G3 (row-group stride ld).  x is the column's LDS exchange buffer.
template <int NY, bool SCR, int GS, bool OPQ>
__device__ __forceinline__ void csh_yinv_column(double2 (&wv)[RY<NY>::E], int i, bool live, int t, double2* x,
                                                const double2* tw, double lam, int ld,
                                                double2* __restrict__ g0, double2* __restrict__ g1,
                                                double2* __restrict__ g2, double2* __restrict__ g3)
{
    constexpr int E = RY<NY>::E;
    const double kxv = (double)i;  // derive2csh's kx = i-1 exactly (checked on the host)
    // screened: psi = omega/(k^2 + lambda) with psi(0,0) = 0, omega_d = k^2 psi;
    // the derivatives below act on omega_d, psi is kept in ps.
    double2 ps[SCR ? E : 1];
    if constexpr (SCR) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int j = rf_first_idx<NY, csh_yem(NY)>(t, e);
            const double kyv = csh_ky<NY>(j);
            const double k2 = kxv * kxv + kyv * kyv;
            const double den = k2 + lam;
            ps[e] = (i == 0 && j == 0) ? make_double2(0.0, 0.0) : make_double2(wv[e].x / den, wv[e].y / den);
            wv[e] = make_double2(k2 * ps[e].x, k2 * ps[e].y);
        }
    }
    // i kx omega, i ky omega first; then omega -> psi in place; then +i ky psi, -i kx psi.
#pragma unroll 1
    for (int f = 0; f < 4; ++f) {
        double2 v[E];
        // OPQ: the slot index is made opaque per transform so that the index
        // arithmetic of the transform and the stores is recomputed here rather
        // than hoisted out of the loop and held in registers (178 -> 117 VGPRs
        // at Ny = 4096, which lets a 1024-thread workgroup run unspilled).
        const int tt = OPQ ? rf_opaque(t) : t;
        if (f == 2) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                if constexpr (SCR) {
                    wv[e] = ps[e];
                } else {
                    const double kyv = csh_ky<NY>(rf_first_idx<NY, csh_yem(NY)>(tt, e));
                    const double k2 = kxv * kxv + kyv * kyv;
                    wv[e] = (k2 != 0.0) ? make_double2(wv[e].x / k2, wv[e].y / k2) : make_double2(0.0, 0.0);
                }
            }
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const double kyv = csh_ky<NY>(rf_first_idx<NY, csh_yem(NY)>(tt, e));
            const double kk = (f == 0 || f == 3) ? kxv : kyv;
            const double2 z = wv[e];
            v[e] = (f == 3) ? make_double2(kk * z.y, -(kk * z.x))     // -i kx psi
                            : make_double2(-(kk * z.y), kk * z.x);    // i k (omega | psi)
        }
        rf_fft<NY, +1, csh_yem(NY)>(v, tt, x, tw);
        // G0 = ux (f 2), G1 = uy (f 3), G2 = d omega/dx (f 0), G3 = d omega/dy (f 1)
        double2* o = f == 0 ? g2 : f == 1 ? g3 : f == 2 ? g0 : g1;
        if (live) {
#pragma unroll
            for (int e = 0; e < E; ++e) o[blkg(i, rf_final_idx<NY, csh_yem(NY)>(tt, e), ld, GS)] = v[e];
        }
    }
}

template <int NY, bool SCR, int GS>
__global__ __launch_bounds__((YiShape<NY, SCR>::NT), (YiShape<NY, SCR>::IL ? 4 : 1)) void k_csh_yinv(const double2* w, int nh,
                                                                    double2* __restrict__ g0, double2* __restrict__ g1,
                                                                    double2* __restrict__ g2, double2* __restrict__ g3,
                                                                    const double2* tw, double lam, int ld, size_t bs)
{
    // batched runs: blockIdx.y selects the run, every field is strided by bs
    const size_t rb = blockIdx.y * bs;
    w += rb;
    g0 += rb;
    g1 += rb;
    g2 += rb;
    g3 += rb;
    using S = YiShape<NY, SCR>;
    constexpr int E = RY<NY>::E, T = S::T;
    __shared__ double2 lds[S::LDS_ELEMS];
    const int c = S::IL ? (int)threadIdx.x % S::CB : (int)threadIdx.x / T;
    const int t = S::IL ? (int)threadIdx.x / S::CB : (int)threadIdx.x % T;
    const int i = xcd_remap(blockIdx.x, gridDim.x) * S::CB + c;
    const bool live = i < nh;
    const size_t col = (size_t)(live ? i : 0) * NY;
    double2 wv[E];
#pragma unroll
    for (int e = 0; e < E; ++e) wv[e] = live ? w[col + rf_first_idx<NY, csh_yem(NY)>(t, e)] : make_double2(0.0, 0.0);
    csh_yinv_column<NY, SCR, GS, S::IL>(wv, i, live, t, lds + c * S::LDSC, tw, lam, ld, g0, g1, g2, g3);
}

// --------------------------------------------------------------------- pass 2
// Two real c2r transforms in one complex FFT: with X_a, X_b the Hermitian half
// spectra (kx = 0..M) of two real rows, Y[k] = X_a[k] + i X_b[k] over the full
// length Nx (X[Nx-k] = conj X[k]) transforms to y = x_a + i x_b, so the product
// x_a * x_b is Re(y) Im(y).  The imaginary parts of X[0] and X[M] are dropped,
// as FFTW's c2r does.  Registers end with y at rf_final_idx<NX> positions.
template <int NX, int EM, bool SPLIT>
__device__ __forceinline__ void x_c2r_pair(const double2* ga, const double2* gb, int j, int t,
                                           double2* xl, const double2* tw, double2 (&v)[RF<NX, EM>::E])
{
    constexpr int M = NX / 2;
    constexpr int LD = csh_gld(M + 1), GS = csh_gshift(M + 1);
#pragma unroll
    for (int e = 0; e < RF<NX, EM>::E; ++e) {
        const int k = rf_first_idx<NX, EM>(t, e);
        const int kk = k <= M ? k : NX - k;
        double2 a = ga[blkg(kk, j, LD, GS)], b = gb[blkg(kk, j, LD, GS)];
        if (kk == 0 || kk == M) {
            a.y = 0.0;
            b.y = 0.0;
        }
        if (k > M) {
            a.y = -a.y;
            b.y = -b.y;
        }
        v[e] = make_double2(a.x - b.y, a.y + b.x);
    }
    rf_fft<NX, +1, EM, false, SPLIT>(v, t, xl, tw);
}

// Shape of the x pass: EM complex elements per thread (4, 8, 16: radix-4, -8,
// -16 passes), RB rows per workgroup, SPLIT: the FFT exchanges go through LDS as
// doubles in two rounds (half the LDS, so that two workgroups share a CU).
template <int NX, int EM, int RB, bool SPLIT>
struct XShape {
    static constexpr int M = NX / 2;
    static constexpr int T = RF<NX, EM>::T;    // threads per row (backward transforms, length NX)
    static constexpr int TM = RF<M, EM>::T;    // of which active in the forward half-length transform
    static constexpr int NT = RB * T;
    // Per-row exchange buffer (complex elements).  Rows are interleaved across
    // lanes, so successive row buffers are offset by 4 elements (64 B) to land
    // on different LDS banks.
    static constexpr int LDSR = (SPLIT ? RF<NX, EM>::LDS / 2 : RF<NX, EM>::LDS) + 4;
    // the forward half-length transform must use exactly half of the row's threads
    static constexpr bool OK = RB * LDSR * 16 <= 160 * 1024 && NT <= 1024 && NT >= 64 && 2 * TM == T;
};

template <int NX, int EM, int RB, bool SPLIT>
__global__ __launch_bounds__((XShape<NX, EM, RB, SPLIT>::NT)) void k_csh_x(
    const double2* g0, const double2* g1, const double2* g2, const double2* g3, double scale,
    double2* __restrict__ h, const double2* tw_n, const double2* tw_m, const double2* rctw, size_t bs)
{
    const size_t rb = blockIdx.y * bs;  // batched runs
    g0 += rb;
    g1 += rb;
    g2 += rb;
    g3 += rb;
    h += rb;
    using S = XShape<NX, EM, RB, SPLIT>;
    static_assert(S::OK, "x pass shape does not fit");
    constexpr int E = RF<NX, EM>::E, EH = RF<S::M, EM>::E, T = S::T, M = S::M;
    __shared__ double2 lds[RB * S::LDSR];
    const int r = threadIdx.x % RB, t = threadIdx.x / RB;  // row fastest: RB-row chunks per kx
    const int j = xcd_remap(blockIdx.x, gridDim.x) * RB + r;
    double2* xl = lds + r * S::LDSR;
    double nl[E];
    double2 v[E];
    x_c2r_pair<NX, EM, SPLIT>(g0, g2, j, t, xl, tw_n, v);   // ux + i d omega/dx
#pragma unroll
    for (int e = 0; e < E; ++e) nl[e] = v[e].x * v[e].y;
    x_c2r_pair<NX, EM, SPLIT>(g1, g3, j, t, xl, tw_n, v);   // uy + i d omega/dy
#pragma unroll
    for (int e = 0; e < E; ++e) nl[e] = (nl[e] + v[e].x * v[e].y) * scale;
    // forward x transform of the real product row: pack z[m] = nl[2m] + i nl[2m+1]
    // for the half-length FFT (threads t < TM carry it; the others only meet its
    // barriers where they fill whole waves, else run the same code on the
    // unused upper half of the row buffer so every barrier matches).
    double* xr = reinterpret_cast<double*>(xl);
#pragma unroll
    for (int e = 0; e < E; ++e) xr[rf_final_idx<NX, EM>(t, e)] = nl[e];
    __syncthreads();
    double2 z[EH];
    const int tm = t % S::TM;
#pragma unroll
    for (int e = 0; e < EH; ++e) {
        const int m = rf_first_idx<M, EM>(tm, e);
        z[e] = make_double2(xr[2 * m], xr[2 * m + 1]);
    }
    // The unpacking roots of this thread's kx = t + u T, fetched now, before
    // any store (loads and stores share one in-order counter: a root loaded
    // between the stores would also wait for the stores before it).  4096^2
    // x pass 0.274 against 0.278 ms.
    constexpr int NK = M / T;
    double2 wk[NK];
#pragma unroll
    for (int u = 0; u < NK; ++u) wk[u] = rctw[t + u * T];
    __syncthreads();
    if constexpr ((S::TM * RB) % 64 == 0) {
        // the threads t >= TM (threadIdx.x >= TM RB) fill whole waves: they meet
        // the transform's barriers without running it
        if (t < S::TM)
            rf_fft<M, -1, EM, false, SPLIT>(z, tm, xl, tw_m);
        else
            rf_fft_barriers_only<M, EM, SPLIT>();
    } else {
        constexpr int HALF = SPLIT ? RF<M, EM>::LDS / 2 : RF<M, EM>::LDS;  // upper half, complex elements
        double2* zl = xl + (t < S::TM ? 0 : HALF);
        rf_fft<M, -1, EM, false, SPLIT>(z, tm, zl, tw_m);
    }
    if (t < S::TM) {
#pragma unroll
        for (int e = 0; e < EH; ++e) xl[rf_pad(rf_final_idx<M, EM>(tm, e))] = z[e];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NK; ++u) {  // consecutive t -> consecutive kx
        const int k = t + u * T;
        const double2 zk = xl[rf_pad(k)];
        const double2 zmk = xl[rf_pad(k == 0 ? 0 : M - k)];
        double2 xm;
        h[blk8(k, j, M + 1)] = r2c_unpack_w(zk, zmk, k, wk[u], &xm);
        if (k == 0) h[blk8(M, j, M + 1)] = xm;
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
    int nh;
    double nu, dt;
    double alpha, lam;     // screened form: alpha * NL, viscous term on k^2 ws/(k^2 + lam)
    int screened;
    size_t bs;             // batched runs: field stride per run (blockIdx.y)
    const double* nus;     // batched runs: viscosity per run (null: nu)
    int* bad;              // per run: set to 1 when a new state value is not finite (AB2, RK4 stage 4)
};

template <int NY, int MODE>
__global__ __launch_bounds__(YShape<NY>::NT) void k_csh_yfwd(YfwdArgs a, const double2* tw)
{
    const size_t rb = blockIdx.y * a.bs;
    a.h += rb;
    a.ws += rb;
    if (a.base) a.base += rb;
    if (a.aux) a.aux += rb;
    if (a.next) a.next += rb;
    if (a.nus) a.nu = a.nus[blockIdx.y];
    using S = YShape<NY>;
    constexpr int E = RY<NY>::E, T = S::T;
    __shared__ double2 lds[S::CB * RY<NY>::LDS];
    const int c = threadIdx.x / T, t = threadIdx.x % T;
    const int i = xcd_remap(blockIdx.x, gridDim.x) * S::CB + c;
    const bool live = i < a.nh;
    const size_t col = (size_t)(live ? i : 0) * NY;
    double2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = live ? a.h[blk8(i, rf_first_idx<NY, csh_yem(NY)>(t, e), a.nh)] : make_double2(0.0, 0.0);
    rf_fft<NY, -1, csh_yem(NY)>(v, t, lds + c * RY<NY>::LDS, tw);
    if (!live) return;
    const double kxv = (double)i;
    bool finite = true;  // the new state's values (AB2, RK4 stage 4)
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int j = rf_final_idx<NY, csh_yem(NY)>(t, e);
        const size_t q = col + j;
        double2 nl = v[e];
        if (i >= a.imin[j]) nl = make_double2(0.0, 0.0);
        const double kyv = csh_ky<NY>(j);
        const double k2 = kxv * kxv + kyv * kyv;
        const double vk = a.nu * k2;
        const double2 ws = a.ws[q];
        double2 f;
        if (a.screened) {
            // derive2css.f95:118-130: -(alpha (NLx + NLy) + nu k^2 omega_d)
            const double den = k2 + a.lam;
            const double2 wd = (i == 0 && j == 0) ? make_double2(0.0, 0.0)
                                                  : make_double2(k2 * (ws.x / den), k2 * (ws.y / den));
            f = make_double2(-(a.alpha * nl.x + vk * wd.x), -(a.alpha * nl.y + vk * wd.y));
        } else {
            f = make_double2(-(nl.x + vk * ws.x), -(nl.y + vk * ws.y));
        }
        if (MODE == CSH_RHS) {
            a.next[q] = f;
        } else if (MODE == CSH_AB2) {
            const double2 fo = a.aux[q];
            const double2 wn = make_double2(ws.x + (1.5 * f.x - 0.5 * fo.x) * a.dt, ws.y + (1.5 * f.y - 0.5 * fo.y) * a.dt);
            a.base[q] = wn;
            a.aux[q] = f;
            finite = finite && isfinite(wn.x) && isfinite(wn.y);
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
            const double2 wn = make_double2(acc.x + c6 * f.x, acc.y + c6 * f.y);
            a.base[q] = wn;
            finite = finite && isfinite(wn.x) && isfinite(wn.y);
        }
    }
    if (!finite && a.bad) a.bad[blockIdx.y] = 1;
}

// ------------------------------------------------------- passes 3 + 1 fused
// The forward y pass of one stage followed by the inverse y pass of the stage
// that comes next, per kx column (derive2csh only).  The update leaves the next
// stage's input (RK4: the stage field; AB2 and the last RK4 stage: the new
// state) in registers, and the four spectra of it are transformed and written
// to G0..G3 without reading it back: one field read less per stage.  Shape and
// LDS as the inverse y pass; the forward y pass of H uses the same lanes.
template <int N, int EM>
constexpr bool rf_final_is_first()
{
    return fft_radix(N, RF<N, EM>::E) == RF<N, EM>::E && N / rf_last_span(N, RF<N, EM>::E) == RF<N, EM>::E;
}

template <int NY, int MODE, int GS>
__global__ __launch_bounds__((YiShape<NY, false>::NT), (YiShape<NY, false>::IL ? 4 : 1)) void k_csh_yfwd_inv(
    YfwdArgs a, const double2* tw, double2* __restrict__ g0, double2* __restrict__ g1, double2* __restrict__ g2,
    double2* __restrict__ g3, int ld)
{
    static_assert(MODE != CSH_RHS, "the fused pass needs a stage update");
    const size_t rb = blockIdx.y * a.bs;
    a.h += rb;
    a.ws += rb;
    a.base += rb;
    a.aux += rb;
    if (a.next) a.next += rb;
    g0 += rb;
    g1 += rb;
    g2 += rb;
    g3 += rb;
    if (a.nus) a.nu = a.nus[blockIdx.y];
    using S = YiShape<NY, false>;
    constexpr int E = RY<NY>::E, T = S::T;
    __shared__ double2 lds[S::LDS_ELEMS];
    const int c = S::IL ? (int)threadIdx.x % S::CB : (int)threadIdx.x / T;
    const int t = S::IL ? (int)threadIdx.x / S::CB : (int)threadIdx.x % T;
    const int i = xcd_remap(blockIdx.x, gridDim.x) * S::CB + c;
    const bool live = i < a.nh;
    const size_t col = (size_t)(live ? i : 0) * NY;
    double2* xl = lds + c * S::LDSC;
    double2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = live ? a.h[blk8(i, rf_first_idx<NY, csh_yem(NY)>(t, e), a.nh)] : make_double2(0.0, 0.0);
    rf_fft<NY, -1, csh_yem(NY)>(v, t, xl, tw);
    const double kxv = (double)i;
    bool finite = true;  // the new state's values (AB2, RK4 stage 4)
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int j = rf_final_idx<NY, csh_yem(NY)>(t, e);
        const size_t q = col + j;
        double2 nl = v[e];
        if (!live || i >= a.imin[j]) nl = make_double2(0.0, 0.0);
        const double kyv = csh_ky<NY>(j);
        const double vk = a.nu * (kxv * kxv + kyv * kyv);
        const double2 ws = live ? a.ws[q] : make_double2(0.0, 0.0);
        const double2 f = make_double2(-(nl.x + vk * ws.x), -(nl.y + vk * ws.y));
        double2 wn;  // the next stage's input
        if (MODE == CSH_AB2) {
            const double2 fo = live ? a.aux[q] : make_double2(0.0, 0.0);
            wn = make_double2(ws.x + (1.5 * f.x - 0.5 * fo.x) * a.dt, ws.y + (1.5 * f.y - 0.5 * fo.y) * a.dt);
            if (live) {
                a.base[q] = wn;
                a.aux[q] = f;
            }
            finite = finite && isfinite(wn.x) && isfinite(wn.y);
        } else if (MODE == CSH_RK4_1) {
            const double c6 = a.dt / 6.0, c2 = 0.5 * a.dt;
            wn = make_double2(ws.x + c2 * f.x, ws.y + c2 * f.y);
            if (live) {
                a.aux[q] = make_double2(ws.x + c6 * f.x, ws.y + c6 * f.y);
                a.next[q] = wn;
            }
        } else if (MODE == CSH_RK4_2 || MODE == CSH_RK4_3) {
            const double c3 = a.dt / 3.0, cn = (MODE == CSH_RK4_2) ? 0.5 * a.dt : a.dt;
            const double2 w0 = live ? a.base[q] : make_double2(0.0, 0.0);
            const double2 acc = live ? a.aux[q] : make_double2(0.0, 0.0);
            wn = make_double2(w0.x + cn * f.x, w0.y + cn * f.y);
            if (live) {
                a.aux[q] = make_double2(acc.x + c3 * f.x, acc.y + c3 * f.y);
                a.next[q] = wn;
            }
        } else {  // CSH_RK4_4
            const double c6 = a.dt / 6.0;
            const double2 acc = live ? a.aux[q] : make_double2(0.0, 0.0);
            wn = make_double2(acc.x + c6 * f.x, acc.y + c6 * f.y);
            if (live) a.base[q] = wn;
            finite = finite && isfinite(wn.x) && isfinite(wn.y);
        }
        v[e] = wn;
    }
    if (!finite && a.bad) a.bad[blockIdx.y] = 1;
    if constexpr (!rf_final_is_first<NY, csh_yem(NY)>()) rf_final_to_first<NY, csh_yem(NY)>(v, t, xl);
    csh_yinv_column<NY, false, GS, S::IL>(v, i, live, t, xl, tw, 0.0, ld, g0, g1, g2, g3);
}

// ------------------------------------------------------------ pointwise AB2
__global__ void k_ab2(const double2* w, double2* __restrict__ wn, const double2* fo,
                      const double2* fn, double dt, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const double2 a = w[q], o = fo[q], f = fn[q];
        wn[q] = make_double2(a.x + (1.5 * f.x - 0.5 * o.x) * dt, a.y + (1.5 * f.y - 0.5 * o.y) * dt);
    }
}

// ------------------------------------------------------------------ dispatch
template <int NY, bool SCR, int GS>
static int launch_yinv_t(const CshPlan& p, const double2* w, const CshPhys& ph, hipStream_t st)
{
    using S = YiShape<NY, SCR>;
    hipLaunchKernelGGL((k_csh_yinv<NY, SCR, GS>), dim3((p.nh + S::CB - 1) / S::CB, p.batch), dim3(S::NT), 0, st, w,
                       p.nh, p.g[0].as<double2>(), p.g[1].as<double2>(), p.g[2].as<double2>(), p.g[3].as<double2>(),
                       p.tw_y, ph.lambda, p.gld, p.bstride);
    return TARA_OK;
}

template <int NY>
static int launch_yinv(const CshPlan& p, const double2* w, const CshPhys& ph, hipStream_t st)
{
    if (p.gsh == 1)
        return ph.screened ? launch_yinv_t<NY, true, 1>(p, w, ph, st) : launch_yinv_t<NY, false, 1>(p, w, ph, st);
    return ph.screened ? launch_yinv_t<NY, true, 2>(p, w, ph, st) : launch_yinv_t<NY, false, 2>(p, w, ph, st);
}

// x-pass shapes the autotuner chooses among: {EM, RB, SPLIT}.  Shape 0 is
// RB = min(8, 1024 / T) at EM = 8 without splitting (the largest workgroup).
struct XCand {
    int em, rb, split;
};
constexpr XCand kXCands[] = {{8, 0, 0}, {16, 2, 1}, {8, 1, 1}, {16, 4, 1}, {8, 2, 1}, {16, 1, 0},
                              {4, 1, 0}, {4, 2, 0}, {4, 4, 1}};
constexpr int kNumXCands = sizeof(kXCands) / sizeof(kXCands[0]);

template <int NX, int EM>
constexpr int default_rb()
{
    return (1024 / RF<NX, EM>::T) < 1 ? 1 : (1024 / RF<NX, EM>::T) > 8 ? 8 : (1024 / RF<NX, EM>::T);
}

template <int NX, int EM, int RB, bool SPLIT>
static int launch_x_shape(const CshPlan& p, hipStream_t st)
{
    using S = XShape<NX, EM, RB, SPLIT>;
    if constexpr (!S::OK) {
        set_error("x pass shape unsupported at Nx=%d", NX);
        return TARA_ERR_UNSUPPORTED;
    } else {
        TARA_CHECK(p.ny % RB == 0, TARA_ERR_UNSUPPORTED, "Ny=%d not a multiple of %d", p.ny, RB);
        const double scale = 1.0 / ((double)p.nx * (double)p.ny * (double)p.nx * (double)p.ny);
        const double2* twn = EM == 16 ? p.tw_xn16 : EM == 4 ? p.tw_xn4 : p.tw_xn;
        const double2* twm = EM == 16 ? p.tw_x16 : EM == 4 ? p.tw_x4 : p.tw_x;
        TARA_CHECK(twn && twm, TARA_ERR_UNSUPPORTED, "x pass shape without its twiddle tables");
        hipLaunchKernelGGL((k_csh_x<NX, EM, RB, SPLIT>), dim3(p.ny / RB, p.batch), dim3(S::NT), 0, st,
                           p.g[0].as<double2>(), p.g[1].as<double2>(), p.g[2].as<double2>(), p.g[3].as<double2>(),
                           scale, p.h.as<double2>(), twn, twm, p.rctw, p.bstride);
        return TARA_OK;
    }
}

template <int NX>
static int launch_x(const CshPlan& p, hipStream_t st)
{
    switch (p.xcfg) {
        case 0: return launch_x_shape<NX, 8, default_rb<NX, 8>(), false>(p, st);
        case 1: return launch_x_shape<NX, 16, 2, true>(p, st);
        case 2: return launch_x_shape<NX, 8, 1, true>(p, st);
        case 3: return launch_x_shape<NX, 16, 4, true>(p, st);
        case 4: return launch_x_shape<NX, 8, 2, true>(p, st);
        case 5: return launch_x_shape<NX, 16, 1, false>(p, st);
        case 6: return launch_x_shape<NX, 4, 1, false>(p, st);
        case 7: return launch_x_shape<NX, 4, 2, false>(p, st);
        case 8: return launch_x_shape<NX, 4, 4, true>(p, st);
        default: set_error("x pass shape %d unknown", p.xcfg); return TARA_ERR_ARG;
    }
}

template <int NX>
static bool x_shape_ok(int c, int ny)
{
    switch (c) {
        case 0: return XShape<NX, 8, default_rb<NX, 8>(), false>::OK && ny % default_rb<NX, 8>() == 0;
        case 1: return XShape<NX, 16, 2, true>::OK && ny % 2 == 0;
        case 2: return XShape<NX, 8, 1, true>::OK;
        case 3: return XShape<NX, 16, 4, true>::OK && ny % 4 == 0;
        case 4: return XShape<NX, 8, 2, true>::OK && ny % 2 == 0;
        case 5: return XShape<NX, 16, 1, false>::OK;
        case 6: return XShape<NX, 4, 1, false>::OK;
        case 7: return XShape<NX, 4, 2, false>::OK && ny % 2 == 0;
        case 8: return XShape<NX, 4, 4, true>::OK && ny % 4 == 0;
        default: return false;
    }
}

template <int NY>
static int launch_yfwd(const CshPlan& p, int mode, const YfwdArgs& a, hipStream_t st)
{
    using S = YShape<NY>;
    dim3 g((p.nh + S::CB - 1) / S::CB, p.batch), b(S::NT);
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

#define TARA_CSH_SWITCH(n, FN, ...)                 \
    switch (n) {                                    \
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

static int yinv(const CshPlan& p, const double2* w, const CshPhys& ph, hipStream_t st)
{
    TARA_CSH_SWITCH(p.ny, launch_yinv, p, w, ph, st)
}
static int xpass(const CshPlan& p, hipStream_t st) { TARA_CSH_SWITCH(p.nx, launch_x, p, st) }
template <int NX>
static int shape_ok_n(int c, int ny)
{
    return x_shape_ok<NX>(c, ny) ? 1 : 0;
}
static int shape_ok(const CshPlan& p, int c) { TARA_CSH_SWITCH(p.nx, shape_ok_n, c, p.ny) }
template <int NY, int GS>
static int launch_yfwd_inv_g(const CshPlan& p, int mode, const YfwdArgs& a, hipStream_t st)
{
    using S = YiShape<NY, false>;
    const dim3 g((p.nh + S::CB - 1) / S::CB, p.batch), b(S::NT);
    double2* const g0 = p.g[0].as<double2>();
    double2* const g1 = p.g[1].as<double2>();
    double2* const g2 = p.g[2].as<double2>();
    double2* const g3 = p.g[3].as<double2>();
    switch (mode) {
        case CSH_AB2: hipLaunchKernelGGL((k_csh_yfwd_inv<NY, CSH_AB2, GS>), g, b, 0, st, a, p.tw_y, g0, g1, g2, g3, p.gld); break;
        case CSH_RK4_1: hipLaunchKernelGGL((k_csh_yfwd_inv<NY, CSH_RK4_1, GS>), g, b, 0, st, a, p.tw_y, g0, g1, g2, g3, p.gld); break;
        case CSH_RK4_2: hipLaunchKernelGGL((k_csh_yfwd_inv<NY, CSH_RK4_2, GS>), g, b, 0, st, a, p.tw_y, g0, g1, g2, g3, p.gld); break;
        case CSH_RK4_3: hipLaunchKernelGGL((k_csh_yfwd_inv<NY, CSH_RK4_3, GS>), g, b, 0, st, a, p.tw_y, g0, g1, g2, g3, p.gld); break;
        case CSH_RK4_4: hipLaunchKernelGGL((k_csh_yfwd_inv<NY, CSH_RK4_4, GS>), g, b, 0, st, a, p.tw_y, g0, g1, g2, g3, p.gld); break;
        default: set_error("fused y pass: mode %d has no stage update", mode); return TARA_ERR_ARG;
    }
    return TARA_OK;
}

template <int NY>
static int launch_yfwd_inv(const CshPlan& p, int mode, const YfwdArgs& a, hipStream_t st)
{
    return p.gsh == 1 ? launch_yfwd_inv_g<NY, 1>(p, mode, a, st) : launch_yfwd_inv_g<NY, 2>(p, mode, a, st);
}

static int yfwd_inv(const CshPlan& p, int mode, const YfwdArgs& a, hipStream_t st)
{
    TARA_CSH_SWITCH(p.ny, launch_yfwd_inv, p, mode, a, st)
}

static int yfwd(const CshPlan& p, int mode, const YfwdArgs& a, hipStream_t st)
{
    TARA_CSH_SWITCH(p.ny, launch_yfwd, p, mode, a, st)
}

// ------------------------------------------------------------------ host side
int csh_host_tables(int nx, int ny, std::vector<double>& kx, std::vector<double>& ky, std::vector<int>& imin)
{
    // wavenumber and dealias tables, with derive2csh's own expressions
    const int nh = nx / 2 + 1;
    kx.resize(nh);
    ky.resize(ny);
    imin.resize(ny);
    const double pi_s = (double)3.1428f;  // real(8), parameter :: pi = 3.1428 (default-real literal)
    const double L = 2.0 * pi_s;
    for (int i = 0; i < nh; ++i) kx[i] = 2.0 * pi_s * (double)i / L;
    for (int j = 0; j < ny; ++j) ky[j] = (j < ny / 2) ? 2.0 * pi_s * (double)j / L : 2.0 * pi_s * ((double)j - ny) / L;
    for (int i = 0; i < nh; ++i)
        TARA_CHECK(kx[i] == (double)i, TARA_ERR_UNSUPPORTED, "derive2csh kx(%d) is not an integer", i);
    for (int j = 0; j < ny; ++j)
        TARA_CHECK(ky[j] == (double)(j < ny / 2 ? j : j - ny), TARA_ERR_UNSUPPORTED, "derive2csh ky(%d) not an integer", j);
    const double cut = ((double)(nx + ny) / 2.0) / 3.0;
    for (int j = 0; j < ny; ++j) {
        int m = nh;
        for (int i = 0; i < nh; ++i)
            if (std::sqrt(kx[i] * kx[i] + ky[j] * ky[j]) >= cut) {
                m = i;
                break;
            }
        imin[j] = m;  // kx increases with i, so every i >= m is cut as well
    }
    return TARA_OK;
}

int csh_plan_init(CshPlan& p, int nx, int ny, int batch)
{
    TARA_CHECK(batch >= 1 && batch <= 65535, TARA_ERR_ARG, "batch %d out of range", batch);
    TARA_CHECK(is_pow2(nx) && is_pow2(ny) && nx >= 32 && ny >= 16 && nx <= 4096 && ny <= 4096,
               TARA_ERR_UNSUPPORTED, "derive2csh: grid %dx%d unsupported (powers of two, 32..4096 x 16..4096)",
               nx, ny);
    p.nx = nx;
    p.ny = ny;
    p.nh = nx / 2 + 1;
    p.gld = csh_gld(p.nh);
    p.gsh = csh_gshift(p.nh);
    const size_t ns = (size_t)p.gld * ny;  // G0..G3 hold gld columns per row group
    p.batch = batch;
    p.bstride = ns;
    int rc;
    for (int f = 0; f < 4; ++f) {
        if ((rc = p.g[f].alloc(batch * ns * sizeof(double2)))) return rc;
        TARA_HIP(hipMemset(p.g[f].p, 0, batch * ns * sizeof(double2)));
    }
    if ((rc = p.h.alloc(batch * ns * sizeof(double2)))) return rc;
    if ((rc = p.bad.alloc(batch * sizeof(int)))) return rc;
    TARA_HIP(hipMemset(p.bad.p, 0, batch * sizeof(int)));
    std::vector<double> kx, ky;
    std::vector<int> imin;
    if ((rc = csh_host_tables(nx, ny, kx, ky, imin))) return rc;
    if ((rc = p.kx.alloc(p.nh * sizeof(double)))) return rc;
    if ((rc = p.ky.alloc(ny * sizeof(double)))) return rc;
    if ((rc = p.imin.alloc(ny * sizeof(int)))) return rc;
    TARA_HIP(hipMemcpy(p.kx.p, kx.data(), p.nh * sizeof(double), hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(p.ky.p, ky.data(), ny * sizeof(double), hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(p.imin.p, imin.data(), ny * sizeof(int), hipMemcpyHostToDevice));
    p.tw_y = twiddles(ny, csh_yem(ny));
    p.tw_x = twiddles(nx / 2);
    p.tw_xn = twiddles(nx);
    p.tw_x16 = twiddles(nx / 2, 16);
    p.tw_xn16 = twiddles(nx, 16);
    p.tw_x4 = twiddles(nx / 2, 4);
    p.tw_xn4 = twiddles(nx, 4);
    p.rctw = rc_twiddles(nx);
    TARA_CHECK(p.tw_y && p.tw_x && p.tw_xn && p.tw_x16 && p.tw_xn16 && p.tw_x4 && p.tw_xn4 && p.rctw, TARA_ERR_ALLOC,
               "twiddle table allocation failed");
    p.xcfg = csh_default_xshape(nx);
    if (shape_ok(p, p.xcfg) != 1) p.xcfg = 0;
    return TARA_OK;
}

int csh_default_xshape(int nx)
{
    (void)nx;
    return 0;
}

int csh_num_xshapes() { return kNumXCands; }

int csh_xshape_ok(const CshPlan& p, int c) { return c >= 0 && c < kNumXCands && shape_ok(p, c) == 1; }

int csh_xpass(const CshPlan& p, hipStream_t st) { return xpass(p, st); }

int csh_nonfinite(const CshPlan& p, int* flags, hipStream_t st)
{
    std::vector<int> f(p.batch, 0);
    TARA_HIP(hipMemcpyAsync(f.data(), p.bad.p, p.batch * sizeof(int), hipMemcpyDeviceToHost, st));
    TARA_HIP(hipStreamSynchronize(st));
    int n = 0;
    for (int r = 0; r < p.batch; ++r) {
        n += f[r] != 0;
        if (flags) flags[r] = f[r] != 0;
    }
    return n;
}

int csh_clear_nonfinite(const CshPlan& p, int r, hipStream_t st)
{
    if (r < 0) {
        TARA_HIP(hipMemsetAsync(p.bad.p, 0, p.batch * sizeof(int), st));
    } else {
        TARA_CHECK(r < p.batch, TARA_ERR_ARG, "run %d out of range", r);
        TARA_HIP(hipMemsetAsync(p.bad.as<int>() + r, 0, sizeof(int), st));
    }
    return TARA_OK;
}

int csh_rhs_stage(const CshPlan& p, const double2* ws, int mode, const CshPhys& ph, double dt, double2* base,
                  double2* aux, double2* next, hipStream_t st, EventTimer* tm)
{
    int rc;
    if (tm) tm->begin(0, st);
    if ((rc = yinv(p, ws, ph, st))) return rc;
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
    a.nh = p.nh;
    a.nu = ph.nu;
    a.dt = dt;
    a.alpha = ph.alpha;
    a.lam = ph.lambda;
    a.screened = ph.screened;
    a.bs = p.bstride;
    a.nus = ph.nus;
    a.bad = p.bad.as<int>();
    if (tm) tm->begin(2, st);
    if ((rc = yfwd(p, mode, a, st))) return rc;
    if (tm) tm->end(2, st);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

// Fusing the forward y pass with the next inverse y pass pays from Ny = 1024
// (1024^2 RK4 step 0.181 -> 0.168 ms, 4096^2 2.41 -> 2.36 ms); below, the fused
// pass holds fewer waves per CU than the separate forward y pass, and the
// 128-512 ensemble step measured 6.65 ms fused against 6.43 ms separate.
static bool fuse_y(const CshPlan& p) { return p.ny >= 1024; }

int csh_yinv_pass(const CshPlan& p, const double2* w, const CshPhys& ph, hipStream_t st)
{
    if (!fuse_y(p)) return TARA_OK;  // the unfused step runs its own inverse y passes
    int rc = yinv(p, w, ph, st);
    if (rc) return rc;
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

int csh_fused_step(const CshPlan& p, int scheme, const CshPhys& ph, double dt, double2* w, double2* aux,
                   double2* next, bool plain_last, hipStream_t st, EventTimer* tm)
{
    TARA_CHECK(!ph.screened, TARA_ERR_ARG, "the fused step covers derive2csh only");
    int rc;
    if (!fuse_y(p)) {
        if (scheme == TARA_SCHEME_AB2) return csh_rhs_stage(p, w, CSH_AB2, ph, dt, w, aux, nullptr, st, tm);
        for (int k = 0; k < 4; ++k)
            if ((rc = csh_rhs_stage(p, k == 0 ? w : next, CSH_RK4_1 + k, ph, dt, w, aux, next, st, tm))) return rc;
        return TARA_OK;
    }
    YfwdArgs a;
    a.h = p.h.as<double2>();
    a.base = w;
    a.aux = aux;
    a.kx = p.kx.as<double>();
    a.ky = p.ky.as<double>();
    a.imin = p.imin.as<int>();
    a.nh = p.nh;
    a.nu = ph.nu;
    a.dt = dt;
    a.alpha = ph.alpha;
    a.lam = ph.lambda;
    a.screened = 0;
    a.bs = p.bstride;
    a.nus = ph.nus;
    a.bad = p.bad.as<int>();
    const int nst = scheme == TARA_SCHEME_AB2 ? 1 : 4;
    for (int k = 0; k < nst; ++k) {
        const int mode = scheme == TARA_SCHEME_AB2 ? CSH_AB2 : CSH_RK4_1 + k;
        a.ws = (scheme == TARA_SCHEME_AB2 || k == 0) ? w : next;
        a.next = (scheme == TARA_SCHEME_AB2 || k == 3) ? nullptr : next;
        if (tm) tm->begin(1, st);
        if ((rc = xpass(p, st))) return rc;
        if (tm) tm->end(1, st);
        const bool plain = plain_last && k == nst - 1;
        if (tm) tm->begin(2, st);
        if ((rc = plain ? yfwd(p, mode, a, st) : yfwd_inv(p, mode, a, st))) return rc;
        if (tm) tm->end(2, st);
    }
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

int csh_advance(const CshPlan& p, int scheme, int nsteps, const CshPhys& ph, double dt, double2* w,
                double2* aux, double2* next, hipStream_t st, EventTimer* tm)
{
    TARA_CHECK(!ph.screened, TARA_ERR_ARG, "the fused advance covers derive2csh only");
    if (nsteps <= 0) return TARA_OK;
    int rc;
    if (tm) tm->begin(0, st);
    if ((rc = csh_yinv_pass(p, w, ph, st))) return rc;  // the only stand-alone inverse y pass of the call
    if (tm) tm->end(0, st);
    for (int s = 0; s < nsteps; ++s)
        if ((rc = csh_fused_step(p, scheme, ph, dt, w, aux, next, s == nsteps - 1, st, tm))) return rc;
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
