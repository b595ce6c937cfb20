// Shared device pieces of the spectral passes: tile geometry, the half-length
// complex packing used by every x (halved-axis) pass, and launch-shape tables.
//
// Device arrays of half spectra use a padded leading dimension `ld` >= Nh
// (complex elements, multiple of 8) so that column tiles of TC = 1, 2, 4, 8
// never straddle the array edge; the padding columns are kept at zero.
#pragma once
#include "fft_core.h"

namespace tara {

// Workgroups are dealt round-robin over the 8 XCDs (each with its own L2).
// xcd_remap gives XCD x a contiguous range of logical tiles, so tiles that
// share cache lines (neighbouring columns or rows) run on the same L2.  It is a
// bijection on [0, n) for any dealing order; only locality depends on it.
__device__ __forceinline__ int xcd_remap(int w, int n)
{
    const int per = n / 8, rem = n % 8;
    const int x = w % 8, q = w / 8;
    return x * per + (x < rem ? x : rem) + q;
}

// Launch shapes of the column (y / z) passes, by transform length:
// TC columns per workgroup, NT threads.
template <int N>
struct ColShape {
    static constexpr int TC = N >= 4096 ? 2 : N >= 2048 ? 4 : 8;
    static constexpr int NT = N >= 1024 ? 512 : N >= 512 ? 256 : N >= 256 ? 128 : 64;
    static constexpr int PER = TC * N / NT;  // elements owned per thread
    static_assert(PER * NT == TC * N, "column tile must split evenly");
};

// Launch shapes of the row (x) passes, by half length M = Nx/2:
// RB rows per workgroup, NT threads, LDS row pitch (complex).
template <int M>
struct RowShape {
    static constexpr int RB = M >= 2048 ? 1 : M >= 1024 ? 2 : M >= 512 ? 4 : 8;
    static constexpr int NT = (RB * M) >= 2048 ? 256 : (RB * M) >= 1024 ? 128 : 64;
    static constexpr int PITCH = M + 1;
    static constexpr int PER = RB * M / NT;
    static_assert(PER * NT == RB * M, "row batch must split evenly");
};

// c2r packing along x (length Nx = 2M): given the half spectrum X[0..M] of a
// Hermitian sequence, Z[k] = (X[k] + conj X[M-k]) + i e^{+2 pi i k/Nx} (X[k] - conj X[M-k])
// for k = 0..M-1; the backward length-M FFT of Z then holds x[2m] + i x[2m+1].
// The imaginary parts of X[0] and X[M] are ignored, as FFTW's c2r does.
__device__ __forceinline__ double2 c2r_pack(double2 xk, double2 xmk, int k, const double2* __restrict__ rctw)
{
    if (k == 0) {
        xk.y = 0.0;
        xmk.y = 0.0;
    }
    const double2 a = make_double2(xk.x + xmk.x, xk.y - xmk.y);   // X[k] + conj X[M-k]
    const double2 d = make_double2(xk.x - xmk.x, xk.y + xmk.y);   // X[k] - conj X[M-k]
    const double2 w = cconj(rctw[k]);                              // e^{+2 pi i k/Nx}
    const double2 t = cmul(w, d);
    return make_double2(a.x - t.y, a.y + t.x);                     // a + i t
}

// r2c unpacking along x: Z = forward length-M FFT of z[m] = x[2m] + i x[2m+1];
// X[k] = 1/2 (Z[k] + conj Z[M-k]) - i/2 e^{-2 pi i k/Nx} (Z[k] - conj Z[M-k]),
// with Z[M] == Z[0].  For k = 0 also returns X[M] = Re Z0 - Im Z0 in *xm.
// The same with the root e^{-2 pi i k/Nx} passed in (wk = rctw[k]), so that a
// caller can fetch its roots before it issues stores: on gfx950 loads and
// stores share one in-order counter, and waiting for a load issued after a
// store also waits for the store.
__device__ __forceinline__ double2 r2c_unpack_w(double2 zk, double2 zmk, int k, double2 wk, double2* xm)
{
    if (k == 0) {
        *xm = make_double2(zk.x - zk.y, 0.0);
        return make_double2(zk.x + zk.y, 0.0);
    }
    const double2 e = make_double2(0.5 * (zk.x + zmk.x), 0.5 * (zk.y - zmk.y));
    const double2 d = make_double2(0.5 * (zk.x - zmk.x), 0.5 * (zk.y + zmk.y));
    const double2 t = cmul(wk, d);                                 // e^{-2 pi i k/Nx} d
    return make_double2(e.x + t.y, e.y - t.x);                     // e - i t
}

__device__ __forceinline__ double2 r2c_unpack(double2 zk, double2 zmk, int k, const double2* __restrict__ rctw,
                                              double2* xm)
{
    return r2c_unpack_w(zk, zmk, k, k == 0 ? make_double2(1.0, 0.0) : rctw[k], xm);
}

}  // namespace tara
