// FP64 Stockham FFT building blocks for gfx950 (CDNA4), used by every
// spectral kernel of the library.
//
// Design: a workgroup transforms NSEQ sequences of length N that sit in LDS.
// Each radix-R pass reads R strided elements per butterfly into registers,
// applies the inter-pass twiddles, runs an in-register DFT_R, and writes the
// outputs back in Stockham self-sorting order (in place: all reads of a pass
// complete before a barrier, then all writes), so the result is in natural
// order without a bit-reversal step.  Element (s, n) of the batch lives at
// lds[s*ss + n*es]: column tiles use es = TC (interleaved columns), row batches
// use es = 1, ss = padded row length.
//
// Sign convention (FFTW's): DIR = -1 is the forward transform e^{-2 pi i jk/N},
// DIR = +1 the backward one; both unnormalised.
//
// Twiddles: one table per (N, schedule), laid out pass after pass; pass p with
// span L and radix R owns L*(R-1) entries T[k*(R-1) + s-1] = e^{-2 pi i k s/(L R)}
// (forward sign, conjugated on the fly for DIR = +1).  Tables are generated on
// the host from long-double sinl/cosl (plan.cpp) in the same order as
// fft_schedule() below, which host and device share.
#pragma once
#include <hip/hip_runtime.h>

namespace tara {

// ---------------------------------------------------------------- schedule
// Radix of the pass that starts with remaining length REM = N / L, for a
// schedule capped at radix RMAX (8 or 4): RMAX while it divides, then the
// remainder.  A twiddle table belongs to one (N, RMAX) schedule.
__host__ __device__ constexpr int fft_radix(int rem, int rmax = 8)
{
    return rem % rmax == 0 ? rmax : rem;
}
// Number of twiddle entries used by all passes before span L (L a product of
// the schedule's radices).
__host__ __device__ constexpr int fft_tw_offset(int n, int l, int rmax = 8)
{
    int off = 0, span = 1;
    while (span < l) {
        int r = fft_radix(n / span, rmax);
        off += span * (r - 1);
        span *= r;
    }
    return off;
}
__host__ __device__ constexpr int fft_tw_size(int n, int rmax = 8) { return fft_tw_offset(n, n, rmax); }

// --------------------------------------------------------------- complex ops
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// a * (DIR * i)
template <int DIR>
__device__ __forceinline__ double2 mul_di(double2 a)
{
    return DIR > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// ----------------------------------------------------- in-register DFT_R
// y_q = sum_s v_s e^{DIR 2 pi i s q / R}, radix-2 decimation in time.
template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<1, DIR> {
    static __device__ __forceinline__ void run(double2*) {}
};

template <int DIR>
struct Dft<2, DIR> {
    static __device__ __forceinline__ void run(double2* v)
    {
        double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int DIR>
struct Dft<4, DIR> {
    static __device__ __forceinline__ void run(double2* v)
    {
        double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        double2 b0 = cadd(v[1], v[3]), b1 = mul_di<DIR>(csub(v[1], v[3]));
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, b1);
        v[3] = csub(a1, b1);
    }
};

// W_R^q = e^{DIR 2 pi i q / R} for R = 8, 16 (q < R/2), exact decimal constants.
template <int R, int DIR>
__device__ __forceinline__ double2 wr(int q)
{
    constexpr double C1 = 0.92387953251128675613, S1 = 0.38268343236508977173;  // cos, sin(pi/8)
    constexpr double H = 0.70710678118654752440;                                // cos(pi/4)
    double c, s;
    if (R == 8) {
        c = (q == 0) ? 1.0 : (q == 1) ? H : (q == 2) ? 0.0 : -H;
        s = (q == 0) ? 0.0 : (q == 1) ? H : (q == 2) ? 1.0 : H;
    } else {  // R == 16
        const double cs[8] = {1.0, C1, H, S1, 0.0, -S1, -H, -C1};
        const double sn[8] = {0.0, S1, H, C1, 1.0, C1, H, S1};
        c = cs[q];
        s = sn[q];
    }
    return make_double2(c, DIR * s);
}

template <int R, int DIR>
struct DftSplit {
    static __device__ __forceinline__ void run(double2* v)
    {
        constexpr int H = R / 2;
        double2 e[H], o[H];
#pragma unroll
        for (int s = 0; s < H; ++s) {
            e[s] = v[2 * s];
            o[s] = v[2 * s + 1];
        }
        Dft<H, DIR>::run(e);
        Dft<H, DIR>::run(o);
#pragma unroll
        for (int q = 0; q < H; ++q) {
            double2 t;
            if (q == 0)
                t = o[0];
            else if (2 * q == H)
                t = mul_di<DIR>(o[q]);
            else
                t = cmul(o[q], wr<R, DIR>(q));
            v[q] = cadd(e[q], t);
            v[q + H] = csub(e[q], t);
        }
    }
};
template <int DIR>
struct Dft<8, DIR> : DftSplit<8, DIR> {};
template <int DIR>
struct Dft<16, DIR> : DftSplit<16, DIR> {};

// ------------------------------------------------------------ Stockham pass
// Butterfly g of the batch: with SEQ_FAST the sequence index varies fastest
// across threads (conflict-free for interleaved column tiles), otherwise the
// butterfly index does (row batches).
template <int N, int L, int DIR, int NT, int NSEQ, bool SEQ_FAST>
__device__ __forceinline__ void stockham_pass(double2* buf, int es, int ss, const double2* __restrict__ tw)
{
    constexpr int R = fft_radix(N / L);
    constexpr int NBF = NSEQ * (N / R);
    constexpr int NB = (NBF + NT - 1) / NT;
    constexpr int TWO = fft_tw_offset(N, L);
    double2 v[NB][R];
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        const int g = threadIdx.x + m * NT;
        if ((NBF % NT == 0) || g < NBF) {
            const int seq = SEQ_FAST ? g % NSEQ : g / (N / R);
            const int b = SEQ_FAST ? g / NSEQ : g % (N / R);
            const double2* src = buf + seq * ss + b * es;
#pragma unroll
            for (int s = 0; s < R; ++s) v[m][s] = src[s * (N / R) * es];
            if (L > 1) {
                const int k = b % L;
                const double2* t = tw + TWO + k * (R - 1);
#pragma unroll
                for (int s = 1; s < R; ++s) {
                    double2 w = t[s - 1];
                    if (DIR > 0) w.y = -w.y;
                    v[m][s] = cmul(v[m][s], w);
                }
            }
            Dft<R, DIR>::run(v[m]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < NB; ++m) {
        const int g = threadIdx.x + m * NT;
        if ((NBF % NT == 0) || g < NBF) {
            const int seq = SEQ_FAST ? g % NSEQ : g / (N / R);
            const int b = SEQ_FAST ? g / NSEQ : g % (N / R);
            const int j = b / L, k = b % L;
            double2* dst = buf + seq * ss + (j * L * R + k) * es;
#pragma unroll
            for (int q = 0; q < R; ++q) dst[q * L * es] = v[m][q];
        }
    }
    __syncthreads();
}

template <int N, int L, int DIR, int NT, int NSEQ, bool SEQ_FAST>
__device__ __forceinline__ void lds_fft_from(double2* buf, int es, int ss, const double2* __restrict__ tw)
{
    if constexpr (L < N) {
        stockham_pass<N, L, DIR, NT, NSEQ, SEQ_FAST>(buf, es, ss, tw);
        lds_fft_from<N, L * fft_radix(N / L), DIR, NT, NSEQ, SEQ_FAST>(buf, es, ss, tw);
    }
}

// In-place FFT of NSEQ length-N sequences in LDS.  Caller must have issued a
// barrier after filling buf; on return a barrier has been passed.
template <int N, int DIR, int NT, int NSEQ, bool SEQ_FAST>
__device__ __forceinline__ void lds_fft(double2* buf, int es, int ss, const double2* __restrict__ tw)
{
    lds_fft_from<N, 1, DIR, NT, NSEQ, SEQ_FAST>(buf, es, ss, tw);
}

}  // namespace tara
