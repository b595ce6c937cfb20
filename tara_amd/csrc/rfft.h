// Register-resident Stockham FFT for gfx950 (FP64).
//
// A length-N sequence is spread over T = N/E threads (E = min(EM, N) complex
// elements per thread; EM = 8 or 4, which is also the radix cap of the
// schedule).  Before pass p (span L, radix R) the thread with slot t holds, for
// each of its E/R butterflies u (b = t + u*T), the R inputs x[b + s*N/R] in
// v[u*R + s]; after the in-register twiddle + DFT_R, v[u*R + q] holds output q
// of butterfly b, destined for Stockham position (b/L)*L*R + q*L + b%L.
// Between passes the threads exchange through LDS (write outputs, sync, read
// next inputs, sync); the first pass takes its inputs straight from global
// memory and the last pass leaves its outputs in registers, so a transform
// costs (passes - 1) LDS round trips and no LDS staging of the global traffic.
//
// Sync scope: WAVE = false uses workgroup barriers; WAVE = true is for a
// sequence whose T threads all sit in one wavefront (T <= 64, wave-aligned),
// where LDS visibility only needs the wave's own LDS operations to complete.
//
// LDS positions are padded by one element per 8 (rf_pad): the LDS serves 8 lanes
// of 16-byte accesses per cycle (128 B), and with this padding the radix-8
// scatter (stride 8 elements) and every gather of the schedule put 8
// consecutive lanes on 8 distinct 16-byte bank groups.
// Radix schedule and twiddle tables are those of fft_core.h for (N, EM);
// EM = 16 runs radix-16 passes (4096 = 16^3: two exchanges instead of three).
#pragma once
#include "fft_core.h"

namespace tara {

template <int N, int EM = 8>
struct RF {
    static constexpr int E = N < EM ? N : EM;  // elements per thread (also the radix cap)
    static constexpr int T = N / E;            // threads per sequence
    static constexpr int LDS = N + N / 8;      // padded exchange length
};

__device__ __forceinline__ int rf_pad(int n) { return n + (n >> 3); }

// v, through an empty asm statement the compiler cannot see into: index
// arithmetic derived from the result is recomputed where it is used instead of
// being hoisted out of a loop of transforms and held in registers.
__device__ __forceinline__ int rf_opaque(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

template <bool WAVE>
__device__ __forceinline__ void rf_sync()
{
    if (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// Span of the last pass.
__host__ __device__ constexpr int rf_last_span(int n, int rmax = 8)
{
    int span = 1;
    while (span * fft_radix(n / span, rmax) < n) span *= fft_radix(n / span, rmax);
    return span;
}

// Global index of register e's element before the pass of span L.
template <int N, int L, int EM = 8>
__device__ __forceinline__ int rf_in_idx(int t, int e)
{
    constexpr int R = fft_radix(N / L, RF<N, EM>::E);
    const int b = t + (e / R) * RF<N, EM>::T;
    return b + (e % R) * (N / R);
}

// Global index of register e's element after the pass of span L.
template <int N, int L, int EM = 8>
__device__ __forceinline__ int rf_out_idx(int t, int e)
{
    constexpr int R = fft_radix(N / L, RF<N, EM>::E);
    const int b = t + (e / R) * RF<N, EM>::T;
    return (b / L) * L * R + (e % R) * L + (b % L);
}

// Natural-order index of register e after the whole transform.
template <int N, int EM = 8>
__device__ __forceinline__ int rf_final_idx(int t, int e)
{
    return rf_out_idx<N, rf_last_span(N, RF<N, EM>::E), EM>(t, e);
}
// Index of register e that the transform expects before its first pass.
template <int N, int EM = 8>
__device__ __forceinline__ int rf_first_idx(int t, int e)
{
    return rf_in_idx<N, 1, EM>(t, e);
}

// Twiddles of pass L for the E/R radix-R butterflies a thread computes (none
// in the first pass).  Of the R - 1 powers w^s of a butterfly's twiddle only
// the power-of-two ones (w, w^2, w^4, w^8) are fetched from the table; the
// others are formed as one product of two already available powers (w^s =
// w^(s - low) w^low, low = lowest set bit of s), at most three products deep.
// That is log2(R) loads instead of R - 1 (the memory-latency-bound x pass
// issued most of its loads for twiddles) at a cost of a few ulp.  Loaded one
// pass ahead, before the LDS exchange, so that their latency overlaps the
// exchange (the exchange barriers wait on LDS operations only).
template <int N, int L, int EM>
struct RfTw {
    static constexpr int E = RF<N, EM>::E;
    static constexpr int R = fft_radix(N / L, E);
    static constexpr int LOGR = R >= 16 ? 4 : R >= 8 ? 3 : R >= 4 ? 2 : 1;
    static_assert((1 << LOGR) == R, "radix must be a power of two");
    static constexpr int CNT = L > 1 ? (E / R) * LOGR : 1;
    double2 w[CNT];
    __device__ __forceinline__ void load(int t, const double2* tw)
    {
        if constexpr (L > 1) {
            constexpr int TWO = fft_tw_offset(N, L, E);
#pragma unroll
            for (int u = 0; u < E / R; ++u) {
                const int b = t + u * RF<N, EM>::T;
                const double2* p = tw + TWO + (b % L) * (R - 1);
#pragma unroll
                for (int q = 0; q < LOGR; ++q) w[u * LOGR + q] = p[(1 << q) - 1];
            }
        }
    }
};

template <int N, int L, int DIR, int EM>
__device__ __forceinline__ void rf_compute(double2 (&v)[(RF<N, EM>::E)], const RfTw<N, L, EM>& tw)
{
    constexpr int E = RF<N, EM>::E;
    constexpr int R = fft_radix(N / L, E);
    constexpr int LOGR = RfTw<N, L, EM>::LOGR;
    static_assert(R <= E, "radix exceeds the elements a thread holds");
#pragma unroll
    for (int u = 0; u < E / R; ++u) {
        if constexpr (L > 1) {
            double2 x[R];
#pragma unroll
            for (int s = 1; s < R; ++s) {
                const int low = s & -s;
                if (low == s) {
                    int q = 0;
                    while ((1 << q) < s) ++q;
                    x[s] = tw.w[u * LOGR + q];
                } else {
                    x[s] = cmul(x[s - low], x[low]);
                }
            }
#pragma unroll
            for (int s = 1; s < R; ++s) {
                double2 y = x[s];
                if (DIR > 0) y.y = -y.y;
                v[u * R + s] = cmul(v[u * R + s], y);
            }
        }
        Dft<R, DIR>::run(&v[u * R]);
    }
}

template <int N, int L, int DIR, int EM, bool WAVE, bool SPLIT, bool TWPRE>
__device__ __forceinline__ void rf_rec(double2 (&v)[(RF<N, EM>::E)], int t, double2* lds, const double2* tw,
                                       const RfTw<N, L, EM>& cur)
{
    constexpr int E = RF<N, EM>::E;
    constexpr int R = fft_radix(N / L, E);
    rf_compute<N, L, DIR, EM>(v, cur);
    if constexpr (L * R < N) {
        RfTw<N, L * R, EM> nxt;
        if constexpr (TWPRE) nxt.load(t, tw);
        if constexpr (SPLIT) {
            // real and imaginary parts in two rounds through a buffer of
            // doubles: half the LDS of the double2 exchange
            double* d = reinterpret_cast<double*>(lds);
#pragma unroll
            for (int e = 0; e < E; ++e) d[rf_pad(rf_out_idx<N, L, EM>(t, e))] = v[e].x;
            rf_sync<WAVE>();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e].x = d[rf_pad(rf_in_idx<N, L * R, EM>(t, e))];
            rf_sync<WAVE>();
#pragma unroll
            for (int e = 0; e < E; ++e) d[rf_pad(rf_out_idx<N, L, EM>(t, e))] = v[e].y;
            rf_sync<WAVE>();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e].y = d[rf_pad(rf_in_idx<N, L * R, EM>(t, e))];
            rf_sync<WAVE>();
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) lds[rf_pad(rf_out_idx<N, L, EM>(t, e))] = v[e];
            rf_sync<WAVE>();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = lds[rf_pad(rf_in_idx<N, L * R, EM>(t, e))];
            rf_sync<WAVE>();
        }
        if constexpr (!TWPRE) nxt.load(t, tw);
        rf_rec<N, L * R, DIR, EM, WAVE, SPLIT, TWPRE>(v, t, lds, tw, nxt);
    }
}

// Transform one sequence.  v: inputs at rf_first_idx positions; on return the
// outputs at rf_final_idx positions.  lds: this sequence's exchange buffer of
// RF<N, EM>::LDS complex elements (SPLIT = false) or RF<N, EM>::LDS doubles
// (SPLIT = true: real and imaginary parts exchanged in two rounds).  tw: the
// twiddle table of the (N, RF<N, EM>::E) schedule.  With WAVE = false it
// contains workgroup barriers: every thread of the workgroup must call it the
// same number of times.  TWPRE: fetch each pass's twiddles before the
// preceding exchange (more live registers across the exchange).
template <int N, int DIR, int EM = 8, bool WAVE = false, bool SPLIT = false, bool TWPRE = true>
__device__ __forceinline__ void rf_fft(double2 (&v)[(RF<N, EM>::E)], int t, double2* lds, const double2* tw)
{
    const RfTw<N, 1, EM> first{};
    rf_rec<N, 1, DIR, EM, WAVE, SPLIT, TWPRE>(v, t, lds, tw, first);
}

// Number of LDS exchanges of a length-n transform with radix cap e: one
// between consecutive passes of the schedule rf_rec walks (span L, radix
// fft_radix(n / L, e), exchange while L * radix < n).
__host__ __device__ constexpr int rf_exchanges(int n, int e)
{
    int x = 0;
    for (int span = 1; span * fft_radix(n / span, e) < n; span *= fft_radix(n / span, e)) ++x;
    return x;
}

// The workgroup barriers of rf_fft<N, DIR, EM, false, SPLIT> without its
// work: for the waves of a workgroup that hold no sequence of this transform,
// so that every wave meets the same barriers (2 per exchange, 4 when SPLIT).
// Callers must branch on a wave-uniform condition.
template <int N, int EM = 8, bool SPLIT = false>
__device__ __forceinline__ void rf_fft_barriers_only()
{
    constexpr int nb = rf_exchanges(N, RF<N, EM>::E) * (SPLIT ? 4 : 2);
#pragma unroll
    for (int b = 0; b < nb; ++b) __syncthreads();
}

// Reorders registers from final (output) positions to first-pass (input)
// positions through LDS, e.g. to feed a transform's output into another
// transform of the same length.
template <int N, int EM = 8, bool WAVE = false>
__device__ __forceinline__ void rf_final_to_first(double2 (&v)[(RF<N, EM>::E)], int t, double2* lds)
{
    constexpr int E = RF<N, EM>::E;
#pragma unroll
    for (int e = 0; e < E; ++e) lds[rf_pad(rf_final_idx<N, EM>(t, e))] = v[e];
    rf_sync<WAVE>();
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = lds[rf_pad(rf_first_idx<N, EM>(t, e))];
    rf_sync<WAVE>();
}

}  // namespace tara
