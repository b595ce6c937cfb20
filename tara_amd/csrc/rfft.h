// Register-resident Stockham FFT for gfx950 (FP64).
//
// A length-N sequence is spread over T = N/E threads (E = min(8, N) complex
// elements per thread, 32 VGPRs).  Before pass p (span L, radix R) the thread
// with slot t holds, for each of its E/R butterflies u (b = t + u*T), the R
// inputs x[b + s*N/R] in v[u*R + s]; after the in-register twiddle + DFT_R,
// v[u*R + q] holds output q of butterfly b, destined for Stockham position
// (b/L)*L*R + q*L + b%L.  Between passes the block exchanges through LDS
// (write outputs, barrier, read next inputs, barrier); the first pass takes its
// inputs straight from global memory and the last pass leaves its outputs in
// registers, so a transform costs (passes - 1) LDS round trips and no LDS
// staging of the global traffic.
//
// LDS positions are padded by one element per 8 (rf_pad): the LDS serves 8 lanes
// of 16-byte accesses per cycle (128 B), and with this padding the radix-8
// scatter (stride 8 elements) and every gather of the schedule put 8
// consecutive lanes on 8 distinct 16-byte bank groups.
// Radix schedule and twiddle tables are those of fft_core.h.
#pragma once
#include "fft_core.h"

namespace tara {

template <int N>
struct RF {
    static constexpr int E = N < 8 ? N : 8;  // elements per thread
    static constexpr int T = N / E;            // threads per sequence
    static constexpr int LDS = N + N / 8;      // padded exchange length
};

__device__ __forceinline__ int rf_pad(int n) { return n + (n >> 3); }

// Span of the last pass.
__host__ __device__ constexpr int rf_last_span(int n)
{
    int span = 1;
    while (span * fft_radix(n / span) < n) span *= fft_radix(n / span);
    return span;
}

// Global index of register e's element before the pass of span L.
template <int N, int L>
__device__ __forceinline__ int rf_in_idx(int t, int e)
{
    constexpr int R = fft_radix(N / L);
    const int b = t + (e / R) * RF<N>::T;
    return b + (e % R) * (N / R);
}

// Global index of register e's element after the pass of span L.
template <int N, int L>
__device__ __forceinline__ int rf_out_idx(int t, int e)
{
    constexpr int R = fft_radix(N / L);
    const int b = t + (e / R) * RF<N>::T;
    return (b / L) * L * R + (e % R) * L + (b % L);
}

// Natural-order index of register e after the whole transform.
template <int N>
__device__ __forceinline__ int rf_final_idx(int t, int e)
{
    return rf_out_idx<N, rf_last_span(N)>(t, e);
}
// Index of register e that the transform expects before its first pass.
template <int N>
__device__ __forceinline__ int rf_first_idx(int t, int e)
{
    return rf_in_idx<N, 1>(t, e);
}

template <int N, int L, int DIR>
__device__ __forceinline__ void rf_compute_check()
{
    static_assert(fft_radix(N / L) <= RF<N>::E, "radix exceeds the elements a thread holds");
}

template <int N, int L, int DIR>
__device__ __forceinline__ void rf_compute(double2 (&v)[RF<N>::E], int t, const double2* tw)
{
    constexpr int R = fft_radix(N / L);
    constexpr int E = RF<N>::E;
    constexpr int TWO = fft_tw_offset(N, L);
    rf_compute_check<N, L, DIR>();
#pragma unroll
    for (int u = 0; u < E / R; ++u) {
        if (L > 1) {
            const int b = t + u * RF<N>::T;
            const double2* w = tw + TWO + (b % L) * (R - 1);
#pragma unroll
            for (int s = 1; s < R; ++s) {
                double2 x = w[s - 1];
                if (DIR > 0) x.y = -x.y;
                v[u * R + s] = cmul(v[u * R + s], x);
            }
        }
        Dft<R, DIR>::run(&v[u * R]);
    }
}

template <int N, int L, int DIR>
__device__ __forceinline__ void rf_rec(double2 (&v)[RF<N>::E], int t, double2* lds, const double2* tw)
{
    constexpr int R = fft_radix(N / L);
    constexpr int E = RF<N>::E;
    rf_compute<N, L, DIR>(v, t, tw);
    if constexpr (L * R < N) {
#pragma unroll
        for (int e = 0; e < E; ++e) lds[rf_pad(rf_out_idx<N, L>(t, e))] = v[e];
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = lds[rf_pad(rf_in_idx<N, L * R>(t, e))];
        __syncthreads();
        rf_rec<N, L * R, DIR>(v, t, lds, tw);
    }
}

// Transform one sequence.  v: inputs at rf_first_idx positions; on return the
// outputs at rf_final_idx positions.  lds: this sequence's RF<N>::LDS-element
// exchange buffer.  Contains block-wide barriers: every thread of the block
// must call it the same number of times.
template <int N, int DIR>
__device__ __forceinline__ void rf_fft(double2 (&v)[RF<N>::E], int t, double2* lds, const double2* tw)
{
    rf_rec<N, 1, DIR>(v, t, lds, tw);
}

// Reorders registers from final (output) positions to first-pass (input)
// positions through LDS, e.g. to feed a transform's output into another
// transform of the same length.
template <int N>
__device__ __forceinline__ void rf_final_to_first(double2 (&v)[RF<N>::E], int t, double2* lds)
{
    constexpr int E = RF<N>::E;
#pragma unroll
    for (int e = 0; e < E; ++e) lds[rf_pad(rf_final_idx<N>(t, e))] = v[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = lds[rf_pad(rf_first_idx<N>(t, e))];
    __syncthreads();
}

}  // namespace tara
