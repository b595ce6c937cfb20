// Column FFTs over tiles of 8 adjacent kx columns of kx-fastest arrays (row
// length padded to a multiple of 8, so a tile row is one 128-byte line).
// Shared by the 3D MHD passes and the 2D compressible families.
#pragma once
#include "passes.h"
#include "rfft.h"

namespace tara {

template <int N, int EM = 8>
struct CShape {
    static constexpr int T = RF<N, EM>::T;        // threads per column
    static constexpr int NT = 8 * T;              // 8 kx columns per tile (128-byte rows)
    static constexpr int LP = RF<N, EM>::LDS + 1; // per-column buffer; +1 shifts banks between columns
};

struct ColArgs {
    const double2* src[14];
    double2* dst[14];
    long es;       // element stride along the transform (complex elements)
    long pstride;  // plane stride
    int tpp;       // tiles (of 8 columns) per plane
};

// Column FFT over tiles of 8 adjacent kx columns; blockIdx.y selects the field.
template <int N, int DIR>
__global__ __launch_bounds__(CShape<N>::NT) void k_col_fft(ColArgs a, const double2* tw)
{
    using S = CShape<N>;
    constexpr int E = RF<N>::E;
    __shared__ double2 lds[8 * S::LP];
    const int c = threadIdx.x % 8, t = threadIdx.x / 8;  // column fastest: 8 x 16 B contiguous
    const int tile = xcd_remap(blockIdx.x, gridDim.x);
    const size_t off = (size_t)(tile / a.tpp) * a.pstride + (tile % a.tpp) * 8 + c;
    const double2* src = a.src[blockIdx.y] + off;
    double2* dst = a.dst[blockIdx.y] + off;
    double2 v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = src[(size_t)rf_first_idx<N>(t, e) * a.es];
    rf_fft<N, DIR>(v, t, lds + c * S::LP, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) dst[(size_t)rf_final_idx<N>(t, e) * a.es] = v[e];
}

// Launch over ntiles tiles x nfields fields (lengths 16..256).
int col_fft_tiles(int n, const ColArgs& a, int ntiles, int nfields, int dir, const double2* tw, hipStream_t st);

}  // namespace tara
