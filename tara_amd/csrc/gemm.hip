// Batched strided GEMM on the CDNA matrix cores (MFMA), the building block of
// the Fourier-neural-operator surrogate:
//
//   C[z](m, n) (=|+=) sum_k A[z](m, k) B[z](k, n),   z = (z1, z2)
//
// A(m, k) at A + m*sam + k*sak (+ z1*sa1 + z2*sa2), likewise B and C, so any
// transpose is a choice of strides.  Operands are bfloat16 (gfx950's MFMA
// 32x32x16 bf16, FP32 accumulation) or float (MFMA 32x32x2 f32); the result is
// stored as float, accumulated into float, or rounded to bfloat16.
//
// Tiling: a 256-thread workgroup computes a 64 x 64 tile of C, each wavefront
// a 32 x 32 sub-tile; K advances 64 at a time through LDS (A tile row-major in
// k, B tile stored transposed so that both MFMA operand reads are
// k-contiguous).  Global fetches walk the operand's unit-stride dimension in
// 16-byte loads, and the next K tile is fetched into registers while the
// current one is multiplied.  Out-of-range rows, columns and k are zero-filled.
#include <cstdint>

#include "gemm.h"

namespace tara {

typedef __bf16 __attribute__((ext_vector_type(8))) v8bf;
typedef float __attribute__((ext_vector_type(16))) v16f;

__device__ __forceinline__ float bf2f(uint16_t h)
{
    return __uint_as_float((uint32_t)h << 16);
}
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

constexpr int GBM = 64, GBN = 64, GBK = 64;
constexpr int GPAD = 8;  // LDS row padding (elements)

// How a workgroup fetches a 64 x 64 operand tile (16 elements per thread):
//   GL_KFAST: unit stride along k; GL_MFAST: unit stride along m (n for B);
//   either in groups of 8 contiguous elements (one 16-byte load for bfloat16,
//   two for float when the addresses are aligned), else GL_SCALAR.
enum GemmLoad { GL_SCALAR = 0, GL_KFAST = 1, GL_MFAST = 2 };

template <typename T>
__device__ __forceinline__ void load8(const T* p, T* r)
{
    if constexpr (sizeof(T) == 2) {
        const uint4 v = *reinterpret_cast<const uint4*>(p);
        __builtin_memcpy(r, &v, 16);
    } else {
        const float4 a = reinterpret_cast<const float4*>(p)[0], b = reinterpret_cast<const float4*>(p)[1];
        __builtin_memcpy(r, &a, 16);
        __builtin_memcpy(r + 4, &b, 16);
    }
}

// Tile (rows r0.., k0..) of an operand whose element (row, k) is at
// base[row*srow + k*sk], rows < nrow, k < nk, into 16 registers.
template <typename T, int mode>
__device__ __forceinline__ void fetch_tile(const T* base, long srow, long sk, long r0, long k0, long nrow, long nk,
                                           bool vec, T* reg)
{
    const int tid = threadIdx.x;
    if constexpr (mode == GL_SCALAR) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = tid + q * 256, row = e % GBM, kk = e / GBM;
            const long gr = r0 + row, gk = k0 + kk;
            reg[q] = (gr < nrow && gk < nk) ? base[gr * srow + gk * sk] : T(0);
        }
        return;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int g = tid + h * 256;
        // KFAST: row g/8, k 8 (g%8) .. +7; MFAST: k g/8, rows 8 (g%8) .. +7
        const int a = g >> 3, b8 = (g & 7) * 8;
        const long gr = r0 + (mode == GL_KFAST ? a : b8), gk = k0 + (mode == GL_KFAST ? b8 : a);
        T* r = reg + 8 * h;
        if constexpr (mode == GL_KFAST) {
            if (vec && gr < nrow && gk + 8 <= nk) {
                load8(base + gr * srow + gk, r);
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) r[i] = (gr < nrow && gk + i < nk) ? base[gr * srow + (gk + i) * sk] : T(0);
            }
        } else {
            if (vec && gk < nk && gr + 8 <= nrow) {
                load8(base + gk * sk + gr, r);
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) r[i] = (gr + i < nrow && gk < nk) ? base[(gr + i) * srow + gk * sk] : T(0);
            }
        }
    }
}

// The 16 registers of fetch_tile into the LDS tile S[row][k].
template <typename T, int mode>
__device__ __forceinline__ void stash_tile(T (*S)[GBK + GPAD], const T* reg)
{
    const int tid = threadIdx.x;
    if constexpr (mode == GL_SCALAR) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = tid + q * 256;
            S[e % GBM][e / GBM] = reg[q];
        }
        return;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int g = tid + h * 256, a = g >> 3, b8 = (g & 7) * 8;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if constexpr (mode == GL_KFAST)
                S[a][b8 + i] = reg[8 * h + i];
            else
                S[b8 + i][a] = reg[8 * h + i];
        }
    }
}

template <typename T, int EPI, int AM, int BM>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g, int avec, int bvec)
{
    __shared__ __align__(16) T As[GBM][GBK + GPAD];
    __shared__ __align__(16) T Bs[GBN][GBK + GPAD];  // transposed: Bs[n][k]
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const int z1 = blockIdx.z % g.nb1, z2 = blockIdx.z / g.nb1;
    const T* A = reinterpret_cast<const T*>(g.a) + z1 * g.sa1 + z2 * g.sa2;
    const T* B = reinterpret_cast<const T*>(g.b) + z1 * g.sb1 + z2 * g.sb2;
    const long m0 = (long)blockIdx.y * GBM, n0 = (long)blockIdx.x * GBN;
    const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
    v16f acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    // result register i of lane l: row 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32
    const long gn = n0 + wn + lane % 32;
    // the epilogue's reads of C, issued with the first operand tiles so that
    // their latency overlaps the products instead of following them
    constexpr bool READC = EPI == GEMM_ACCUM_F32 || EPI == GEMM_ADD_TO;
    float cin[READC ? 16 : 1];
    if constexpr (READC) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const long gm = m0 + wm + 8 * (i / 4) + 4 * (lane / 32) + (i % 4);
            cin[i] = (gm < g.m && gn < g.n)
                         ? reinterpret_cast<const float*>(g.c)[z1 * g.sc1 + z2 * g.sc2 + gm * g.scm + gn * g.scn]
                         : 0.f;
        }
    }
    T ra[16], rb[16];
    // B is fetched as rows n, columns k: element (n, k) at B[n*sbn + k*sbk]
    fetch_tile<T, AM>(A, g.sam, g.sak, m0, 0, g.m, g.k, avec != 0, ra);
    fetch_tile<T, BM>(B, g.sbn, g.sbk, n0, 0, g.n, g.k, bvec != 0, rb);
    for (long k0 = 0; k0 < g.k; k0 += GBK) {
        stash_tile<T, AM>(As, ra);
        stash_tile<T, BM>(Bs, rb);
        __syncthreads();
        if (k0 + GBK < g.k) {  // next tile's loads overlap this tile's products
            fetch_tile<T, AM>(A, g.sam, g.sak, m0, k0 + GBK, g.m, g.k, avec != 0, ra);
            fetch_tile<T, BM>(B, g.sbn, g.sbk, n0, k0 + GBK, g.n, g.k, bvec != 0, rb);
        }
        if constexpr (sizeof(T) == 2) {
            // gfx950's 32x32x16 bf16: lane l supplies row/col l % 32, k = 8 (l / 32) .. +7, one 16-byte
            // LDS read per operand (rows of 144 bytes: 16-byte aligned, 8 lanes on distinct banks)
#pragma unroll
            for (int kk = 0; kk < GBK; kk += 16) {
                const int ko = kk + 8 * (lane / 32);
                const v8bf a = *reinterpret_cast<const v8bf*>(&As[wm + lane % 32][ko]);
                const v8bf b = *reinterpret_cast<const v8bf*>(&Bs[wn + lane % 32][ko]);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
            }
        } else {
            // 32x32x2 f32: lane l supplies row/col l % 32, k = l / 32
#pragma unroll
            for (int kk = 0; kk < GBK; kk += 2) {
                const int ko = kk + lane / 32;
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(As[wm + lane % 32][ko], Bs[wn + lane % 32][ko], acc, 0, 0,
                                                           0);
            }
        }
        __syncthreads();
    }
    if (gn >= g.n) return;
    const bool extra = g.bias || g.act || g.pre || g.c_f32 || g.c_bf16;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const long gm = m0 + wm + 8 * (i / 4) + 4 * (lane / 32) + (i % 4);
        if (gm >= g.m) continue;
        const long off = z1 * g.sc1 + z2 * g.sc2 + gm * g.scm + gn * g.scn;
        float v = acc[i] * g.alpha;
        if constexpr (READC) v += cin[i];
        if (extra) {
            if (g.bias) v += g.bias[gm];
            if (g.pre) g.pre[off] = v;
            if (g.act) v = 0.5f * v * (1.f + erff(v * 0.70710678118654752f));
            if (g.c_f32) g.c_f32[off] = v;
            if (g.c_bf16) reinterpret_cast<uint16_t*>(g.c_bf16)[off] = f2bf(v);
        }
        if constexpr (EPI == GEMM_STORE_F32 || EPI == GEMM_ACCUM_F32) {
            reinterpret_cast<float*>(g.c)[off] = v;
        } else if constexpr (EPI == GEMM_STORE_BF16) {
            reinterpret_cast<uint16_t*>(g.c)[off] = f2bf(v);
        }
    }
}


// bfloat16 stored as its 16 bits; MFMA reads the bits.
struct bf16raw {
    uint16_t v;
    __device__ bf16raw() = default;
    __device__ constexpr explicit bf16raw(int) : v(0) {}
    __device__ explicit operator short() const { return (short)v; }
};

template <typename T, int EPI, int AM>
int launch_b(const GemmArgs& g, int bm, int av, int bv, dim3 grid, hipStream_t st)
{
    switch (bm) {
        case GL_KFAST: hipLaunchKernelGGL((k_gemm<T, EPI, AM, GL_KFAST>), grid, dim3(256), 0, st, g, av, bv); break;
        case GL_MFAST: hipLaunchKernelGGL((k_gemm<T, EPI, AM, GL_MFAST>), grid, dim3(256), 0, st, g, av, bv); break;
        default: hipLaunchKernelGGL((k_gemm<T, EPI, AM, GL_SCALAR>), grid, dim3(256), 0, st, g, av, bv); break;
    }
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

template <typename T, int EPI>
int launch_a(const GemmArgs& g, int am, int bm, int av, int bv, dim3 grid, hipStream_t st)
{
    return am == GL_KFAST ? launch_b<T, EPI, GL_KFAST>(g, bm, av, bv, grid, st)
                          : launch_b<T, EPI, GL_MFAST>(g, bm, av, bv, grid, st);
}

template <typename T>
int launch_t(const GemmArgs& g, int epi, int am, int bm, int av, int bv, dim3 grid, hipStream_t st)
{
    switch (epi) {
        case GEMM_STORE_F32: return launch_a<T, GEMM_STORE_F32>(g, am, bm, av, bv, grid, st);
        case GEMM_ACCUM_F32: return launch_a<T, GEMM_ACCUM_F32>(g, am, bm, av, bv, grid, st);
        case GEMM_STORE_BF16: return launch_a<T, GEMM_STORE_BF16>(g, am, bm, av, bv, grid, st);
        case GEMM_ADD_TO: return launch_a<T, GEMM_ADD_TO>(g, am, bm, av, bv, grid, st);
        default: set_error("gemm: bad epilogue %d", epi); return TARA_ERR_ARG;
    }
}

int gemm(const GemmArgs& g, int dtype, int epi, hipStream_t st)
{
    TARA_CHECK(g.m > 0 && g.n > 0 && g.k > 0 && g.nb1 > 0 && g.nb2 > 0, TARA_ERR_ARG, "gemm: bad shape");
    TARA_CHECK((long)g.nb1 * g.nb2 <= 65535, TARA_ERR_ARG, "gemm: too many batches");
    dim3 grid((unsigned)((g.n + GBN - 1) / GBN), (unsigned)((g.m + GBM - 1) / GBM), (unsigned)(g.nb1 * g.nb2));
    TARA_CHECK(grid.y <= 65535, TARA_ERR_ARG, "gemm: m too large");
    const long es = dtype == GEMM_BF16 ? 2 : 4;
    // fetch mode per operand (rows m for A, n for B) and whether 16-byte loads are aligned
    auto mode = [](long srow, long sk) { return sk == 1 ? GL_KFAST : (srow == 1 ? GL_MFAST : GL_SCALAR); };
    auto aligned = [es](const void* p, long s_other, long s1, long s2) {
        return ((uintptr_t)p % 16 == 0) && (s_other * es) % 16 == 0 && (s1 * es) % 16 == 0 && (s2 * es) % 16 == 0;
    };
    const int am = mode(g.sam, g.sak), bm = mode(g.sbn, g.sbk);
    const int av = am == GL_KFAST ? aligned(g.a, g.sam, g.sa1, g.sa2) : am == GL_MFAST ? aligned(g.a, g.sak, g.sa1, g.sa2) : 0;
    const int bv = bm == GL_KFAST ? aligned(g.b, g.sbn, g.sb1, g.sb2) : bm == GL_MFAST ? aligned(g.b, g.sbk, g.sb1, g.sb2) : 0;
    TARA_CHECK(am != GL_SCALAR, TARA_ERR_UNSUPPORTED, "gemm: A needs a unit stride along m or k");
    if (dtype == GEMM_BF16) return launch_t<bf16raw>(g, epi, am, bm, av, bv, grid, st);
    return launch_t<float>(g, epi, am, bm, av, bv, grid, st);
}

}  // namespace tara

using namespace tara;

// Device pointers; strides in elements.  Exposed for testing and for callers
// that compose their own networks.
extern "C" int tara_gemm(int dtype, int epi, long m, long n, long k, const void* a, long sam, long sak,
                         const void* b, long sbk, long sbn, void* c, long scm, long scn, int nb1, long sa1, long sb1,
                         long sc1, int nb2, long sa2, long sb2, long sc2, float alpha)
{
    TARA_CHECK(a && b && c, TARA_ERR_ARG, "gemm: null pointer");
    TARA_CHECK(epi >= GEMM_STORE_F32 && epi <= GEMM_STORE_BF16, TARA_ERR_ARG, "gemm: bad epilogue %d", epi);
    GemmArgs g{};
    g.m = m;
    g.n = n;
    g.k = k;
    g.a = a;
    g.sam = sam;
    g.sak = sak;
    g.b = b;
    g.sbk = sbk;
    g.sbn = sbn;
    g.c = c;
    g.scm = scm;
    g.scn = scn;
    g.nb1 = nb1;
    g.sa1 = sa1;
    g.sb1 = sb1;
    g.sc1 = sc1;
    g.nb2 = nb2;
    g.sa2 = sa2;
    g.sb2 = sb2;
    g.sc2 = sc2;
    g.alpha = alpha;
    int rc = gemm(g, dtype, epi, nullptr);
    if (rc) return rc;
    TARA_HIP(hipDeviceSynchronize());
    return TARA_OK;
}
