// Batched strided GEMM on the CDNA matrix cores (MFMA), the building block of
// the Fourier-neural-operator surrogate:
//
//   C[z](m, n) (=|+=) sum_k A[z](m, k) B[z](k, n),   z = (z1, z2)
//
// A(m, k) at A + m*sam + k*sak (+ z1*sa1 + z2*sa2), likewise B and C, so any
// transpose is a choice of strides.  Operands are bfloat16 (MFMA
// 32x32x8 bf16, FP32 accumulation) or float (MFMA 32x32x2 f32); the result is
// stored as float, accumulated into float, or rounded to bfloat16.
//
// Tiling: a 256-thread workgroup computes a 64 x 64 tile of C, each wavefront
// a 32 x 32 sub-tile; K advances 32 at a time through LDS (A tile row-major in
// k, B tile stored transposed so that both operand fetches read k-contiguous
// elements).  Out-of-range rows, columns and k are zero-filled.
#include <cstdint>

#include "gemm.h"

namespace tara {

typedef short __attribute__((ext_vector_type(4))) v4s;
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

constexpr int GBM = 64, GBN = 64, GBK = 32;
constexpr int GPAD = 8;  // LDS row padding (elements)

template <typename T, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
    __shared__ T As[GBM][GBK + GPAD];
    __shared__ T Bs[GBN][GBK + GPAD];  // transposed: Bs[n][k]
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const int z1 = blockIdx.z % g.nb1, z2 = blockIdx.z / g.nb1;
    const T* A = reinterpret_cast<const T*>(g.a) + z1 * g.sa1 + z2 * g.sa2;
    const T* B = reinterpret_cast<const T*>(g.b) + z1 * g.sb1 + z2 * g.sb2;
    const long m0 = (long)blockIdx.y * GBM, n0 = (long)blockIdx.x * GBN;
    const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
    v16f acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    // Lanes walk the operand's unit-stride dimension when it has one, so that
    // each fetch instruction reads contiguous memory.
    const bool a_kfast = g.sak == 1, b_kfast = g.sbk == 1;
    for (long k0 = 0; k0 < g.k; k0 += GBK) {
        // stage A (64 x 32) and B (32 x 64): 8 elements per thread each
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int e = tid + r * 256;
            const int am = a_kfast ? e / GBK : e % GBM, ak = a_kfast ? e % GBK : e / GBM;
            const long gm = m0 + am, gk = k0 + ak;
            T va = T(0);
            if (gm < g.m && gk < g.k) va = A[gm * g.sam + gk * g.sak];
            As[am][ak] = va;
            const int bn = b_kfast ? e / GBK : e % GBN, bk = b_kfast ? e % GBK : e / GBN;
            const long gn = n0 + bn, gk2 = k0 + bk;
            T vb = T(0);
            if (gn < g.n && gk2 < g.k) vb = B[gk2 * g.sbk + gn * g.sbn];
            Bs[bn][bk] = vb;
        }
        __syncthreads();
        if constexpr (sizeof(T) == 2) {
            // 32x32x8 bf16: lane l supplies row/col l % 32, k = 4 (l / 32) .. +3
#pragma unroll
            for (int kk = 0; kk < GBK; kk += 8) {
                const int ko = kk + 4 * (lane / 32);
                v4s a, b;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    a[q] = (short)As[wm + lane % 32][ko + q];
                    b[q] = (short)Bs[wn + lane % 32][ko + q];
                }
                acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
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
    // result register i of lane l: row 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32
    const long gn = n0 + wn + lane % 32;
    if (gn >= g.n) return;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const long gm = m0 + wm + 8 * (i / 4) + 4 * (lane / 32) + (i % 4);
        if (gm >= g.m) continue;
        const long off = z1 * g.sc1 + z2 * g.sc2 + gm * g.scm + gn * g.scn;
        const float v = acc[i] * g.alpha;
        if constexpr (EPI == GEMM_STORE_F32) {
            reinterpret_cast<float*>(g.c)[off] = v;
        } else if constexpr (EPI == GEMM_ACCUM_F32) {
            reinterpret_cast<float*>(g.c)[off] += v;
        } else {
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

int gemm(const GemmArgs& g, int dtype, int epi, hipStream_t st)
{
    TARA_CHECK(g.m > 0 && g.n > 0 && g.k > 0 && g.nb1 > 0 && g.nb2 > 0, TARA_ERR_ARG, "gemm: bad shape");
    TARA_CHECK((long)g.nb1 * g.nb2 <= 65535, TARA_ERR_ARG, "gemm: too many batches");
    dim3 grid((unsigned)((g.n + GBN - 1) / GBN), (unsigned)((g.m + GBM - 1) / GBM), (unsigned)(g.nb1 * g.nb2));
    TARA_CHECK(grid.y <= 65535, TARA_ERR_ARG, "gemm: m too large");
    if (dtype == GEMM_BF16) {
        switch (epi) {
            case GEMM_STORE_F32: hipLaunchKernelGGL((k_gemm<bf16raw, GEMM_STORE_F32>), grid, dim3(256), 0, st, g); break;
            case GEMM_ACCUM_F32: hipLaunchKernelGGL((k_gemm<bf16raw, GEMM_ACCUM_F32>), grid, dim3(256), 0, st, g); break;
            case GEMM_STORE_BF16: hipLaunchKernelGGL((k_gemm<bf16raw, GEMM_STORE_BF16>), grid, dim3(256), 0, st, g); break;
            default: set_error("gemm: bad epilogue %d", epi); return TARA_ERR_ARG;
        }
    } else {
        switch (epi) {
            case GEMM_STORE_F32: hipLaunchKernelGGL((k_gemm<float, GEMM_STORE_F32>), grid, dim3(256), 0, st, g); break;
            case GEMM_ACCUM_F32: hipLaunchKernelGGL((k_gemm<float, GEMM_ACCUM_F32>), grid, dim3(256), 0, st, g); break;
            case GEMM_STORE_BF16: hipLaunchKernelGGL((k_gemm<float, GEMM_STORE_BF16>), grid, dim3(256), 0, st, g); break;
            default: set_error("gemm: bad epilogue %d", epi); return TARA_ERR_ARG;
        }
    }
    TARA_HIP(hipGetLastError());
    return TARA_OK;
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
