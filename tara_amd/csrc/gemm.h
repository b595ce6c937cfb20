// Batched strided MFMA GEMM (gemm.hip).
#pragma once
#include "common.h"

namespace tara {

enum GemmType { GEMM_BF16 = 0, GEMM_F32 = 1 };
enum GemmEpi { GEMM_STORE_F32 = 0, GEMM_ACCUM_F32 = 1, GEMM_STORE_BF16 = 2 };

struct GemmArgs {
    long m = 0, n = 0, k = 0;
    const void* a = nullptr;
    long sam = 0, sak = 0;
    const void* b = nullptr;
    long sbk = 0, sbn = 0;
    void* c = nullptr;
    long scm = 0, scn = 0;
    int nb1 = 1;
    long sa1 = 0, sb1 = 0, sc1 = 0;
    int nb2 = 1;
    long sa2 = 0, sb2 = 0, sc2 = 0;
    float alpha = 1.f;
};

// C[z](m, n) (=, +=) alpha * sum_k A[z](m, k) B[z](k, n) over z = (z1 < nb1, z2 < nb2).
int gemm(const GemmArgs& g, int dtype, int epi, hipStream_t st);

}  // namespace tara
