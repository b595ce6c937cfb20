// Batched strided MFMA GEMM (gemm.hip).
#pragma once
#include "common.h"

namespace tara {

enum GemmType { GEMM_BF16 = 0, GEMM_F32 = 1 };
// Epilogues: r = alpha * sum (+ C[off] for ACCUM and ADD_TO) (+ bias[m]) -> act -> stored.
//   STORE_F32 / ACCUM_F32: C[off] = r (float); STORE_BF16: C[off] = bf16(r);
//   ADD_TO: C is only read; r goes to the optional outputs below.
// Optional outputs at the same offset: pre (float, value before the activation),
// c_f32 (float), c_bf16 (bfloat16 copy).
enum GemmEpi { GEMM_STORE_F32 = 0, GEMM_ACCUM_F32 = 1, GEMM_STORE_BF16 = 2, GEMM_ADD_TO = 3 };

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
    const float* bias = nullptr;  // per row m
    int act = 0;                  // 1: GELU (erf form)
    float* pre = nullptr;
    float* c_f32 = nullptr;
    void* c_bf16 = nullptr;
};

// C[z](m, n) (=, +=) alpha * sum_k A[z](m, k) B[z](k, n) over z = (z1 < nb1, z2 < nb2).
int gemm(const GemmArgs& g, int dtype, int epi, hipStream_t st);

}  // namespace tara
