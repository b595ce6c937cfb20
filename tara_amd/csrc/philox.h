// Philox-4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11): multipliers 0xD2511F53,
// 0xCD9E8D57, Weyl key increments 0x9E3779B9, 0xBB67AE85 (the published
// constants of the method).  A draw depends only on (counter, key), so every
// mode's random phase is fixed by (seed, mode index), whatever the launch shape.
#pragma once
#include <hip/hip_runtime.h>

namespace tara {

struct U4 {
    unsigned x, y, z, w;
};

__host__ __device__ inline void mulhilo(unsigned a, unsigned b, unsigned* hi, unsigned* lo)
{
    const unsigned long long p = (unsigned long long)a * b;
    *hi = (unsigned)(p >> 32);
    *lo = (unsigned)p;
}

__host__ __device__ inline U4 philox4x32_10(U4 c, unsigned k0, unsigned k1)
{
    for (int r = 0; r < 10; ++r) {
        unsigned h0, l0, h1, l1;
        mulhilo(0xD2511F53u, c.x, &h0, &l0);
        mulhilo(0xCD9E8D57u, c.z, &h1, &l1);
        c = U4{h1 ^ c.y ^ k0, l1, h0 ^ c.w ^ k1, l0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Uniform in [0, 1) with 53 bits from two 32-bit words.
__host__ __device__ inline double u53(unsigned a, unsigned b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

}  // namespace tara
