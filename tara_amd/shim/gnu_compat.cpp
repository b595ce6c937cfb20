// GNU Fortran runtime extensions that the reference's main programs call and
// flang's runtime lacks: RAND (src/main2csh.f95:94-95, tracer start positions)
// and SRAND (src/main2coh.f95:43, main2com.f95:57, main3com.f95:71).  Neither
// touches the solver's fields.
//
// The generator is Park and Miller's "minimal standard" multiplicative
// congruential generator (Comm. ACM 31(10), 1988): x <- 16807 x mod (2^31 - 1),
// seed 1 until SRAND sets one, RAND returning (x - 1) / (2^31 - 2) in [0, 1)
// as a default real.  Only the argument-less form RAND() is provided.
#include <cstdint>

namespace {
std::uint32_t g_state = 1;
constexpr std::uint64_t kA = 16807, kM = 2147483647;  // 2^31 - 1
}  // namespace

extern "C" float rand_(void)
{
    g_state = (std::uint32_t)((kA * g_state) % kM);
    return (float)((double)(g_state - 1) / (double)(kM - 1));
}

extern "C" void srand_(const int* seed)
{
    std::int64_t s = seed ? *seed : 0;
    s %= (std::int64_t)kM;
    if (s <= 0) s += (std::int64_t)kM - 1;  // the generator's state must lie in [1, 2^31 - 2]
    g_state = (std::uint32_t)s;
}
