// Host-side plumbing shared by every translation unit of libtara_hip.so:
// error reporting across the C ABI, per-device cached tables (FFT twiddles,
// wavenumbers, dealias bounds) and small launch helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "tara_hip.h"

namespace tara {

// Last error message, per thread.
void set_error(const char* fmt, ...);

// Turns a HIP error into a negative status and records the message.
#define TARA_HIP(expr)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            ::tara::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
            return TARA_ERR_HIP;                                                              \
        }                                                                                     \
    } while (0)

#define TARA_CHECK(cond, code, ...)      \
    do {                                 \
        if (!(cond)) {                   \
            ::tara::set_error(__VA_ARGS__); \
            return (code);               \
        }                                \
    } while (0)

inline bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
inline int ilog2(int n)
{
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

// Device buffer owned by RAII (hipMalloc / hipFree), or a view of memory
// owned elsewhere (view(): never freed here).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    bool own = true;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    int alloc(size_t n);
    void view(void* q, size_t n)
    {
        release();
        p = q;
        bytes = n;
        own = false;
    }
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Per-device immutable tables, built once and kept for the process lifetime.
//   twiddles(n, rmax): Stockham per-pass table for a length-n complex FFT whose
//                      schedule caps the radix at rmax (8 or 4; fft_core.h)
//   rc_twiddles(n)   : e^{-2 pi i k/n}, k = 0..n/2 (real <-> half-length complex packing)
//   unit_roots(n)    : e^{-2 pi i k/n}, k = 0..n-1 (any n; the mixed-radix FFT)
const double2* twiddles(int n, int rmax = 8);
const double2* rc_twiddles(int n);
const double2* unit_roots(int n);
// e^{-2 pi i num/den} in long double, rounded once (exact at multiples of pi/4).
double2 unit_root(long num, long den);

// Host-side table generators.
void host_twiddles(int n, int rmax, std::vector<double2>& out);
void host_rc_twiddles(int n, std::vector<double2>& out);

int current_device();

}  // namespace tara
