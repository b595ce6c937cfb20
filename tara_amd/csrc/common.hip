// Host plumbing: error strings, device buffers, cached per-device tables.
#include <cmath>
#include <cstdarg>
#include <cstring>

#include "common.h"
#include "fft_core.h"

namespace tara {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int DevBuf::alloc(size_t n)
{
    release();
    if (n == 0) return TARA_OK;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        p = nullptr;
        set_error("hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
        return TARA_ERR_ALLOC;
    }
    bytes = n;
    return TARA_OK;
}

void DevBuf::release()
{
    if (p && own) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    own = true;
}

int current_device()
{
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) return -1;
    return d;
}

// e^{-2 pi i num/den} from long-double libm, reduced to the first octant so the
// argument passed to sinl/cosl is small.
double2 unit_root(long num, long den)
{
    num %= den;
    if (num < 0) num += den;
    // angle = 2 pi num/den in [0, 2pi); use symmetries on 8*num vs den.
    long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)num / (long double)den;
    long double c = cosl(a), s = sinl(a);
    // exact special points
    long e8 = 8 * num;
    if (e8 % den == 0) {
        long oct = e8 / den;  // multiple of pi/4
        static const long double h = 0.707106781186547524400844362104849039L;
        static const long double cs[8] = {1, h, 0, -h, -1, -h, 0, h};
        static const long double sn[8] = {0, h, 1, h, 0, -h, -1, -h};
        c = cs[oct];
        s = sn[oct];
    }
    return make_double2((double)c, (double)(-s));
}

void host_twiddles(int n, int rmax, std::vector<double2>& out)
{
    out.assign(std::max(1, fft_tw_size(n, rmax)), make_double2(0.0, 0.0));
    int span = 1, off = 0;
    while (span < n) {
        int r = fft_radix(n / span, rmax);
        for (int k = 0; k < span; ++k)
            for (int s = 1; s < r; ++s) out[off + k * (r - 1) + (s - 1)] = unit_root((long)k * s, (long)span * r);
        off += span * (r - 1);
        span *= r;
    }
}

void host_rc_twiddles(int n, std::vector<double2>& out)
{
    out.resize(n / 2 + 1);
    for (int k = 0; k <= n / 2; ++k) out[k] = unit_root(k, n);
}

namespace {
std::mutex g_tab_mu;
std::map<std::tuple<int, int, int>, void*> g_tabs;  // (device, kind, n) -> device ptr

// kind: 0 = Stockham table with radix cap 8, 1 = real<->complex packing roots,
// 2 = all n roots e^{-2 pi i k/n}, 4 = Stockham table with radix cap 4,
// 16 = Stockham table with radix cap 16.
const double2* cached_table(int kind, int n)
{
    int dev = current_device();
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto key = std::make_tuple(dev, kind, n);
    auto it = g_tabs.find(key);
    if (it != g_tabs.end()) return static_cast<const double2*>(it->second);
    std::vector<double2> h;
    if (kind == 1)
        host_rc_twiddles(n, h);
    else if (kind == 2) {
        h.resize(n);
        for (int k = 0; k < n; ++k) h[k] = unit_root(k, n);
    } else
        host_twiddles(n, kind == 4 ? 4 : kind == 16 ? 16 : 8, h);
    void* d = nullptr;
    if (hipMalloc(&d, h.size() * sizeof(double2)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), h.size() * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    g_tabs[key] = d;
    return static_cast<const double2*>(d);
}
}  // namespace

const double2* twiddles(int n, int rmax) { return cached_table(rmax == 4 ? 4 : rmax == 16 ? 16 : 0, n); }
const double2* rc_twiddles(int n) { return cached_table(1, n); }
const double2* unit_roots(int n) { return cached_table(2, n); }

}  // namespace tara

extern "C" const char* tara_last_error(void) { return tara::g_err; }
extern "C" int tara_version(void) { return 100; }
