// derive2csh on device arrays with no device allocation: for callers that
// cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  Every intermediate of the derive2csh
// pipeline (the ky-major input and output copies, G0..G3, H and the FFT /
// wavenumber / dealias tables) lives in one file-scope device array, whose
// address is taken once per device with hipGetSymbolAddress; the kernels are
// those of tara_derive2csh, with the same tables and x-pass shape, so the
// results are the same bit for bit.
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"
#include "csh2d.h"
#include "fft_core.h"
#include "layout.h"

namespace tara {
namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per device
constexpr size_t kTableBytes = (size_t)1 << 20;     // first MiB: tables
constexpr int kMaxDevices = 64;

// The scratch array.  In host code this name is only a stand-in for the
// symbol: its device address (one per device) comes from hipGetSymbolAddress
// and is the only form in which the memory is used.
static __device__ __align__(256) double2 scratch[kScratchBytes / sizeof(double2)];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};            // scratch's address on each device
int g_tab_nx[kMaxDevices] = {}, g_tab_ny[kMaxDevices] = {};  // grid whose tables the scratch holds

size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte offsets of the regions of one (nx, ny) grid in the scratch.
struct Layout {
    size_t kx, ky, imin, tw_y, tw_x, tw_xn, tw_x16, tw_xn16, rctw, tables_end;
    size_t w, g[4], h, f, end;
};

Layout layout_of(int nx, int ny)
{
    const int nh = nx / 2 + 1;
    const size_t ld = (size_t)csh_gld(nh);
    Layout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = up256(o + bytes);
        return at;
    };
    L.kx = take(nh * sizeof(double));
    L.ky = take(ny * sizeof(double));
    L.imin = take(ny * sizeof(int));
    L.tw_y = take(std::max(1, fft_tw_size(ny, csh_yem(ny))) * sizeof(double2));
    L.tw_x = take(std::max(1, fft_tw_size(nx / 2, 8)) * sizeof(double2));
    L.tw_xn = take(std::max(1, fft_tw_size(nx, 8)) * sizeof(double2));
    L.tw_x16 = take(std::max(1, fft_tw_size(nx / 2, 16)) * sizeof(double2));
    L.tw_xn16 = take(std::max(1, fft_tw_size(nx, 16)) * sizeof(double2));
    L.rctw = take((nx / 2 + 1) * sizeof(double2));
    L.tables_end = o;
    o = kTableBytes;
    L.w = take((size_t)nh * ny * sizeof(double2));
    for (int k = 0; k < 4; ++k) L.g[k] = take(ld * ny * sizeof(double2));
    L.h = take(ld * ny * sizeof(double2));
    L.f = take((size_t)nh * ny * sizeof(double2));
    L.end = o;
    return L;
}

// The scratch's device address on the current device (taken once per device).
int scratch_ptr(int dev, char** out)
{
    TARA_CHECK(dev >= 0 && dev < kMaxDevices, TARA_ERR_UNSUPPORTED, "device %d out of range", dev);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        TARA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        TARA_CHECK(p, TARA_ERR_HIP, "no device address for the scratch array");
        g_ptr[dev] = p;
    }
    *out = static_cast<char*>(g_ptr[dev]);
    return TARA_OK;
}

// Writes the tables of (nx, ny) into the scratch when it holds another grid's.
int put_tables(int dev, char* base, const Layout& L, int nx, int ny, hipStream_t st)
{
    if (g_tab_nx[dev] == nx && g_tab_ny[dev] == ny) return TARA_OK;
    std::vector<double> kx, ky;
    std::vector<int> imin;
    int rc = csh_host_tables(nx, ny, kx, ky, imin);
    if (rc) return rc;
    std::vector<double2> t;
    // earlier calls on the stream may still read the old tables
    TARA_HIP(hipStreamSynchronize(st));
    TARA_HIP(hipMemcpy(base + L.kx, kx.data(), kx.size() * sizeof(double), hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(base + L.ky, ky.data(), ky.size() * sizeof(double), hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(base + L.imin, imin.data(), imin.size() * sizeof(int), hipMemcpyHostToDevice));
    const struct {
        size_t off;
        int n, rmax;
    } tws[] = {{L.tw_y, ny, csh_yem(ny)}, {L.tw_x, nx / 2, 8}, {L.tw_xn, nx, 8}, {L.tw_x16, nx / 2, 16}, {L.tw_xn16, nx, 16}};
    for (const auto& w : tws) {
        host_twiddles(w.n, w.rmax, t);
        TARA_HIP(hipMemcpy(base + w.off, t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice));
    }
    host_rc_twiddles(nx, t);
    TARA_HIP(hipMemcpy(base + L.rctw, t.data(), t.size() * sizeof(double2), hipMemcpyHostToDevice));
    g_tab_nx[dev] = nx;
    g_tab_ny[dev] = ny;
    return TARA_OK;
}

}  // namespace
}  // namespace tara

using namespace tara;

extern "C" int tara_derive2csh_noalloc(int nx, int ny, int nh, double nu, const double* omegak,
                                       double* dt_omegak_new, void* stream)
{
    TARA_CHECK(omegak && dt_omegak_new, TARA_ERR_ARG, "derive2csh_noalloc: null array");
    TARA_CHECK(nh == nx / 2 + 1, TARA_ERR_ARG, "derive2csh_noalloc: Nh=%d but Nx/2+1=%d", nh, nx / 2 + 1);
    TARA_CHECK(is_pow2(nx) && is_pow2(ny) && nx >= 32 && ny >= 16 && nx <= 1024 && ny <= 1024, TARA_ERR_UNSUPPORTED,
               "derive2csh_noalloc: grid %dx%d unsupported (powers of two, 32..1024 x 16..1024)", nx, ny);
    const Layout L = layout_of(nx, ny);
    TARA_CHECK(L.tables_end <= kTableBytes && L.end <= kScratchBytes, TARA_ERR_UNSUPPORTED,
               "derive2csh_noalloc: grid %dx%d needs %zu bytes of scratch, more than its %zu", nx, ny, L.end,
               kScratchBytes);
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(g_mu);
    const int dev = current_device();
    char* base = nullptr;
    int rc;
    if ((rc = scratch_ptr(dev, &base)) || (rc = put_tables(dev, base, L, nx, ny, st))) return rc;
    // a plan whose buffers and tables are views of the scratch
    CshPlan p;
    p.nx = nx;
    p.ny = ny;
    p.nh = nh;
    p.batch = 1;
    p.gld = csh_gld(nh);
    p.gsh = csh_gshift(nh);
    p.bstride = (size_t)p.gld * ny;
    const size_t fb = p.bstride * sizeof(double2);
    for (int k = 0; k < 4; ++k) p.g[k].view(base + L.g[k], fb);
    p.h.view(base + L.h, fb);
    p.kx.view(base + L.kx, nh * sizeof(double));
    p.ky.view(base + L.ky, ny * sizeof(double));
    p.imin.view(base + L.imin, ny * sizeof(int));
    p.tw_y = reinterpret_cast<const double2*>(base + L.tw_y);
    p.tw_x = reinterpret_cast<const double2*>(base + L.tw_x);
    p.tw_xn = reinterpret_cast<const double2*>(base + L.tw_xn);
    p.tw_x16 = reinterpret_cast<const double2*>(base + L.tw_x16);
    p.tw_xn16 = reinterpret_cast<const double2*>(base + L.tw_xn16);
    p.rctw = reinterpret_cast<const double2*>(base + L.rctw);
    p.xcfg = csh_default_xshape(nx);  // the drop-in plan's shape (csh_plan_init)
    if (!csh_xshape_ok(p, p.xcfg)) p.xcfg = 0;
    double2* w = reinterpret_cast<double2*>(base + L.w);
    double2* f = reinterpret_cast<double2*>(base + L.f);
    CshPhys ph;
    ph.nu = nu;
    if ((rc = to_kymajor(reinterpret_cast<const double2*>(omegak), w, nh, ny, 1, st))) return rc;
    if ((rc = csh_rhs_stage(p, w, CSH_RHS, ph, 0.0, nullptr, nullptr, f, st, nullptr))) return rc;
    if ((rc = from_kymajor(f, reinterpret_cast<double2*>(dt_omegak_new), nh, ny, 1, st))) return rc;
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}
