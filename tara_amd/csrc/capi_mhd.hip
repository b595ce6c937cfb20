// C ABI of the 3D MHD path: drop-in derive3com / ab3com on host arrays, and
// the device-resident stepping context.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "fft_api.h"
#include "mhd3d.h"

namespace tara {
namespace {

// live state fields, in the reference's argument order minus E (index 4)
constexpr int kLive[7] = {0, 1, 2, 3, 5, 6, 7};

int check_dims(int nx, int ny, int nz, int nh)
{
    TARA_CHECK(nh == nx / 2 + 1, TARA_ERR_ARG, "3D MHD: Nh=%d but Nx/2+1=%d", nh, nx / 2 + 1);
    return TARA_OK;
}

// host (nh, ny, nz) contiguous <-> device (ld, ny, nz)
int h2d(const MhdPlan& p, double2* d, const double* h, hipStream_t s)
{
    TARA_HIP(hipMemcpy2DAsync(d, p.ld * sizeof(double2), h, p.nh * sizeof(double2), p.nh * sizeof(double2),
                              (size_t)p.ny * p.nz, hipMemcpyHostToDevice, s));
    return TARA_OK;
}
int d2h(const MhdPlan& p, double* h, const double2* d, hipStream_t s)
{
    TARA_HIP(hipMemcpy2DAsync(h, p.nh * sizeof(double2), d, p.ld * sizeof(double2), p.nh * sizeof(double2),
                              (size_t)p.ny * p.nz, hipMemcpyDeviceToHost, s));
    return TARA_OK;
}

struct DropIn {
    MhdPlan plan;
    DevBuf st[7], out[7];
};
std::mutex g_mu;
std::map<std::tuple<int, int, int, int, double, double, double>, std::unique_ptr<DropIn>> g_dropin;

int get_dropin(int nx, int ny, int nz, const tara_mhd_params* prm, DropIn** out)
{
    auto key = std::make_tuple(current_device(), nx, ny, nz, prm->lx, prm->ly, prm->lz);
    auto it = g_dropin.find(key);
    if (it == g_dropin.end()) {
        auto d = std::make_unique<DropIn>();
        int rc = mhd_plan_init(d->plan, nx, ny, nz, prm->lx, prm->ly, prm->lz, prm->mu, prm->eta, prm->cs);
        if (rc) return rc;
        const size_t bytes = d->plan.fsz * sizeof(double2);
        for (int f = 0; f < 7; ++f) {
            if ((rc = d->st[f].alloc(bytes)) || (rc = d->out[f].alloc(bytes))) return rc;
            TARA_HIP(hipMemset(d->st[f].p, 0, bytes));
        }
        it = g_dropin.emplace(key, std::move(d)).first;
    }
    it->second->plan.mu = prm->mu;
    it->second->plan.eta = prm->eta;
    it->second->plan.cs = prm->cs;
    *out = it->second.get();
    return TARA_OK;
}

}  // namespace
}  // namespace tara

using namespace tara;

extern "C" int tara_derive3com(int nx, int ny, int nz, int nh, const tara_mhd_params* prm,
                               const double* const state[8], double* const rhs[8])
{
    TARA_CHECK(prm && state && rhs, TARA_ERR_ARG, "derive3com: null argument");
    for (int f = 0; f < 8; ++f) TARA_CHECK(state[f] && rhs[f], TARA_ERR_ARG, "derive3com: null array %d", f);
    int rc = check_dims(nx, ny, nz, nh);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_mu);
    DropIn* d = nullptr;
    if ((rc = get_dropin(nx, ny, nz, prm, &d))) return rc;
    const MhdPlan& p = d->plan;
    double2 *st[7], *out[7];
    for (int f = 0; f < 7; ++f) {
        st[f] = d->st[f].as<double2>();
        out[f] = d->out[f].as<double2>();
        if ((rc = h2d(p, st[f], state[kLive[f]], nullptr))) return rc;
    }
    if ((rc = mhd_rhs_stage(p, st, MHD_RHS, 0.0, out, nullptr, nullptr, nullptr))) return rc;
    for (int f = 0; f < 7; ++f)
        if ((rc = d2h(p, rhs[kLive[f]], out[f], nullptr))) return rc;
    TARA_HIP(hipStreamSynchronize(nullptr));
    std::memset(rhs[4], 0, (size_t)nh * ny * nz * 2 * sizeof(double));
    return TARA_OK;
}

extern "C" int tara_ab3com(int nx, int ny, int nz, int nh, double dt, const double* const state[8],
                           double* const state_new[8], const double* const f_old[8], const double* const f_new[8])
{
    TARA_CHECK(state && state_new && f_old && f_new, TARA_ERR_ARG, "ab3com: null argument");
    int rc = check_dims(nx, ny, nz, nh);
    if (rc) return rc;
    const size_t n = (size_t)nh * ny * nz, bytes = n * sizeof(double2);
    DevBuf w, wn, fo, fn;
    if ((rc = w.alloc(bytes)) || (rc = wn.alloc(bytes)) || (rc = fo.alloc(bytes)) || (rc = fn.alloc(bytes)))
        return rc;
    for (int f = 0; f < 8; ++f) {
        TARA_CHECK(state[f] && state_new[f] && f_old[f] && f_new[f], TARA_ERR_ARG, "ab3com: null array %d", f);
        if (f == 4) {  // energy frozen (src/ab3com.f95:51)
            std::memcpy(state_new[f], state[f], bytes);
            continue;
        }
        TARA_HIP(hipMemcpy(w.p, state[f], bytes, hipMemcpyHostToDevice));
        TARA_HIP(hipMemcpy(fo.p, f_old[f], bytes, hipMemcpyHostToDevice));
        TARA_HIP(hipMemcpy(fn.p, f_new[f], bytes, hipMemcpyHostToDevice));
        if ((rc = launch_ab2(w.as<double2>(), wn.as<double2>(), fo.as<double2>(), fn.as<double2>(), dt, n, nullptr)))
            return rc;
        TARA_HIP(hipMemcpy(state_new[f], wn.p, bytes, hipMemcpyDeviceToHost));
    }
    return TARA_OK;
}

struct tara_mhd {
    MhdPlan plan;
    DevBuf st[8], hist[7];
    hipStream_t s = nullptr;
    double dt = 0;
    std::unique_ptr<EventTimer> tm;
};

extern "C" int tara_mhd_create(int nx, int ny, int nz, const tara_mhd_params* prm, tara_mhd** out)
{
    TARA_CHECK(prm && out, TARA_ERR_ARG, "null argument");
    *out = nullptr;
    auto c = std::make_unique<tara_mhd>();
    int rc = mhd_plan_init(c->plan, nx, ny, nz, prm->lx, prm->ly, prm->lz, prm->mu, prm->eta, prm->cs);
    if (rc) return rc;
    const size_t bytes = c->plan.fsz * sizeof(double2);
    for (int f = 0; f < 8; ++f) {
        if ((rc = c->st[f].alloc(bytes))) return rc;
        TARA_HIP(hipMemset(c->st[f].p, 0, bytes));
    }
    for (int f = 0; f < 7; ++f) {
        if ((rc = c->hist[f].alloc(bytes))) return rc;
        TARA_HIP(hipMemset(c->hist[f].p, 0, bytes));
    }
    TARA_HIP(hipStreamCreate(&c->s));
    c->dt = prm->dt;
    *out = c.release();
    return TARA_OK;
}

extern "C" int tara_mhd_upload(tara_mhd* c, const double* const state[8])
{
    TARA_CHECK(c && state, TARA_ERR_ARG, "null argument");
    int rc;
    for (int f = 0; f < 8; ++f) {
        TARA_CHECK(state[f], TARA_ERR_ARG, "null array %d", f);
        if ((rc = h2d(c->plan, c->st[f].as<double2>(), state[f], c->s))) return rc;
    }
    for (int f = 0; f < 7; ++f) TARA_HIP(hipMemsetAsync(c->hist[f].p, 0, c->hist[f].bytes, c->s));
    TARA_HIP(hipStreamSynchronize(c->s));
    return TARA_OK;
}

extern "C" int tara_mhd_download(tara_mhd* c, double* const state[8])
{
    TARA_CHECK(c && state, TARA_ERR_ARG, "null argument");
    int rc;
    for (int f = 0; f < 8; ++f) {
        TARA_CHECK(state[f], TARA_ERR_ARG, "null array %d", f);
        if ((rc = d2h(c->plan, state[f], c->st[f].as<double2>(), c->s))) return rc;
    }
    TARA_HIP(hipStreamSynchronize(c->s));
    return TARA_OK;
}

extern "C" int tara_mhd_advance(tara_mhd* c, int nsteps)
{
    TARA_CHECK(c && nsteps >= 0, TARA_ERR_ARG, "bad argument");
    double2 *st[7], *h[7];
    for (int f = 0; f < 7; ++f) {
        st[f] = c->st[kLive[f]].as<double2>();
        h[f] = c->hist[f].as<double2>();
    }
    for (int k = 0; k < nsteps; ++k) {
        int rc = mhd_rhs_stage(c->plan, st, MHD_AB2, c->dt, nullptr, h, c->s, c->tm.get());
        if (rc) return rc;
    }
    return TARA_OK;
}

extern "C" int tara_mhd_sync(tara_mhd* c)
{
    TARA_CHECK(c, TARA_ERR_ARG, "null context");
    TARA_HIP(hipStreamSynchronize(c->s));
    return TARA_OK;
}

namespace {

// Field f (reference order rho, rho ux, rho uy, rho uz, E, Bx, By, Bz) of the
// ABC-flow state of src/main3com.f95:290-324 on the grid x(i) = (i-1) dx:
// uniform rho0 and P0, u = U0 (A sin kz + C cos ky, B sin kx + A cos kz,
// C sin ky + B cos kx) with wavenumber kf, uniform B = B0 on every axis,
// E = P + rho u^2/2 + B^2/2.
__global__ void k_abc_field(double* out, int f, int nx, int ny, int nz, double dx, double dy, double dz,
                            double rho0, double u0, double b0, double p0, double a, double b, double cc, double kf)
{
    const size_t n = (size_t)nx * ny * nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((size_t)nx * ny));
        const double x = i * dx, y = j * dy, z = k * dz;
        const double ux = u0 * (a * sin(kf * z) + cc * cos(kf * y));
        const double uy = u0 * (b * sin(kf * x) + a * cos(kf * z));
        const double uz = u0 * (cc * sin(kf * y) + b * cos(kf * x));
        double v;
        switch (f) {
            case 0: v = rho0; break;
            case 1: v = rho0 * ux; break;
            case 2: v = rho0 * uy; break;
            case 3: v = rho0 * uz; break;
            case 4: v = p0 + 0.5 * rho0 * (ux * ux + uy * uy + uz * uz) + 0.5 * (3.0 * b0 * b0); break;
            default: v = b0; break;
        }
        out[q] = v;
    }
}

}  // namespace

extern "C" int tara_mhd_ic_abc(tara_mhd* c, double rho0, double u0, double b0, double p0, double a, double b,
                               double cc, double kf)
{
    TARA_CHECK(c, TARA_ERR_ARG, "null context");
    const MhdPlan& p = c->plan;
    const size_t n = (size_t)p.nx * p.ny * p.nz;
    DevBuf real;
    int rc;
    if ((rc = real.alloc(n * sizeof(double)))) return rc;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    for (int f = 0; f < 8; ++f) {
        double2* s = c->st[f].as<double2>();
        TARA_HIP(hipMemsetAsync(s, 0, c->st[f].bytes, c->s));
        hipLaunchKernelGGL(k_abc_field, dim3((unsigned)blocks), dim3(256), 0, c->s, real.as<double>(), f, p.nx,
                           p.ny, p.nz, p.lx / p.nx, p.ly / p.ny, p.lz / p.nz, rho0, u0, b0, p0, a, b, cc, kf);
        TARA_HIP(hipGetLastError());
        if ((rc = row_r2c(real.as<double>(), s, p.nx, p.ld, p.ny * p.nz, c->s))) return rc;
        if ((rc = col_fft(s, p.ny, p.ld, p.ld, (long)p.ld * p.ny, p.nz, -1, c->s))) return rc;
        if ((rc = col_fft(s, p.nz, p.ld, (long)p.ld * p.ny, p.ld, p.ny, -1, c->s))) return rc;
    }
    for (int f = 0; f < 7; ++f) TARA_HIP(hipMemsetAsync(c->hist[f].p, 0, c->hist[f].bytes, c->s));
    TARA_HIP(hipStreamSynchronize(c->s));
    return TARA_OK;
}

extern "C" int tara_mhd_destroy(tara_mhd* c)
{
    if (!c) return TARA_OK;
    if (c->s) {
        (void)hipStreamSynchronize(c->s);
        c->tm.reset();
        (void)hipStreamDestroy(c->s);
    }
    delete c;
    return TARA_OK;
}

extern "C" int tara_mhd_set_timing(tara_mhd* c, int on)
{
    TARA_CHECK(c, TARA_ERR_ARG, "null context");
    if (on && !c->tm) c->tm = std::make_unique<EventTimer>();
    if (!on) c->tm.reset();
    return TARA_OK;
}

extern "C" int tara_mhd_kernel_stats(tara_mhd* c, int nclass, long* counts, double* ms)
{
    TARA_CHECK(c && c->tm, TARA_ERR_STATE, "timing not enabled");
    int rc = c->tm->collect();
    if (rc) return rc;
    for (int k = 0; k < nclass && k < EventTimer::kClasses; ++k) {
        if (counts) counts[k] = c->tm->counts[k];
        if (ms) ms[k] = c->tm->ms[k];
        c->tm->counts[k] = 0;
        c->tm->ms[k] = 0;
    }
    return TARA_OK;
}
