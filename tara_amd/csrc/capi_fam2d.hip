// C ABI of the 2D compressible families: drop-in derive2coh / ab2coh /
// euler2coh / rk4_*2coh / pc_*2coh and derive2com / ab2com on host arrays, and
// the device-resident AB2 stepping context shared by both families.
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "fam2d.h"
#include "generic_fam2d.h"

namespace tara {
namespace {

constexpr int kFields[2] = {3, 6};

int check_nh(const char* name, int nx, int nh)
{
    TARA_CHECK(nh == nx / 2 + 1, TARA_ERR_ARG, "%s: Nh=%d but Nx/2+1=%d", name, nh, nx / 2 + 1);
    return TARA_OK;
}

// host (nh, ny) contiguous <-> device (ld, ny)
int h2d(const F2Plan& p, double2* d, const double* h, hipStream_t s)
{
    TARA_HIP(hipMemcpy2DAsync(d, p.ld * sizeof(double2), h, p.nh * sizeof(double2), p.nh * sizeof(double2), p.ny,
                              hipMemcpyHostToDevice, s));
    return TARA_OK;
}
int d2h(const F2Plan& p, double* h, const double2* d, hipStream_t s)
{
    TARA_HIP(hipMemcpy2DAsync(h, p.nh * sizeof(double2), d, p.ld * sizeof(double2), p.nh * sizeof(double2), p.ny,
                              hipMemcpyDeviceToHost, s));
    return TARA_OK;
}

struct DropIn {
    F2Plan plan;
    DevBuf st[6], out[6];
};
std::mutex g_mu;
std::map<std::tuple<int, int, int, int, double, double, double>, std::unique_ptr<DropIn>> g_dropin;

int get_dropin(int family, int nx, int ny, double pi, double lx, double ly, DropIn** out)
{
    auto key = std::make_tuple(current_device(), family, nx, ny, pi, lx, ly);
    auto it = g_dropin.find(key);
    if (it == g_dropin.end()) {
        auto d = std::make_unique<DropIn>();
        int rc = f2_plan_init(d->plan, family, nx, ny, pi, lx, ly);
        if (rc) return rc;
        const size_t bytes = d->plan.fsz * sizeof(double2);
        for (int f = 0; f < kFields[family]; ++f) {
            if ((rc = d->st[f].alloc(bytes)) || (rc = d->out[f].alloc(bytes))) return rc;
            TARA_HIP(hipMemset(d->st[f].p, 0, bytes));
            TARA_HIP(hipMemset(d->out[f].p, 0, bytes));
        }
        it = g_dropin.emplace(key, std::move(d)).first;
    }
    *out = it->second.get();
    return TARA_OK;
}

// Pointwise updates on contiguous host arrays of n complex elements per field.
struct PointBufs {
    DevBuf b[4];
    size_t bytes = 0;
};
PointBufs g_pt;

int pointwise_host(int kind, int nf, size_t n, double dt, const double* const w[], double* const wn[],
                   const double* const fo[], const double* const fn[])
{
    const size_t bytes = n * sizeof(double2);
    if (g_pt.bytes < bytes) {
        for (auto& b : g_pt.b) {
            int rc = b.alloc(bytes);
            if (rc) return rc;
        }
        g_pt.bytes = bytes;
    }
    for (int f = 0; f < nf; ++f) {
        TARA_HIP(hipMemcpy(g_pt.b[0].p, w[f], bytes, hipMemcpyHostToDevice));
        TARA_HIP(hipMemcpy(g_pt.b[3].p, fn[f], bytes, hipMemcpyHostToDevice));
        if (fo) TARA_HIP(hipMemcpy(g_pt.b[2].p, fo[f], bytes, hipMemcpyHostToDevice));
        int rc = f2_pointwise(kind, g_pt.b[0].as<double2>(), g_pt.b[1].as<double2>(), g_pt.b[2].as<double2>(),
                              g_pt.b[3].as<double2>(), dt, n, nullptr);
        if (rc) return rc;
        TARA_HIP(hipMemcpy(wn[f], g_pt.b[1].p, bytes, hipMemcpyDeviceToHost));
    }
    return TARA_OK;
}

template <typename Fn>
int derive_host(int family, int nx, int ny, double pi, double lx, double ly, const double* const state[],
                double* const rhs[], Fn&& stage)
{
    const int nf = kFields[family];
    std::lock_guard<std::mutex> lk(g_mu);
    DropIn* d = nullptr;
    int rc = get_dropin(family, nx, ny, pi, lx, ly, &d);
    if (rc) return rc;
    const F2Plan& p = d->plan;
    double2* st[6];
    double2* out[6];
    for (int f = 0; f < nf; ++f) {
        st[f] = d->st[f].as<double2>();
        out[f] = d->out[f].as<double2>();
        if ((rc = h2d(p, st[f], state[f], nullptr))) return rc;
    }
    if ((rc = stage(p, st, out))) return rc;
    for (int f = 0; f < nf; ++f)
        if ((rc = d2h(p, rhs[f], out[f], nullptr))) return rc;
    TARA_HIP(hipDeviceSynchronize());
    return TARA_OK;
}

}  // namespace
}  // namespace tara

using namespace tara;

// ------------------------------------------------------------------ coh
extern "C" int tara_derive2coh(int nx, int ny, int nh, const tara_coh_params* prm, const double* const state[3],
                               double* const rhs[3])
{
    TARA_CHECK(prm && state && rhs, TARA_ERR_ARG, "derive2coh: null argument");
    for (int f = 0; f < 3; ++f) TARA_CHECK(state[f] && rhs[f], TARA_ERR_ARG, "derive2coh: null array %d", f);
    int rc = check_nh("derive2coh", nx, nh);
    if (rc) return rc;
    if (!f2_fast_grid(nx, ny)) return gen_derive2coh_host(nx, ny, prm, state, rhs);
    CohPhys ph;
    ph.mux = prm->mux;
    ph.muy = prm->muy;
    ph.cs = prm->cs;
    return derive_host(0, nx, ny, prm->pi, prm->lx, prm->ly, state, rhs,
                       [&](const F2Plan& p, double2* const st[], double2* const out[]) {
                           return coh_rhs_stage(p, ph, st, F2_RHS, 0.0, out, nullptr, nullptr);
                       });
}

extern "C" int tara_ab2coh(int nx, int ny, int nh, double dt, const double* const state[3],
                           double* const state_new[3], const double* const f_old[3], const double* const f_new[3])
{
    TARA_CHECK(state && state_new && f_old && f_new, TARA_ERR_ARG, "ab2coh: null argument");
    for (int f = 0; f < 3; ++f)
        TARA_CHECK(state[f] && state_new[f] && f_old[f] && f_new[f], TARA_ERR_ARG, "ab2coh: null array %d", f);
    TARA_CHECK(nx > 0 && ny > 0 && nh == nx / 2 + 1, TARA_ERR_ARG, "ab2coh: bad dims %d %d %d", nx, ny, nh);
    std::lock_guard<std::mutex> lk(g_mu);
    return pointwise_host(F2P_AB2, 3, (size_t)nh * ny, dt, state, state_new, f_old, f_new);
}

extern "C" int tara_euler2coh(int nx, int ny, int nh, double dt, const double* const state[3],
                              double* const state_new[3], const double* const slope[3])
{
    TARA_CHECK(state && state_new && slope, TARA_ERR_ARG, "euler2coh: null argument");
    for (int f = 0; f < 3; ++f)
        TARA_CHECK(state[f] && state_new[f] && slope[f], TARA_ERR_ARG, "euler2coh: null array %d", f);
    TARA_CHECK(nx > 0 && ny > 0 && nh == nx / 2 + 1, TARA_ERR_ARG, "euler2coh: bad dims %d %d %d", nx, ny, nh);
    std::lock_guard<std::mutex> lk(g_mu);
    return pointwise_host(F2P_EULER, 3, (size_t)nh * ny, dt, state, state_new, nullptr, slope);
}

static int coh_field(int kind, const char* name, int nx, int ny, int nh, double dt, const double* field,
                     double* field_new, const double* slope)
{
    TARA_CHECK(field && field_new && slope, TARA_ERR_ARG, "%s: null array", name);
    TARA_CHECK(nx > 0 && ny > 0 && nh == nx / 2 + 1, TARA_ERR_ARG, "%s: bad dims %d %d %d", name, nx, ny, nh);
    const double* w[1] = {field};
    double* wn[1] = {field_new};
    const double* fn[1] = {slope};
    std::lock_guard<std::mutex> lk(g_mu);
    return pointwise_host(kind, 1, (size_t)nh * ny, dt, w, wn, nullptr, fn);
}

extern "C" int tara_rk4coh_field(int nx, int ny, int nh, double dt, const double* field, double* field_new,
                                 const double* slope)
{
    return coh_field(F2P_RK4_DEGEN, "rk4_*2coh", nx, ny, nh, dt, field, field_new, slope);
}

extern "C" int tara_pccoh_field(int nx, int ny, int nh, double dt, const double* field, double* field_new,
                                const double* slope)
{
    return coh_field(F2P_PC_DEGEN, "pc_*2coh", nx, ny, nh, dt, field, field_new, slope);
}

// ------------------------------------------------------------------ com
extern "C" int tara_derive2com(int nx, int ny, int nh, const tara_com_params* prm, const double* const state[6],
                               double* const rhs[6])
{
    TARA_CHECK(prm && state && rhs, TARA_ERR_ARG, "derive2com: null argument");
    for (int f = 0; f < 6; ++f) TARA_CHECK(state[f] && rhs[f], TARA_ERR_ARG, "derive2com: null array %d", f);
    int rc = check_nh("derive2com", nx, nh);
    if (rc) return rc;
    if (!f2_fast_grid(nx, ny)) return gen_derive2com_host(nx, ny, prm, state, rhs);
    ComPhys ph;
    ph.mu = prm->mu;
    ph.eta = prm->eta;
    ph.cs = prm->cs;
    ph.kf = prm->kf;
    ph.amp = prm->a;
    return derive_host(1, nx, ny, prm->pi, prm->lx, prm->ly, state, rhs,
                       [&](const F2Plan& p, double2* const st[], double2* const out[]) {
                           return com_rhs_stage(p, ph, st, F2_RHS, 0.0, out, nullptr, nullptr);
                       });
}

extern "C" int tara_ab2com(int nx, int ny, int nh, double dt, const double* const state[6],
                           double* const state_new[6], const double* const f_old[6], const double* const f_new[6])
{
    TARA_CHECK(state && state_new && f_old && f_new, TARA_ERR_ARG, "ab2com: null argument");
    for (int f = 0; f < 6; ++f)
        TARA_CHECK(state[f] && state_new[f] && f_old[f] && f_new[f], TARA_ERR_ARG, "ab2com: null array %d", f);
    TARA_CHECK(nx > 0 && ny > 0 && nh == nx / 2 + 1, TARA_ERR_ARG, "ab2com: bad dims %d %d %d", nx, ny, nh);
    std::lock_guard<std::mutex> lk(g_mu);
    return pointwise_host(F2P_AB2, 6, (size_t)nh * ny, dt, state, state_new, f_old, f_new);
}

// ------------------------------------------------------- stepping context
struct tara_f2 {
    int family = 0;
    F2Plan plan;
    CohPhys coh;
    ComPhys com;
    double dt = 0;
    DevBuf st[6], hist[6];
    hipStream_t s = nullptr;
    std::unique_ptr<EventTimer> tm;
};

static int f2_create(int family, int nx, int ny, double pi, double lx, double ly, double dt, tara_f2** out,
                     tara_f2** made)
{
    TARA_CHECK(out, TARA_ERR_ARG, "null out");
    *out = nullptr;
    auto c = std::make_unique<tara_f2>();
    c->family = family;
    c->dt = dt;
    int rc = f2_plan_init(c->plan, family, nx, ny, pi, lx, ly);
    if (rc) return rc;
    const size_t bytes = c->plan.fsz * sizeof(double2);
    for (int f = 0; f < kFields[family]; ++f) {
        if ((rc = c->st[f].alloc(bytes)) || (rc = c->hist[f].alloc(bytes))) return rc;
        TARA_HIP(hipMemset(c->st[f].p, 0, bytes));
        TARA_HIP(hipMemset(c->hist[f].p, 0, bytes));
    }
    TARA_HIP(hipStreamCreate(&c->s));
    *made = c.get();
    *out = c.release();
    return TARA_OK;
}

extern "C" int tara_coh_create(int nx, int ny, const tara_coh_params* prm, tara_f2** out)
{
    TARA_CHECK(prm, TARA_ERR_ARG, "null params");
    tara_f2* c = nullptr;
    int rc = f2_create(0, nx, ny, prm->pi, prm->lx, prm->ly, prm->dt, out, &c);
    if (rc) return rc;
    c->coh.mux = prm->mux;
    c->coh.muy = prm->muy;
    c->coh.cs = prm->cs;
    return TARA_OK;
}

extern "C" int tara_com_create(int nx, int ny, const tara_com_params* prm, tara_f2** out)
{
    TARA_CHECK(prm, TARA_ERR_ARG, "null params");
    tara_f2* c = nullptr;
    int rc = f2_create(1, nx, ny, prm->pi, prm->lx, prm->ly, prm->dt, out, &c);
    if (rc) return rc;
    c->com.mu = prm->mu;
    c->com.eta = prm->eta;
    c->com.cs = prm->cs;
    c->com.kf = prm->kf;
    c->com.amp = prm->a;
    return TARA_OK;
}

// Reads the context's non-finite flag (after its stream has drained).
static int f2_check_finite(tara_f2* c)
{
    int bad = 0;
    TARA_HIP(hipMemcpyAsync(&bad, c->plan.bad.p, sizeof(int), hipMemcpyDeviceToHost, c->s));
    TARA_HIP(hipStreamSynchronize(c->s));
    TARA_CHECK(bad == 0, TARA_ERR_NONFINITE, "the state became non-finite (NaN or Inf) during advance");
    return TARA_OK;
}

extern "C" int tara_f2_upload(tara_f2* c, const double* const state[])
{
    TARA_CHECK(c && state, TARA_ERR_ARG, "null argument");
    for (int f = 0; f < kFields[c->family]; ++f) {
        TARA_CHECK(state[f], TARA_ERR_ARG, "null array %d", f);
        int rc = h2d(c->plan, c->st[f].as<double2>(), state[f], c->s);
        if (rc) return rc;
        TARA_HIP(hipMemsetAsync(c->hist[f].p, 0, c->hist[f].bytes, c->s));
    }
    TARA_HIP(hipMemsetAsync(c->plan.bad.p, 0, sizeof(int), c->s));
    TARA_HIP(hipStreamSynchronize(c->s));
    return TARA_OK;
}

extern "C" int tara_f2_download(tara_f2* c, double* const state[])
{
    TARA_CHECK(c && state, TARA_ERR_ARG, "null argument");
    for (int f = 0; f < kFields[c->family]; ++f) {
        TARA_CHECK(state[f], TARA_ERR_ARG, "null array %d", f);
        int rc = d2h(c->plan, state[f], c->st[f].as<double2>(), c->s);
        if (rc) return rc;
    }
    return f2_check_finite(c);
}

extern "C" int tara_f2_advance(tara_f2* c, int nsteps)
{
    TARA_CHECK(c && nsteps >= 0, TARA_ERR_ARG, "bad argument");
    double2* st[6];
    double2* hist[6];
    for (int f = 0; f < 6; ++f) {
        st[f] = c->st[f].as<double2>();
        hist[f] = c->hist[f].as<double2>();
    }
    for (int s = 0; s < nsteps; ++s) {
        int rc = c->family == 0 ? coh_rhs_stage(c->plan, c->coh, st, F2_AB2, c->dt, hist, c->s, c->tm.get())
                                : com_rhs_stage(c->plan, c->com, st, F2_AB2, c->dt, hist, c->s, c->tm.get());
        if (rc) return rc;
    }
    return TARA_OK;
}

extern "C" int tara_f2_sync(tara_f2* c)
{
    TARA_CHECK(c, TARA_ERR_ARG, "null context");
    return f2_check_finite(c);
}

extern "C" int tara_f2_destroy(tara_f2* c)
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

extern "C" int tara_f2_set_timing(tara_f2* c, int on)
{
    TARA_CHECK(c, TARA_ERR_ARG, "null context");
    if (on && !c->tm) c->tm = std::make_unique<EventTimer>();
    if (!on) c->tm.reset();
    return TARA_OK;
}

extern "C" int tara_f2_kernel_stats(tara_f2* c, int nclass, long* counts, double* ms)
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
