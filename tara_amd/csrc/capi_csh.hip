// C ABI of the 2D incompressible paths: drop-in derive2csh / ab2csh and
// derive2css / ab2css / rk42css on host arrays, and the device-resident
// stepping context shared by both.
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "common.h"
#include "csh2d.h"
#include "fft_api.h"
#include "generic2d.h"
#include "layout.h"

namespace tara {

void EventTimer::begin(int cls, hipStream_t st)
{
    Pair p;
    if (!pool.empty()) {
        p = pool.back();
        pool.pop_back();
    } else {
        (void)hipEventCreate(&p.a);
        (void)hipEventCreate(&p.b);
    }
    p.cls = cls;
    (void)hipEventRecord(p.a, st);
    cur[cls] = p;
}

void EventTimer::end(int cls, hipStream_t st)
{
    Pair p = cur[cls];
    (void)hipEventRecord(p.b, st);
    live.push_back(p);
}

int EventTimer::collect()
{
    for (auto& p : live) {
        TARA_HIP(hipEventSynchronize(p.b));
        float t = 0.f;
        TARA_HIP(hipEventElapsedTime(&t, p.a, p.b));
        counts[p.cls] += 1;
        ms[p.cls] += t;
        pool.push_back(p);
    }
    live.clear();
    return TARA_OK;
}

EventTimer::~EventTimer()
{
    for (auto& p : live) pool.push_back(p);
    for (auto& p : pool) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
}

namespace {

// Drop-in calls keep one plan and staging buffers per (device, nx, ny), as the
// reference re-plans per call (src/derive2csh.f95:73-106) but need not.
struct DropIn {
    CshPlan plan;
    DevBuf w, f, stage;  // w, f ky-major; stage in the caller's kx-fastest layout
    DevBuf aux, next;    // RK4 accumulator and stage field (allocated on first RK4 call)
};
std::mutex g_mu;
std::map<std::tuple<int, int, int>, std::unique_ptr<DropIn>> g_dropin;

int get_dropin(int nx, int ny, DropIn** out)
{
    auto key = std::make_tuple(current_device(), nx, ny);
    auto it = g_dropin.find(key);
    if (it != g_dropin.end()) {
        *out = it->second.get();
        return TARA_OK;
    }
    auto d = std::make_unique<DropIn>();
    int rc = csh_plan_init(d->plan, nx, ny);
    if (rc) return rc;
    const size_t bytes = (size_t)d->plan.nh * ny * sizeof(double2);
    if ((rc = d->w.alloc(bytes)) || (rc = d->f.alloc(bytes)) || (rc = d->stage.alloc(bytes))) return rc;
    *out = d.get();
    g_dropin[key] = std::move(d);
    return TARA_OK;
}

struct Ab2Bufs {
    DevBuf b[4];
    size_t bytes = 0;
};
Ab2Bufs g_ab2;

// Grids the fused passes do not take: upload to pitch spec_ld(nx), the
// any-size RHS (or RK4 step), download.
int generic_dropin(int nx, int ny, int nh, const CshPhys& ph, const double* in, double* out, bool rk4, double dt)
{
    TARA_CHECK(nx >= 2 && ny >= 1, TARA_ERR_UNSUPPORTED, "grid %dx%d unsupported", nx, ny);
    const int ld = spec_ld(nx);
    const size_t bytes = (size_t)ld * ny * sizeof(double2);
    DevBuf a, b;
    int rc;
    if ((rc = a.alloc(bytes)) || (rc = b.alloc(bytes))) return rc;
    TARA_HIP(hipMemset(a.p, 0, bytes));
    TARA_HIP(hipMemcpy2D(a.p, ld * sizeof(double2), in, nh * sizeof(double2), nh * sizeof(double2), ny,
                         hipMemcpyHostToDevice));
    double2* res = a.as<double2>();
    if (rk4)
        rc = gen_csh_rk4(nx, ny, ph, dt, res, nullptr);
    else {
        rc = gen_csh_rhs(nx, ny, ph, res, b.as<double2>(), nullptr);
        res = b.as<double2>();
    }
    if (rc) return rc;
    TARA_HIP(hipMemcpy2D(out, nh * sizeof(double2), res, ld * sizeof(double2), nh * sizeof(double2), ny,
                         hipMemcpyDeviceToHost));
    return TARA_OK;
}

int derive_dropin(const char* name, int nx, int ny, int nh, const CshPhys& ph, const double* omegak, double* dt_new)
{
    TARA_CHECK(omegak && dt_new, TARA_ERR_ARG, "%s: null array", name);
    TARA_CHECK(nh == nx / 2 + 1, TARA_ERR_ARG, "%s: Nh=%d but Nx/2+1=%d", name, nh, nx / 2 + 1);
    if (!csh_fast_grid(nx, ny)) return generic_dropin(nx, ny, nh, ph, omegak, dt_new, false, 0.0);
    std::lock_guard<std::mutex> lk(g_mu);
    DropIn* d = nullptr;
    int rc = get_dropin(nx, ny, &d);
    if (rc) return rc;
    const CshPlan& p = d->plan;
    const size_t bytes = (size_t)nh * ny * sizeof(double2);
    TARA_HIP(hipMemcpy(d->stage.p, omegak, bytes, hipMemcpyHostToDevice));
    if ((rc = to_kymajor(d->stage.as<double2>(), d->w.as<double2>(), nh, ny, 1, nullptr))) return rc;
    rc = csh_rhs_stage(p, d->w.as<double2>(), CSH_RHS, ph, 0.0, nullptr, nullptr, d->f.as<double2>(), nullptr,
                       nullptr);
    if (rc) return rc;
    if ((rc = from_kymajor(d->f.as<double2>(), d->stage.as<double2>(), nh, ny, 1, nullptr))) return rc;
    TARA_HIP(hipMemcpy(dt_new, d->stage.p, bytes, hipMemcpyDeviceToHost));
    return TARA_OK;
}

// One classic RK4 step of the RHS selected by ph on host arrays.
int rk4_dropin(const char* name, int nx, int ny, int nh, const CshPhys& ph, double dt, const double* omegak,
               double* omegak_new)
{
    TARA_CHECK(omegak && omegak_new, TARA_ERR_ARG, "%s: null array", name);
    TARA_CHECK(nh == nx / 2 + 1, TARA_ERR_ARG, "%s: Nh=%d but Nx/2+1=%d", name, nh, nx / 2 + 1);
    if (!csh_fast_grid(nx, ny)) return generic_dropin(nx, ny, nh, ph, omegak, omegak_new, true, dt);
    std::lock_guard<std::mutex> lk(g_mu);
    DropIn* d = nullptr;
    int rc = get_dropin(nx, ny, &d);
    if (rc) return rc;
    const CshPlan& p = d->plan;
    const size_t bytes = (size_t)nh * ny * sizeof(double2);
    if (!d->aux.p && ((rc = d->aux.alloc(bytes)) || (rc = d->next.alloc(bytes)))) return rc;
    double2* w = d->w.as<double2>();
    double2* aux = d->aux.as<double2>();
    double2* nxt = d->next.as<double2>();
    TARA_HIP(hipMemcpy(d->stage.p, omegak, bytes, hipMemcpyHostToDevice));
    if ((rc = to_kymajor(d->stage.as<double2>(), w, nh, ny, 1, nullptr))) return rc;
    if ((rc = csh_rhs_stage(p, w, CSH_RK4_1, ph, dt, w, aux, nxt, nullptr, nullptr))) return rc;
    if ((rc = csh_rhs_stage(p, nxt, CSH_RK4_2, ph, dt, w, aux, nxt, nullptr, nullptr))) return rc;
    if ((rc = csh_rhs_stage(p, nxt, CSH_RK4_3, ph, dt, w, aux, nxt, nullptr, nullptr))) return rc;
    if ((rc = csh_rhs_stage(p, nxt, CSH_RK4_4, ph, dt, w, aux, nxt, nullptr, nullptr))) return rc;
    if ((rc = from_kymajor(w, d->stage.as<double2>(), nh, ny, 1, nullptr))) return rc;
    TARA_HIP(hipMemcpy(omegak_new, d->stage.p, bytes, hipMemcpyDeviceToHost));
    return TARA_OK;
}

CshPhys csh_phys(double nu)
{
    CshPhys ph;
    ph.nu = nu;
    return ph;
}

CshPhys css_phys(double alpha, double nu, double lambda)
{
    CshPhys ph;
    ph.nu = nu;
    ph.alpha = alpha;
    ph.lambda = lambda;
    ph.screened = 1;
    return ph;
}

}  // namespace
}  // namespace tara

using namespace tara;

extern "C" int tara_derive2csh(int nx, int ny, int nh, double nu, const double* omegak, double* dt_new)
{
    return derive_dropin("derive2csh", nx, ny, nh, csh_phys(nu), omegak, dt_new);
}

extern "C" int tara_derive2css(int nx, int ny, int nh, double alpha, double nu, double lambda, const double* omegak,
                               double* dt_new)
{
    return derive_dropin("derive2css", nx, ny, nh, css_phys(alpha, nu, lambda), omegak, dt_new);
}

extern "C" int tara_rk42css(int nx, int ny, int nh, double alpha, double nu, double lambda, double dt,
                            const double* omegak, double* omegak_new)
{
    return rk4_dropin("rk42css", nx, ny, nh, css_phys(alpha, nu, lambda), dt, omegak, omegak_new);
}

extern "C" int tara_ab2csh(int nx, int ny, int nh, double dt, const double* omegak, double* omegak_new,
                           const double* f_old, const double* f_new)
{
    TARA_CHECK(omegak && omegak_new && f_old && f_new, TARA_ERR_ARG, "ab2csh: null array");
    TARA_CHECK(nh == nx / 2 + 1 && nx > 0 && ny > 0, TARA_ERR_ARG, "ab2csh: bad dims %d %d %d", nx, ny, nh);
    std::lock_guard<std::mutex> lk(g_mu);
    const size_t n = (size_t)nh * ny, bytes = n * sizeof(double2);
    if (g_ab2.bytes < bytes) {
        for (auto& b : g_ab2.b) {
            int rc = b.alloc(bytes);
            if (rc) return rc;
        }
        g_ab2.bytes = bytes;
    }
    TARA_HIP(hipMemcpy(g_ab2.b[0].p, omegak, bytes, hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(g_ab2.b[2].p, f_old, bytes, hipMemcpyHostToDevice));
    TARA_HIP(hipMemcpy(g_ab2.b[3].p, f_new, bytes, hipMemcpyHostToDevice));
    int rc = launch_ab2(g_ab2.b[0].as<double2>(), g_ab2.b[1].as<double2>(), g_ab2.b[2].as<double2>(),
                        g_ab2.b[3].as<double2>(), dt, n, nullptr);
    if (rc) return rc;
    TARA_HIP(hipMemcpy(omegak_new, g_ab2.b[1].p, bytes, hipMemcpyDeviceToHost));
    return TARA_OK;
}

extern "C" int tara_ab2css(int nx, int ny, int nh, double dt, const double* omegak, double* omegak_new,
                           const double* f_old, const double* f_new)
{
    return tara_ab2csh(nx, ny, nh, dt, omegak, omegak_new, f_old, f_new);
}

// ------------------------------------------------------- stepping context
struct tara_csh {
    CshPlan plan;
    DevBuf w, aux, next, stage;  // ky-major state; stage in the caller's layout
    hipStream_t st = nullptr;
    CshPhys ph;
    double dt = 0;
    int scheme = 0;
    std::unique_ptr<EventTimer> tm;
};

namespace {
int create_ctx(int nx, int ny, const CshPhys& ph, double dt, int scheme, tara_csh** out)
{
    TARA_CHECK(out, TARA_ERR_ARG, "null out");
    TARA_CHECK(scheme == TARA_SCHEME_AB2 || scheme == TARA_SCHEME_RK4, TARA_ERR_ARG, "unknown scheme %d", scheme);
    *out = nullptr;
    auto c = std::make_unique<tara_csh>();
    int rc = csh_plan_init(c->plan, nx, ny);
    if (rc) return rc;
    const size_t bytes = (size_t)c->plan.nh * ny * sizeof(double2);
    if ((rc = c->w.alloc(bytes)) || (rc = c->aux.alloc(bytes)) || (rc = c->stage.alloc(bytes))) return rc;
    if (scheme == TARA_SCHEME_RK4 && (rc = c->next.alloc(bytes))) return rc;
    TARA_HIP(hipMemset(c->w.p, 0, bytes));
    TARA_HIP(hipMemset(c->aux.p, 0, bytes));
    if (c->next.p) TARA_HIP(hipMemset(c->next.p, 0, bytes));
    TARA_HIP(hipStreamCreate(&c->st));
    c->ph = ph;
    c->dt = dt;
    c->scheme = scheme;
    *out = c.release();
    return TARA_OK;
}
}  // namespace

extern "C" int tara_csh_create(int nx, int ny, double nu, double dt, int scheme, tara_csh** out)
{
    return create_ctx(nx, ny, csh_phys(nu), dt, scheme, out);
}

extern "C" int tara_css_create(int nx, int ny, double alpha, double nu, double lambda, double dt, int scheme,
                               tara_csh** out)
{
    return create_ctx(nx, ny, css_phys(alpha, nu, lambda), dt, scheme, out);
}

namespace tara {
int csh_ctx_state(tara_csh* c, double2** w, const CshPlan** p, hipStream_t* st)
{
    TARA_CHECK(c, TARA_ERR_ARG, "null context");
    *w = c->w.as<double2>();
    *p = &c->plan;
    *st = c->st;
    return TARA_OK;
}

int csh_ctx_zero_history(tara_csh* c)
{
    TARA_CHECK(c, TARA_ERR_ARG, "null context");
    TARA_HIP(hipMemsetAsync(c->aux.p, 0, c->aux.bytes, c->st));
    int rc = csh_clear_nonfinite(c->plan, -1, c->st);  // a new state: the guard starts over
    if (rc) return rc;
    TARA_HIP(hipStreamSynchronize(c->st));
    return TARA_OK;
}
}  // namespace tara

extern "C" int tara_csh_upload(tara_csh* c, const double* omegak)
{
    TARA_CHECK(c && omegak, TARA_ERR_ARG, "null argument");
    const CshPlan& p = c->plan;
    TARA_HIP(hipMemcpyAsync(c->stage.p, omegak, c->stage.bytes, hipMemcpyHostToDevice, c->st));
    int rc = to_kymajor(c->stage.as<double2>(), c->w.as<double2>(), p.nh, p.ny, 1, c->st);
    if (rc) return rc;
    return csh_ctx_zero_history(c);
}

namespace {
// The guard of the resident steppers: the last stage of every step flags a
// non-finite new state value; reported here after the stream has drained.
int check_finite(tara_csh* c)
{
    const int n = csh_nonfinite(c->plan, nullptr, c->st);
    if (n < 0) return n;
    TARA_CHECK(n == 0, TARA_ERR_NONFINITE, "the state became non-finite (NaN or Inf) during advance");
    return TARA_OK;
}
}  // namespace

extern "C" int tara_csh_download(tara_csh* c, double* omegak)
{
    TARA_CHECK(c && omegak, TARA_ERR_ARG, "null argument");
    const CshPlan& p = c->plan;
    int rc = from_kymajor(c->w.as<double2>(), c->stage.as<double2>(), p.nh, p.ny, 1, c->st);
    if (rc) return rc;
    TARA_HIP(hipMemcpyAsync(omegak, c->stage.p, c->stage.bytes, hipMemcpyDeviceToHost, c->st));
    TARA_HIP(hipStreamSynchronize(c->st));
    return check_finite(c);  // the state is copied out either way
}

extern "C" int tara_csh_advance(tara_csh* c, int nsteps)
{
    TARA_CHECK(c && nsteps >= 0, TARA_ERR_ARG, "bad argument");
    double2* w = c->w.as<double2>();
    double2* aux = c->aux.as<double2>();
    double2* nx = c->next.as<double2>();
    EventTimer* tm = c->tm.get();
    if (!c->ph.screened) return csh_advance(c->plan, c->scheme, nsteps, c->ph, c->dt, w, aux, nx, c->st, tm);
    for (int s = 0; s < nsteps; ++s) {
        int rc;
        if (c->scheme == TARA_SCHEME_AB2) {
            rc = csh_rhs_stage(c->plan, w, CSH_AB2, c->ph, c->dt, w, aux, nullptr, c->st, tm);
            if (rc) return rc;
        } else {
            if ((rc = csh_rhs_stage(c->plan, w, CSH_RK4_1, c->ph, c->dt, w, aux, nx, c->st, tm))) return rc;
            if ((rc = csh_rhs_stage(c->plan, nx, CSH_RK4_2, c->ph, c->dt, w, aux, nx, c->st, tm))) return rc;
            if ((rc = csh_rhs_stage(c->plan, nx, CSH_RK4_3, c->ph, c->dt, w, aux, nx, c->st, tm))) return rc;
            if ((rc = csh_rhs_stage(c->plan, nx, CSH_RK4_4, c->ph, c->dt, w, aux, nx, c->st, tm))) return rc;
        }
    }
    return TARA_OK;
}

extern "C" int tara_csh_sync(tara_csh* c)
{
    TARA_CHECK(c, TARA_ERR_ARG, "null context");
    TARA_HIP(hipStreamSynchronize(c->st));
    return check_finite(c);
}

extern "C" int tara_csh_destroy(tara_csh* c)
{
    if (!c) return TARA_OK;
    if (c->st) {
        (void)hipStreamSynchronize(c->st);
        c->tm.reset();
        (void)hipStreamDestroy(c->st);
    }
    delete c;
    return TARA_OK;
}

extern "C" int tara_csh_num_xshapes(void) { return csh_num_xshapes(); }

extern "C" int tara_csh_tune(tara_csh* c, int reps, int* chosen, double* ms)
{
    TARA_CHECK(c && reps > 0, TARA_ERR_ARG, "bad argument");
    const int n = csh_num_xshapes();
    hipEvent_t e0, e1;
    TARA_HIP(hipEventCreate(&e0));
    TARA_HIP(hipEventCreate(&e1));
    const int keep = c->plan.xcfg;
    int best = -1;
    float best_ms = 0.f;
    int rc = TARA_OK;
    for (int k = 0; k < n && rc == TARA_OK; ++k) {
        if (ms) ms[k] = -1.0;
        if (!csh_xshape_ok(c->plan, k)) continue;
        c->plan.xcfg = k;
        if ((rc = csh_xpass(c->plan, c->st))) break;  // warm-up
        if (hipEventRecord(e0, c->st) != hipSuccess) { rc = TARA_ERR_HIP; break; }
        for (int r = 0; r < reps && rc == TARA_OK; ++r) rc = csh_xpass(c->plan, c->st);
        if (rc) break;
        float t = 0.f;
        if (hipEventRecord(e1, c->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
            rc = TARA_ERR_HIP;
            break;
        }
        t /= (float)reps;
        if (ms) ms[k] = t;
        if (best < 0 || t < best_ms) {
            best = k;
            best_ms = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        c->plan.xcfg = keep;
        return rc;
    }
    c->plan.xcfg = best >= 0 ? best : keep;
    if (chosen) *chosen = c->plan.xcfg;
    return TARA_OK;
}

extern "C" int tara_csh_set_xshape(tara_csh* c, int shape)
{
    TARA_CHECK(c, TARA_ERR_ARG, "null context");
    TARA_CHECK(csh_xshape_ok(c->plan, shape), TARA_ERR_UNSUPPORTED, "x pass shape %d does not run on %dx%d", shape,
               c->plan.nx, c->plan.ny);
    c->plan.xcfg = shape;
    return TARA_OK;
}

extern "C" int tara_csh_set_timing(tara_csh* c, int on)
{
    TARA_CHECK(c, TARA_ERR_ARG, "null context");
    if (on && !c->tm) c->tm = std::make_unique<EventTimer>();
    if (!on) c->tm.reset();
    return TARA_OK;
}

extern "C" int tara_csh_kernel_stats(tara_csh* c, int nclass, long* counts, double* ms)
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
