// C ABI of the batched ensemble: many independent 2D incompressible runs
// (derive2csh physics) of possibly different grid sizes and viscosities,
// advanced together.  Runs are grouped by grid size into buckets; each bucket's
// passes are single batched launches over all its runs (blockIdx.y = run), the
// buckets run on streams of their own, and one whole step (every bucket, every
// stage) is captured once in a hipGraph and replayed per step.
#include <algorithm>
#include <map>
#include <memory>
#include <vector>

#include "common.h"
#include "csh2d.h"
#include "ens.h"
#include "fft_api.h"
#include "layout.h"

using namespace tara;

namespace {

struct Bucket {
    int n = 0;
    std::vector<int> runs;  // run ids, in bucket order
    CshPlan plan;           // batch = runs.size()
    DevBuf w, aux, next, stage, nus;
    DevBuf spec, real;  // real <-> spectral workspace of one run (allocated on first use)
    CshPhys ph;
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
};

}  // namespace

struct tara_ens {
    std::vector<std::unique_ptr<Bucket>> b;
    std::vector<std::pair<int, int>> where;  // run -> (bucket, slot)
    hipStream_t st = nullptr;
    hipEvent_t fork = nullptr;
    double dt = 0;
    int scheme = 0;
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

void drop_graph(tara_ens* e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

// One step of every bucket, forked onto the bucket streams and joined back.
// Each forward y pass is fused with the following inverse y pass, the last
// stage's too: a step expects G0..G3 to hold the inverse y pass of the state
// (enqueue_head) and leaves them holding that of the new state, so that the
// captured step can be replayed back to back.
int enqueue_step(tara_ens* e)
{
    TARA_HIP(hipEventRecord(e->fork, e->st));
    for (auto& bp : e->b) {
        Bucket& b = *bp;
        TARA_HIP(hipStreamWaitEvent(b.st, e->fork, 0));
        int rc = csh_fused_step(b.plan, e->scheme, b.ph, e->dt, b.w.as<double2>(), b.aux.as<double2>(),
                                b.next.as<double2>(), false, b.st, nullptr);
        if (rc) return rc;
        TARA_HIP(hipEventRecord(b.done, b.st));
        TARA_HIP(hipStreamWaitEvent(e->st, b.done, 0));
    }
    return TARA_OK;
}

// The inverse y pass of every bucket's state, once per advance call (the state
// may have been set or uploaded since the last step).
int enqueue_head(tara_ens* e)
{
    TARA_HIP(hipEventRecord(e->fork, e->st));
    for (auto& bp : e->b) {
        Bucket& b = *bp;
        TARA_HIP(hipStreamWaitEvent(b.st, e->fork, 0));
        int rc = csh_yinv_pass(b.plan, b.w.as<double2>(), b.ph, b.st);
        if (rc) return rc;
        TARA_HIP(hipEventRecord(b.done, b.st));
        TARA_HIP(hipStreamWaitEvent(e->st, b.done, 0));
    }
    return TARA_OK;
}

int capture(tara_ens* e)
{
    drop_graph(e);
    TARA_HIP(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(e);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(e->st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    TARA_HIP(ce);
    e->graph = g;
    TARA_HIP(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    return TARA_OK;
}

int run_slot(tara_ens* e, int run, Bucket** b, int* slot)
{
    TARA_CHECK(e && run >= 0 && run < (int)e->where.size(), TARA_ERR_ARG, "run %d out of range", run);
    *b = e->b[e->where[run].first].get();
    *slot = e->where[run].second;
    return TARA_OK;
}

// w[i*ny + j] = s[j*ld + i]: padded kx-fastest spectrum -> ky-major state
__global__ void k_spec_to_state(const double2* s, int ld, int nh, int ny, double2* w)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)nh * ny;
         q += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(q % ny), i = (int)(q / ny);
        w[q] = s[(size_t)j * ld + i];
    }
}

// s[j*ld + i] = w[i*ny + j] for i < nh, 0 in the padding columns
__global__ void k_state_to_spec(const double2* w, int nh, int ny, int ld, double2* s)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)ld * ny;
         q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % ld), j = (int)(q / ld);
        s[q] = i < nh ? w[(size_t)i * ny + j] : make_double2(0.0, 0.0);
    }
}

__global__ void k_f32_to_f64(const float* a, double* b, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        b[q] = (double)a[q];
}

__global__ void k_f64_to_f32(const double* a, float* b, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        b[q] = (float)a[q];
}

unsigned ens_grid(size_t n)
{
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b ? b : 1));
}

int bucket_ws(Bucket& b)
{
    int rc;
    if (!b.spec.p && (rc = b.spec.alloc((size_t)spec_ld(b.n) * b.n * sizeof(double2)))) return rc;
    if (!b.real.p && (rc = b.real.alloc((size_t)b.n * b.n * sizeof(double)))) return rc;
    return TARA_OK;
}

}  // namespace

extern "C" int tara_ens_create(int nruns, const int* n, const double* nu, double dt, int scheme, tara_ens** out)
{
    TARA_CHECK(out && n && nu && nruns > 0, TARA_ERR_ARG, "bad argument");
    TARA_CHECK(scheme == TARA_SCHEME_AB2 || scheme == TARA_SCHEME_RK4, TARA_ERR_ARG, "unknown scheme %d", scheme);
    *out = nullptr;
    auto e = std::make_unique<tara_ens>();
    e->dt = dt;
    e->scheme = scheme;
    std::map<int, int> index;  // grid -> bucket
    e->where.resize(nruns);
    for (int r = 0; r < nruns; ++r) {
        auto it = index.find(n[r]);
        if (it == index.end()) {
            it = index.emplace(n[r], (int)e->b.size()).first;
            e->b.push_back(std::make_unique<Bucket>());
            e->b.back()->n = n[r];
        }
        Bucket& b = *e->b[it->second];
        e->where[r] = {it->second, (int)b.runs.size()};
        b.runs.push_back(r);
    }
    int rc;
    for (auto& bp : e->b) {
        Bucket& b = *bp;
        const int cnt = (int)b.runs.size();
        if ((rc = csh_plan_init(b.plan, b.n, b.n, cnt))) return rc;
        // per-run slots are bstride apart (padded); the host-side staging
        // buffer holds exactly one reference-layout spectrum (nh x ny)
        const size_t bytes = b.plan.bstride * sizeof(double2);
        const size_t spec_bytes = (size_t)b.plan.nh * b.plan.ny * sizeof(double2);
        if ((rc = b.w.alloc(cnt * bytes)) || (rc = b.aux.alloc(cnt * bytes)) || (rc = b.stage.alloc(spec_bytes)) ||
            (rc = b.nus.alloc(cnt * sizeof(double))))
            return rc;
        if (scheme == TARA_SCHEME_RK4 && (rc = b.next.alloc(cnt * bytes))) return rc;
        TARA_HIP(hipMemset(b.w.p, 0, b.w.bytes));
        TARA_HIP(hipMemset(b.aux.p, 0, b.aux.bytes));
        if (b.next.p) TARA_HIP(hipMemset(b.next.p, 0, b.next.bytes));
        std::vector<double> v(cnt);
        for (int k = 0; k < cnt; ++k) v[k] = nu[b.runs[k]];
        TARA_HIP(hipMemcpy(b.nus.p, v.data(), cnt * sizeof(double), hipMemcpyHostToDevice));
        b.ph.nu = v[0];
        b.ph.nus = b.nus.as<double>();
        TARA_HIP(hipStreamCreate(&b.st));
        TARA_HIP(hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
    }
    TARA_HIP(hipStreamCreate(&e->st));
    TARA_HIP(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
    *out = e.release();
    return TARA_OK;
}

extern "C" int tara_ens_upload(tara_ens* e, int run, const double* omegak)
{
    TARA_CHECK(omegak, TARA_ERR_ARG, "null array");
    Bucket* b = nullptr;
    int slot = 0, rc;
    if ((rc = run_slot(e, run, &b, &slot))) return rc;
    const CshPlan& p = b->plan;
    TARA_HIP(hipStreamSynchronize(e->st));
    TARA_HIP(hipMemcpy(b->stage.p, omegak, b->stage.bytes, hipMemcpyHostToDevice));
    if ((rc = to_kymajor(b->stage.as<double2>(), b->w.as<double2>() + slot * p.bstride, p.nh, p.ny, 1, nullptr)))
        return rc;
    TARA_HIP(hipMemset(b->aux.as<double2>() + slot * p.bstride, 0, p.bstride * sizeof(double2)));
    if ((rc = csh_clear_nonfinite(p, slot, nullptr))) return rc;
    TARA_HIP(hipDeviceSynchronize());
    return TARA_OK;
}

extern "C" int tara_ens_download(tara_ens* e, int run, double* omegak)
{
    TARA_CHECK(omegak, TARA_ERR_ARG, "null array");
    Bucket* b = nullptr;
    int slot = 0, rc;
    if ((rc = run_slot(e, run, &b, &slot))) return rc;
    const CshPlan& p = b->plan;
    TARA_HIP(hipStreamSynchronize(e->st));
    if ((rc = from_kymajor(b->w.as<double2>() + slot * p.bstride, b->stage.as<double2>(), p.nh, p.ny, 1, nullptr)))
        return rc;
    TARA_HIP(hipMemcpy(omegak, b->stage.p, b->stage.bytes, hipMemcpyDeviceToHost));
    return TARA_OK;
}

extern "C" int tara_ens_ic_random(tara_ens* e, int run, unsigned seed, double k0, double energy)
{
    Bucket* b = nullptr;
    int slot = 0, rc;
    if ((rc = run_slot(e, run, &b, &slot))) return rc;
    const CshPlan& p = b->plan;
    if ((rc = csh_random_phase(b->w.as<double2>() + slot * p.bstride, b->n, b->n, seed, k0, energy, e->st)))
        return rc;
    TARA_HIP(hipMemsetAsync(b->aux.as<double2>() + slot * p.bstride, 0, p.bstride * sizeof(double2), e->st));
    if ((rc = csh_clear_nonfinite(p, slot, e->st))) return rc;
    TARA_HIP(hipStreamSynchronize(e->st));
    return TARA_OK;
}

extern "C" int tara_ens_set_omega(tara_ens* e, int run, const float* omega)
{
    TARA_CHECK(omega, TARA_ERR_ARG, "null array");
    Bucket* b = nullptr;
    int slot = 0, rc;
    if ((rc = run_slot(e, run, &b, &slot)) || (rc = bucket_ws(*b))) return rc;
    const CshPlan& p = b->plan;
    const int n = b->n, ld = spec_ld(n);
    const size_t np = (size_t)n * n;
    hipLaunchKernelGGL(k_f32_to_f64, dim3(ens_grid(np)), dim3(256), 0, e->st, omega, b->real.as<double>(), np);
    if ((rc = row_r2c(b->real.as<double>(), b->spec.as<double2>(), n, ld, n, e->st))) return rc;
    if ((rc = col_fft(b->spec.as<double2>(), n, ld, ld, (long)ld * n, 1, -1, e->st))) return rc;
    hipLaunchKernelGGL(k_spec_to_state, dim3(ens_grid((size_t)p.nh * n)), dim3(256), 0, e->st,
                       b->spec.as<double2>(), ld, p.nh, n, b->w.as<double2>() + slot * p.bstride);
    TARA_HIP(hipGetLastError());
    TARA_HIP(hipMemsetAsync(b->aux.as<double2>() + slot * p.bstride, 0, p.bstride * sizeof(double2), e->st));
    return csh_clear_nonfinite(p, slot, e->st);
}

extern "C" int tara_ens_get_omega(tara_ens* e, int run, float* omega)
{
    TARA_CHECK(omega, TARA_ERR_ARG, "null array");
    Bucket* b = nullptr;
    int slot = 0, rc;
    if ((rc = run_slot(e, run, &b, &slot)) || (rc = bucket_ws(*b))) return rc;
    const CshPlan& p = b->plan;
    const int n = b->n, ld = spec_ld(n);
    const size_t np = (size_t)n * n;
    hipLaunchKernelGGL(k_state_to_spec, dim3(ens_grid((size_t)ld * n)), dim3(256), 0, e->st,
                       b->w.as<double2>() + slot * p.bstride, p.nh, n, ld, b->spec.as<double2>());
    TARA_HIP(hipGetLastError());
    if ((rc = col_fft(b->spec.as<double2>(), n, ld, ld, (long)ld * n, 1, +1, e->st))) return rc;
    if ((rc = row_c2r(b->spec.as<double2>(), b->real.as<double>(), n, ld, n, 1.0 / (double)np, e->st))) return rc;
    hipLaunchKernelGGL(k_f64_to_f32, dim3(ens_grid(np)), dim3(256), 0, e->st, b->real.as<double>(), omega, np);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

extern "C" int tara_ens_advance(tara_ens* e, int nsteps)
{
    TARA_CHECK(e && nsteps >= 0, TARA_ERR_ARG, "bad argument");
    int rc;
    // the step graph is captured on the first call, also with nsteps = 0 (a
    // caller can then capture before any work is in flight)
    if (e->use_graph && !e->exec && (rc = capture(e))) return rc;
    if (nsteps == 0) return TARA_OK;
    if ((rc = enqueue_head(e))) return rc;
    for (int s = 0; s < nsteps; ++s) {
        if (e->use_graph) {
            TARA_HIP(hipGraphLaunch(e->exec, e->st));
        } else if ((rc = enqueue_step(e))) {
            return rc;
        }
    }
    return TARA_OK;
}

extern "C" int tara_ens_graph_nodes(tara_ens* e, long* nodes, long* kernels)
{
    TARA_CHECK(e && nodes, TARA_ERR_ARG, "null argument");
    *nodes = 0;
    if (kernels) *kernels = 0;
    if (!e->graph) return TARA_OK;
    size_t n = 0;
    TARA_HIP(hipGraphGetNodes(e->graph, nullptr, &n));
    *nodes = (long)n;
    if (kernels && n) {
        std::vector<hipGraphNode_t> v(n);
        TARA_HIP(hipGraphGetNodes(e->graph, v.data(), &n));
        long k = 0;
        for (size_t i = 0; i < n; ++i) {
            hipGraphNodeType t;
            TARA_HIP(hipGraphNodeGetType(v[i], &t));
            k += t == hipGraphNodeTypeKernel ? 1 : 0;
        }
        *kernels = k;
    }
    return TARA_OK;
}

extern "C" int tara_ens_nonfinite(tara_ens* e, int* flags)
{
    TARA_CHECK(e, TARA_ERR_ARG, "null context");
    int total = 0;
    std::vector<int> f;
    for (auto& bp : e->b) {
        Bucket& b = *bp;
        f.assign(b.runs.size(), 0);
        const int n = csh_nonfinite(b.plan, f.data(), e->st);
        if (n < 0) return n;
        total += n;
        if (flags)
            for (size_t k = 0; k < b.runs.size(); ++k) flags[b.runs[k]] = f[k];
    }
    return total;
}

extern "C" int tara_ens_sync(tara_ens* e)
{
    TARA_CHECK(e, TARA_ERR_ARG, "null context");
    TARA_HIP(hipStreamSynchronize(e->st));
    const int n = tara_ens_nonfinite(e, nullptr);
    if (n < 0) return n;
    TARA_CHECK(n == 0, TARA_ERR_NONFINITE, "%d run(s) of the ensemble became non-finite (NaN or Inf)", n);
    return TARA_OK;
}

extern "C" int tara_ens_use_graph(tara_ens* e, int on)
{
    TARA_CHECK(e, TARA_ERR_ARG, "null context");
    TARA_HIP(hipStreamSynchronize(e->st));
    e->use_graph = on != 0;
    drop_graph(e);
    return TARA_OK;
}

extern "C" int tara_ens_tune(tara_ens* e, int reps)
{
    TARA_CHECK(e && reps > 0, TARA_ERR_ARG, "bad argument");
    TARA_HIP(hipStreamSynchronize(e->st));
    hipEvent_t e0, e1;
    TARA_HIP(hipEventCreate(&e0));
    TARA_HIP(hipEventCreate(&e1));
    int rc = TARA_OK;
    for (auto& bp : e->b) {
        Bucket& b = *bp;
        int best = b.plan.xcfg;
        float best_ms = -1.f;
        for (int k = 0; k < csh_num_xshapes() && rc == TARA_OK; ++k) {
            if (!csh_xshape_ok(b.plan, k)) continue;
            b.plan.xcfg = k;
            if ((rc = csh_xpass(b.plan, b.st))) break;
            (void)hipEventRecord(e0, b.st);
            for (int r = 0; r < reps && rc == TARA_OK; ++r) rc = csh_xpass(b.plan, b.st);
            (void)hipEventRecord(e1, b.st);
            float t = 0.f;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
                rc = TARA_ERR_HIP;
                break;
            }
            if (best_ms < 0 || t < best_ms) {
                best_ms = t;
                best = k;
            }
        }
        b.plan.xcfg = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    drop_graph(e);  // the captured launches carry the old shapes
    return rc;
}

extern "C" int tara_ens_destroy(tara_ens* e)
{
    if (!e) return TARA_OK;
    if (e->st) (void)hipStreamSynchronize(e->st);
    drop_graph(e);
    for (auto& bp : e->b) {
        if (bp->st) {
            (void)hipStreamSynchronize(bp->st);
            (void)hipStreamDestroy(bp->st);
        }
        if (bp->done) (void)hipEventDestroy(bp->done);
    }
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->st) (void)hipStreamDestroy(e->st);
    delete e;
    return TARA_OK;
}

namespace tara {
hipStream_t ens_stream(tara_ens* e) { return e ? e->st : nullptr; }
int ens_grid_of(tara_ens* e, int run)
{
    if (!e || run < 0 || run >= (int)e->where.size()) return 0;
    return e->b[e->where[run].first]->n;
}
}  // namespace tara
