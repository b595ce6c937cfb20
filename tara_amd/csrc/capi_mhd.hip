// C ABI of the 3D MHD path: drop-in derive3com / ab3com on host arrays, and
// the device-resident stepping context.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "fft_api.h"
#include "generic3d.h"
#include "mhd3d.h"
#include "philox.h"

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
    if (!mhd_fast_grid(nx, ny, nz)) return gen_derive3com_host(nx, ny, nz, prm, state, rhs);
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
    // diagnostics work space (allocated on the first tara_mhd_energy call):
    // spectral A (3), div B, u (3), omega (3) and one copy buffer; 15 real
    // fields; per-block partial sums
    DevBuf dspec[11], dreal[15], dpart;
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
    TARA_HIP(hipMemsetAsync(c->plan.bad.p, 0, sizeof(int), c->s));
    TARA_HIP(hipStreamSynchronize(c->s));
    return TARA_OK;
}

namespace {
// The guard of the resident stepper: the AB2 update flags a non-finite new
// state value; reported after the stream has drained.
int check_finite(tara_mhd* c)
{
    int bad = 0;
    TARA_HIP(hipMemcpyAsync(&bad, c->plan.bad.p, sizeof(int), hipMemcpyDeviceToHost, c->s));
    TARA_HIP(hipStreamSynchronize(c->s));
    TARA_CHECK(bad == 0, TARA_ERR_NONFINITE, "the 3D state became non-finite (NaN or Inf) during advance");
    return TARA_OK;
}
}  // namespace

extern "C" int tara_mhd_download(tara_mhd* c, double* const state[8])
{
    TARA_CHECK(c && state, TARA_ERR_ARG, "null argument");
    int rc;
    for (int f = 0; f < 8; ++f) {
        TARA_CHECK(state[f], TARA_ERR_ARG, "null array %d", f);
        if ((rc = d2h(c->plan, state[f], c->st[f].as<double2>(), c->s))) return rc;
    }
    TARA_HIP(hipStreamSynchronize(c->s));
    return check_finite(c);  // the state is copied out either way
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
    return check_finite(c);
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
    TARA_HIP(hipMemsetAsync(c->plan.bad.p, 0, sizeof(int), c->s));
    TARA_HIP(hipStreamSynchronize(c->s));
    return TARA_OK;
}

namespace {

constexpr int kIcBlocks = 1024;  // grid of the generation kernel (fixed: its partial sums are reduced in order)

// One random-phase solenoidal vector field (v0, v1, v2: padded (ld, Ny, Nz)
// half spectra) with |v(k)| per component k exp(-(k/k0)^2) before the
// projection v <- v - k (k . v)/k^2, i.e. shell spectrum ~ k^4 exp(-2 (k/k0)^2).
// Mode (i, j, l) draws its three phases 2 pi u from Philox-4x32-10 with key
// {seed, tag} and counters {m, 0, 0, 0} (u_0: words 0-1, u_1: words 2-3) and
// {m, 1, 0, 0} (u_2: words 0-1), m = i + Nh (j + Ny l); on the kx = 0 plane a
// mode whose mirror (-ky, -kz) has the smaller index j + Ny l takes the
// conjugate of the mirror's draw.  k = 0, the Nyquist planes and the padding
// columns are zero.  Each block writes its share of sum w |v|^2 (w = 1 on
// kx = 0, 2 elsewhere) to partial[blockIdx.x].
__global__ __launch_bounds__(256) void k_ic3d_field(double2* v0, double2* v1, double2* v2, int nh, int ny, int nz,
                                                   int ld, unsigned seed, unsigned tag, double k0, double* partial)
{
    __shared__ double red[256];
    double s = 0.0;
    const size_t n = (size_t)ld * ny * nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % ld), j = (int)((q / ld) % ny), l = (int)(q / ((size_t)ld * ny));
        double2 o[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
        if (!(i >= nh - 1 || j == ny / 2 || l == nz / 2 || (i == 0 && j == 0 && l == 0))) {
            int jc = j, lc = l;
            bool conj = false;
            if (i == 0) {
                const int jm = (ny - j) % ny, lm = (nz - l) % nz;
                if (jm + ny * lm < j + ny * l) {
                    jc = jm;
                    lc = lm;
                    conj = true;
                }
            }
            const double kv[3] = {(double)i, (double)(j < ny / 2 ? j : j - ny), (double)(l < nz / 2 ? l : l - nz)};
            const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
            const double k = sqrt(k2);
            const double amp = k * exp(-(k / k0) * (k / k0));
            const unsigned m = (unsigned)((size_t)i + (size_t)nh * ((size_t)jc + (size_t)ny * lc));
            const U4 r = philox4x32_10(U4{m, 0u, 0u, 0u}, seed, tag);
            const U4 t = philox4x32_10(U4{m, 1u, 0u, 0u}, seed, tag);
            const double u[3] = {u53(r.x, r.y), u53(r.z, r.w), u53(t.x, t.y)};
            double re[3], im[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double ph = 2.0 * 3.14159265358979323846 * u[c];
                re[c] = amp * cos(ph);
                im[c] = conj ? -(amp * sin(ph)) : amp * sin(ph);
            }
            const double dr = (kv[0] * re[0] + kv[1] * re[1] + kv[2] * re[2]) / k2;
            const double di = (kv[0] * im[0] + kv[1] * im[1] + kv[2] * im[2]) / k2;
            const double w = i == 0 ? 1.0 : 2.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                o[c] = make_double2(re[c] - kv[c] * dr, im[c] - kv[c] * di);
                s += w * (o[c].x * o[c].x + o[c].y * o[c].y);
            }
        }
        v0[q] = o[0];
        v1[q] = o[1];
        v2[q] = o[2];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// scale[0] = sqrt(energy / e) with e = 1/2 sum(partial) / nn^2, the field's
// 1/2 <|v|^2> (Parseval over the full spectrum of an unnormalised transform).
__global__ __launch_bounds__(256) void k_ic3d_scale_factor(const double* partial, int np, double energy, double nn,
                                                          double* scale)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int b = threadIdx.x; b < np; b += blockDim.x) s += partial[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double e = 0.5 * red[0] / (nn * nn);
        scale[0] = e > 0 ? sqrt(energy / e) : 0.0;
    }
}

__global__ void k_ic3d_rescale(double2* v0, double2* v1, double2* v2, size_t n, const double* scale)
{
    const double f = scale[0];
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const double2 a = v0[q], b = v1[q], c = v2[q];
        v0[q] = make_double2(a.x * f, a.y * f);
        v1[q] = make_double2(b.x * f, b.y * f);
        v2[q] = make_double2(c.x * f, c.y * f);
    }
}

// acc = (first ? 0 : acc) + x^2
__global__ void k_ic3d_accum(double* acc, const double* x, size_t n, int first)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const double v = x[q];
        acc[q] = (first ? 0.0 : acc[q]) + v * v;
    }
}

// E = P + rho u^2/2 + B^2/2 with P = 1, rho = 1: acc -> 1 + acc/2 in place
__global__ void k_ic3d_energy(double* acc, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        acc[q] = 1.0 + 0.5 * acc[q];
}

// rho = 1: the unnormalised transform is Nx Ny Nz at k = 0 and zero elsewhere.
__global__ void k_ic3d_density(double2* r, size_t n, double nn)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        r[q] = make_double2(q == 0 ? nn : 0.0, 0.0);
}

unsigned grid_of(size_t n)
{
    const size_t b = (n + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b ? b : 1));
}

}  // namespace

extern "C" int tara_mhd_ic_random(tara_mhd* c, unsigned seed, double k0, double e_kin, double e_mag)
{
    TARA_CHECK(c, TARA_ERR_ARG, "null context");
    TARA_CHECK(k0 > 0 && e_kin >= 0 && e_mag >= 0, TARA_ERR_ARG, "ic: k0 must be positive, energies non-negative");
    const MhdPlan& p = c->plan;
    const size_t n = p.fsz, nr = (size_t)p.nx * p.ny * p.nz;
    const double nn = (double)nr;
    hipStream_t s = c->s;
    // the history buffers are zeroed at the end and serve as scratch meanwhile:
    // hist[0] partial sums + scale, hist[1] spectral copy, hist[2] real field,
    // hist[3] the accumulated u^2 + B^2
    double* partial = c->hist[0].as<double>();
    double2* spec = c->hist[1].as<double2>();
    double* real = c->hist[2].as<double>();
    double* acc = c->hist[3].as<double>();
    const unsigned tags[2] = {1u, 2u};
    const double energies[2] = {e_kin, e_mag};
    const int base[2] = {1, 5};
    for (int v = 0; v < 2; ++v) {
        double2* f0 = c->st[base[v]].as<double2>();
        double2* f1 = c->st[base[v] + 1].as<double2>();
        double2* f2 = c->st[base[v] + 2].as<double2>();
        hipLaunchKernelGGL(k_ic3d_field, dim3(kIcBlocks), dim3(256), 0, s, f0, f1, f2, p.nh, p.ny, p.nz, p.ld, seed,
                           tags[v], k0, partial);
        hipLaunchKernelGGL(k_ic3d_scale_factor, dim3(1), dim3(256), 0, s, partial, kIcBlocks, energies[v], nn,
                           partial + kIcBlocks);
        hipLaunchKernelGGL(k_ic3d_rescale, dim3(grid_of(n)), dim3(256), 0, s, f0, f1, f2, n, partial + kIcBlocks);
        TARA_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_ic3d_density, dim3(grid_of(n)), dim3(256), 0, s, c->st[0].as<double2>(), n, nn);
    int rc, first = 1;
    for (int f : {1, 2, 3, 5, 6, 7}) {
        TARA_HIP(hipMemcpyAsync(spec, c->st[f].p, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
        if ((rc = col_fft(spec, p.nz, p.ld, (long)p.ld * p.ny, p.ld, p.ny, +1, s))) return rc;
        if ((rc = col_fft(spec, p.ny, p.ld, p.ld, (long)p.ld * p.ny, p.nz, +1, s))) return rc;
        if ((rc = row_c2r(spec, real, p.nx, p.ld, p.ny * p.nz, 1.0 / nn, s))) return rc;
        hipLaunchKernelGGL(k_ic3d_accum, dim3(grid_of(nr)), dim3(256), 0, s, acc, real, nr, first);
        first = 0;
    }
    hipLaunchKernelGGL(k_ic3d_energy, dim3(grid_of(nr)), dim3(256), 0, s, acc, nr);
    TARA_HIP(hipGetLastError());
    double2* e = c->st[4].as<double2>();
    TARA_HIP(hipMemsetAsync(e, 0, c->st[4].bytes, s));
    if ((rc = row_r2c(acc, e, p.nx, p.ld, p.ny * p.nz, s))) return rc;
    if ((rc = col_fft(e, p.ny, p.ld, p.ld, (long)p.ld * p.ny, p.nz, -1, s))) return rc;
    if ((rc = col_fft(e, p.nz, p.ld, (long)p.ld * p.ny, p.ld, p.ny, -1, s))) return rc;
    for (int f = 0; f < 7; ++f) TARA_HIP(hipMemsetAsync(c->hist[f].p, 0, c->hist[f].bytes, s));
    TARA_HIP(hipMemsetAsync(c->plan.bad.p, 0, sizeof(int), s));
    TARA_HIP(hipStreamSynchronize(s));
    return TARA_OK;
}

// ------------------------------------------------------------ diagnostics
// The per-step diagnostics of the reference's 3D driver, src/main3com.f95:
// 619-1071, on the resident state: div B, j = curl B and A = j/k^2 in spectral
// space (:621-686), the real fields (:688-835), u = rho u/rho and its
// transform (:837-880), omega = curl u as that loop writes it, including its
// omega_y = i kx uz - i kz ux (:882-975), and the volume averages of
// :1024-1071, written per step to Energy.dat.
namespace {

constexpr int kDiagBlocks = 1024;  // grid of the reduction (its partial sums are added in a fixed order)
constexpr int kDiagSums = 10;

// i k s (the reference's (0,1)*k*s: (0, k) times s)
__device__ __forceinline__ double2 dg_ik(double k, double2 s) { return make_double2(-(k * s.y), k * s.x); }
__device__ __forceinline__ double2 dg_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// div B, A = (i k x B)/k^2 (A = j at k = 0) from the B spectra (src/main3com.f95:623-686)
__global__ void k_diag_spec_b(const double2* bx, const double2* by, const double2* bz, const double* kx,
                              const double* ky, const double* kz, int ld, int ny, int nz, double2* ax, double2* ay,
                              double2* az, double2* divb)
{
    const size_t n = (size_t)ld * ny * nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % ld), j = (int)((q / ld) % ny), k = (int)(q / ((size_t)ld * ny));
        const double kxv = kx[i], kyv = ky[j], kzv = kz[k];
        const double2 x = bx[q], y = by[q], z = bz[q];
        const double2 d1 = dg_ik(kxv, x), d2 = dg_ik(kyv, y), d3 = dg_ik(kzv, z);
        divb[q] = make_double2((d1.x + d2.x) + d3.x, (d1.y + d2.y) + d3.y);
        const double2 jx = dg_sub(dg_ik(kyv, z), dg_ik(kzv, y));
        const double2 jy = dg_sub(dg_ik(kzv, x), dg_ik(kxv, z));
        const double2 jz = dg_sub(dg_ik(kxv, y), dg_ik(kyv, x));
        if (q == 0) {
            ax[q] = jx;
            ay[q] = jy;
            az[q] = jz;
        } else {
            const double k2 = kxv * kxv + kyv * kyv + kzv * kzv;
            // padding columns (kx = 0, k2 = 0 only at q = 0 otherwise) stay zero
            const double2 zero = make_double2(0.0, 0.0);
            ax[q] = k2 != 0.0 ? make_double2(jx.x / k2, jx.y / k2) : zero;
            ay[q] = k2 != 0.0 ? make_double2(jy.x / k2, jy.y / k2) : zero;
            az[q] = k2 != 0.0 ? make_double2(jz.x / k2, jz.y / k2) : zero;
        }
    }
}

// u = rho u / rho in real space (src/main3com.f95:804-806), in place over rho u
__global__ void k_diag_u(const double* rho, double* ux, double* uy, double* uz, size_t n)
{
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const double r = rho[q];
        ux[q] = ux[q] / r;
        uy[q] = uy[q] / r;
        uz[q] = uz[q] / r;
    }
}

// omega from the u spectra as src/main3com.f95:889-975 writes it
__global__ void k_diag_spec_w(const double2* ux, const double2* uy, const double2* uz, const double* kx,
                              const double* ky, const double* kz, int ld, int ny, int nz, double2* wx, double2* wy,
                              double2* wz)
{
    const size_t n = (size_t)ld * ny * nz;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(q % ld), j = (int)((q / ld) % ny), k = (int)(q / ((size_t)ld * ny));
        const double kxv = kx[i], kyv = ky[j], kzv = kz[k];
        const double2 x = ux[q], y = uy[q], z = uz[q];
        wx[q] = dg_sub(dg_ik(kyv, z), dg_ik(kzv, y));
        wy[q] = dg_sub(dg_ik(kxv, z), dg_ik(kzv, x));
        wz[q] = dg_sub(dg_ik(kxv, y), dg_ik(kyv, x));
    }
}

struct DiagFields {
    const double *rho, *ux, *uy, *uz, *e, *bx, *by, *bz, *ax, *ay, *az, *wx, *wy, *wz, *divb;
};

// Per-point terms of the averages of src/main3com.f95:1034-1064 (each divided
// by Nx Ny Nz there; here once, after the sum): P = CS^2 rho, |u|^2/2, E,
// |B|^2/(2 mu_0), |omega|, |omega|^2, u . omega, A . B, (u + A).(omega + B),
// div B; per-block sums into partial[block][10].
__global__ __launch_bounds__(256) void k_diag_reduce(DiagFields f, size_t n, double cs2, double mu0, double* partial)
{
    __shared__ double red[kDiagSums][256];
    double acc[kDiagSums] = {0};
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        const double ux = f.ux[q], uy = f.uy[q], uz = f.uz[q];
        const double bx = f.bx[q], by = f.by[q], bz = f.bz[q];
        const double ax = f.ax[q], ay = f.ay[q], az = f.az[q];
        const double wx = f.wx[q], wy = f.wy[q], wz = f.wz[q];
        const double w2 = wx * wx + wy * wy + wz * wz;
        const double wn = sqrt(w2);
        acc[0] += cs2 * f.rho[q];
        acc[1] += (ux * ux + uy * uy + uz * uz) / 2.0;
        acc[2] += f.e[q];
        acc[3] += (bx * bx + by * by + bz * bz) / (2.0 * mu0);
        acc[4] += wn;
        acc[5] += wn * wn;
        acc[6] += ux * wx + uy * wy + uz * wz;
        acc[7] += ax * bx + ay * by + az * bz;
        acc[8] += (ux + ax) * (wx + bx) + (uy + ay) * (wy + by) + (uz + az) * (wz + bz);
        acc[9] += f.divb[q];
    }
#pragma unroll
    for (int c = 0; c < kDiagSums; ++c) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int c = 0; c < kDiagSums; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + h];
        __syncthreads();
    }
    if ((int)threadIdx.x < kDiagSums) partial[(size_t)blockIdx.x * kDiagSums + threadIdx.x] = red[threadIdx.x][0];
}

// out[c] = sum over blocks of partial[.][c] / nn, in block order
__global__ void k_diag_final(const double* partial, int nblocks, double nn, double* out)
{
    const int c = threadIdx.x;
    if (c >= kDiagSums) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += partial[(size_t)b * kDiagSums + c];
    out[c] = s / nn;
}

unsigned diag_grid(size_t n)
{
    const size_t b = (n + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b ? b : 1));
}

// spectral (padded, destroyed) -> real (nx, ny, nz) / (Nx Ny Nz)
int diag_c2r(const MhdPlan& p, double2* spec, double* real, hipStream_t s)
{
    int rc;
    if ((rc = col_fft(spec, p.nz, p.ld, (long)p.ld * p.ny, p.ld, p.ny, +1, s))) return rc;
    if ((rc = col_fft(spec, p.ny, p.ld, p.ld, (long)p.ld * p.ny, p.nz, +1, s))) return rc;
    return row_c2r(spec, real, p.nx, p.ld, p.ny * p.nz, 1.0 / ((double)p.nx * p.ny * p.nz), s);
}

// real (nx, ny, nz) -> spectral (padded)
int diag_r2c(const MhdPlan& p, const double* real, double2* spec, hipStream_t s)
{
    int rc;
    TARA_HIP(hipMemsetAsync(spec, 0, p.fsz * sizeof(double2), s));
    if ((rc = row_r2c(real, spec, p.nx, p.ld, p.ny * p.nz, s))) return rc;
    if ((rc = col_fft(spec, p.ny, p.ld, p.ld, (long)p.ld * p.ny, p.nz, -1, s))) return rc;
    return col_fft(spec, p.nz, p.ld, (long)p.ld * p.ny, p.ld, p.ny, -1, s);
}

// Enqueues the diagnostics of the current state; the 10 averages land in
// device memory at out (ordered on the context's stream).
int mhd_diag_enqueue(tara_mhd* c, double* out)
{
    const MhdPlan& p = c->plan;
    hipStream_t s = c->s;
    int rc;
    if (!c->dpart.p) {
        for (auto& b : c->dspec)
            if ((rc = b.alloc(p.fsz * sizeof(double2)))) return rc;
        for (auto& b : c->dreal)
            if ((rc = b.alloc((size_t)p.nx * p.ny * p.nz * sizeof(double)))) return rc;
        if ((rc = c->dpart.alloc((size_t)(kDiagBlocks + 1) * kDiagSums * sizeof(double)))) return rc;
    }
    const size_t nsp = p.fsz, nr = (size_t)p.nx * p.ny * p.nz;
    double2 *A[3] = {c->dspec[0].as<double2>(), c->dspec[1].as<double2>(), c->dspec[2].as<double2>()};
    double2* divb = c->dspec[3].as<double2>();
    double2 *U[3] = {c->dspec[4].as<double2>(), c->dspec[5].as<double2>(), c->dspec[6].as<double2>()};
    double2 *W[3] = {c->dspec[7].as<double2>(), c->dspec[8].as<double2>(), c->dspec[9].as<double2>()};
    double2* tmp = c->dspec[10].as<double2>();
    double* R[15];
    for (int f = 0; f < 15; ++f) R[f] = c->dreal[f].as<double>();
    // real fields: 0 rho, 1-3 rho u (then u), 4 E, 5-7 B, 8-10 A, 11-13 omega, 14 div B
    hipLaunchKernelGGL(k_diag_spec_b, dim3(diag_grid(nsp)), dim3(256), 0, s, c->st[5].as<double2>(),
                       c->st[6].as<double2>(), c->st[7].as<double2>(), p.kx.as<double>(), p.ky.as<double>(),
                       p.kz.as<double>(), p.ld, p.ny, p.nz, A[0], A[1], A[2], divb);
    TARA_HIP(hipGetLastError());
    for (int f = 0; f < 8; ++f) {  // rho, rho u, E, B (the state's order)
        TARA_HIP(hipMemcpyAsync(tmp, c->st[f].p, nsp * sizeof(double2), hipMemcpyDeviceToDevice, s));
        if ((rc = diag_c2r(p, tmp, R[f], s))) return rc;
    }
    for (int f = 0; f < 3; ++f)
        if ((rc = diag_c2r(p, A[f], R[8 + f], s))) return rc;
    if ((rc = diag_c2r(p, divb, R[14], s))) return rc;
    hipLaunchKernelGGL(k_diag_u, dim3(diag_grid(nr)), dim3(256), 0, s, R[0], R[1], R[2], R[3], nr);
    TARA_HIP(hipGetLastError());
    for (int f = 0; f < 3; ++f)
        if ((rc = diag_r2c(p, R[1 + f], U[f], s))) return rc;
    hipLaunchKernelGGL(k_diag_spec_w, dim3(diag_grid(nsp)), dim3(256), 0, s, U[0], U[1], U[2], p.kx.as<double>(),
                       p.ky.as<double>(), p.kz.as<double>(), p.ld, p.ny, p.nz, W[0], W[1], W[2]);
    TARA_HIP(hipGetLastError());
    for (int f = 0; f < 3; ++f)
        if ((rc = diag_c2r(p, W[f], R[11 + f], s))) return rc;
    DiagFields df{R[0], R[1], R[2], R[3], R[4], R[5], R[6], R[7], R[8], R[9], R[10], R[11], R[12], R[13], R[14]};
    double* part = c->dpart.as<double>();
    hipLaunchKernelGGL(k_diag_reduce, dim3(kDiagBlocks), dim3(256), 0, s, df, nr, p.cs * p.cs, 1.0, part);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, s, part, kDiagBlocks, (double)nr, out);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

}  // namespace

extern "C" int tara_mhd_energy(tara_mhd* c, double* out)
{
    TARA_CHECK(c && out, TARA_ERR_ARG, "null argument");
    DevBuf row;
    int rc;
    if ((rc = row.alloc(kDiagSums * sizeof(double))) || (rc = mhd_diag_enqueue(c, row.as<double>()))) return rc;
    TARA_HIP(hipMemcpyAsync(out, row.p, kDiagSums * sizeof(double), hipMemcpyDeviceToHost, c->s));
    TARA_HIP(hipStreamSynchronize(c->s));
    return TARA_OK;
}

extern "C" int tara_mhd_record_energy(tara_mhd* c, int nrec, int every, double* rows)
{
    TARA_CHECK(c && rows && nrec >= 0 && every >= 1, TARA_ERR_ARG, "bad argument");
    if (nrec == 0) return TARA_OK;
    DevBuf dev;
    int rc;
    if ((rc = dev.alloc((size_t)nrec * kDiagSums * sizeof(double)))) return rc;
    for (int r = 0; r < nrec; ++r) {
        if ((rc = tara_mhd_advance(c, every))) return rc;
        if ((rc = mhd_diag_enqueue(c, dev.as<double>() + (size_t)r * kDiagSums))) return rc;
    }
    TARA_HIP(hipMemcpyAsync(rows, dev.p, (size_t)nrec * kDiagSums * sizeof(double), hipMemcpyDeviceToHost, c->s));
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
