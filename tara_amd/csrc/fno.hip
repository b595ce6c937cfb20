// Fourier neural operator (2D) on the matrix cores: a surrogate that maps
// real-space vorticity omega(t) on an N x N periodic grid to omega(t + T).  It
// has no counterpart in the reference.
//
// Network (width C, retained modes MX x MY, L Fourier layers, hidden H = 2C),
// per sample with s = rms(omega):
//   lift     v0 = Wl omega / s + bl                            (per pixel)
//   layer l  v  <- act(K v + W v + b)   (act = gelu, identity after the last layer)
//            K v = irfft2(R . rfft2(v) restricted to kx < MX, -MY/2 <= ky < MY/2)
//   project  out = omega + s (w2 . gelu(W1 v + b1) + b2)
// The per-sample scale s makes the map equivariant to the field's amplitude,
// and the residual form leaves the network the increment over T to learn.
//
// Every linear map is a GEMM on the MFMA units (gemm.hip).  With MX x MY
// retained modes, the transforms are truncated DFTs: per channel an
// (N x N) x (N x MX) product along x and an (MY x 2N) x (2N x MX) product along
// y, rather than full FFTs that would compute (N/2+1) N modes to keep MX MY.
// The per-mode contraction (a complex C x C matrix R per mode) is one batched
// real GEMM of size (B x 2C) x (2C x 2C) per mode.  In the bfloat16 precision
// the forward transforms, the contraction and the channel mixing take
// bfloat16 operands with FP32 accumulation; the FP32 precision runs every
// product on FP32 operands (32x32x2 MFMA), which lets the gradients be checked
// tightly.  The backward pass is FP32 MFMA throughout.
//
// Layouts: fields [b][c][y][x]; spectra X, Y [mode = ky_r MX + kx][b][2C]
// (real parts of the C channels, then imaginary parts); retained row ky_r holds
// ky = ky_r for ky_r < MY/2, else ky_r - MY.
//
// Parameters, one flat FP32 vector:
//   Wl[C] bl[C] | R[L][MX MY][2][C][C] (re, im; [in][out]) | W[L][C][C] ([out][in])
//   | b[L][C] | W1[H][C] | b1[H] | w2[H] | b2
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "ens.h"
#include "gemm.h"

namespace tara {
namespace {

__device__ __forceinline__ uint16_t fno_f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float gelu(float x)
{
    return 0.5f * x * (1.f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float x)
{
    return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

#define FNO_LOOP(n) \
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (n); i += (size_t)gridDim.x * blockDim.x)

// Sum over one 256-thread workgroup; every thread gets the result.
__device__ float block_sum(float v, float* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s /= 2) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const float r = sh[0];
    __syncthreads();
    return r;
}

__global__ void k_to_bf16(const float* x, uint16_t* y, size_t n)
{
    FNO_LOOP(n) y[i] = fno_f2bf(x[i]);
}

// sig[b] = rms of sample b (one workgroup per sample)
__global__ __launch_bounds__(256) void k_rms(const float* om, float* sig, size_t np)
{
    __shared__ float sh[256];
    const float* o = om + blockIdx.x * np;
    float a = 0.f;
    for (size_t p = threadIdx.x; p < np; p += 256) a += o[p] * o[p];
    const float s = block_sum(a, sh);
    if (threadIdx.x == 0) sig[blockIdx.x] = sqrtf(s / (float)np) + 1e-20f;
}

// v0[b][c][p] = Wl[c] om[b][p] / sig[b] + bl[c]  (null outputs are skipped)
__global__ void k_lift(const float* om, const float* sig, const float* wl, const float* bl, float* v, uint16_t* vb,
                       int nb, int c, size_t np)
{
    FNO_LOOP((size_t)nb * c * np)
    {
        const size_t p = i % np, ch = (i / np) % c, b = i / (np * c);
        const float r = wl[ch] * (om[b * np + p] / sig[b]) + bl[ch];
        if (v) v[i] = r;
        if (vb) vb[i] = fno_f2bf(r);
    }
}

// pre = acc + bias[c]; v = act ? gelu(pre) : pre.  Null outputs are skipped.
__global__ void k_bias_act(const float* acc, const float* bias, float* pre, float* v, uint16_t* vb, int c, size_t np,
                           size_t total, int act)
{
    FNO_LOOP(total)
    {
        const float p = acc[i] + bias[(i / np) % c];
        if (pre) pre[i] = p;
        const float r = act ? gelu(p) : p;
        if (v) v[i] = r;
        if (vb) vb[i] = fno_f2bf(r);
    }
}

// out[b][p] = om[b][p] + sig[b] (sum_j w2[j] h[b][j][p] + b2)
__global__ void k_proj2(const float* h, const float* w2, const float* b2, const float* om, const float* sig,
                        float* out, int nb, int hdim, size_t np)
{
    FNO_LOOP((size_t)nb * np)
    {
        const size_t b = i / np, p = i % np;
        float s = 0.f;
        for (int j = 0; j < hdim; ++j) s += w2[j] * h[(b * hdim + j) * np + p];
        out[i] = om[i] + sig[b] * (s + b2[0]);
    }
}

// Rs[m][k][n] (2C x 2C) = [[Cr, Ci], [-Ci, Cr]] from Cr = R[m][0], Ci = R[m][1] ([in][out]),
// so that [Xre Xim] Rs = [Re(X R), Im(X R)].
__global__ void k_build_rs(const float* r, float* rs, uint16_t* rsb, int modes, int c)
{
    const int c2 = 2 * c;
    FNO_LOOP((size_t)modes * c2 * c2)
    {
        const int n = (int)(i % c2), k = (int)((i / c2) % c2), m = (int)(i / ((size_t)c2 * c2));
        const float* cr = r + (size_t)m * 2 * c * c;
        const float* ci = cr + (size_t)c * c;
        const int ki = k % c, no = n % c;
        float val;
        if (k < c)
            val = n < c ? cr[ki * c + no] : ci[ki * c + no];
        else
            val = n < c ? -ci[ki * c + no] : cr[ki * c + no];
        rs[i] = val;
        rsb[i] = fno_f2bf(val);
    }
}

// dR[m][0] += dRs[:C,:C] + dRs[C:,C:], dR[m][1] += dRs[:C,C:] - dRs[C:,:C]
__global__ void k_fold_drs(const float* drs, float* dr, int modes, int c)
{
    const int c2 = 2 * c;
    FNO_LOOP((size_t)modes * c * c)
    {
        const int o = (int)(i % c), ii = (int)((i / c) % c), m = (int)(i / ((size_t)c * c));
        const float* d = drs + (size_t)m * c2 * c2;
        float* out = dr + (size_t)m * 2 * c * c;
        out[ii * c + o] += d[ii * c2 + o] + d[(c + ii) * c2 + c + o];
        out[(size_t)c * c + ii * c + o] += d[ii * c2 + c + o] - d[(c + ii) * c2 + o];
    }
}

// Loss = mean over samples of |out - tgt| / |tgt|; writes the per-sample loss
// and dproj[b][p] = dLoss/d(projection) = sig[b] (out - tgt) / (|out - tgt| |tgt| nb).
__global__ __launch_bounds__(256) void k_loss(const float* out, const float* tgt, const float* sig, float* dproj,
                                              float* loss, int nb, size_t np)
{
    __shared__ float sh[256];
    const int b = blockIdx.x;
    float d2 = 0.f, t2 = 0.f;
    for (size_t p = threadIdx.x; p < np; p += 256) {
        const float d = out[b * np + p] - tgt[b * np + p], t = tgt[b * np + p];
        d2 += d * d;
        t2 += t * t;
    }
    const float nd = sqrtf(block_sum(d2, sh)), nt = sqrtf(block_sum(t2, sh));
    const float scale = (nd > 0.f && nt > 0.f) ? sig[b] / (nd * nt * nb) : 0.f;
    for (size_t p = threadIdx.x; p < np; p += 256) dproj[b * np + p] = (out[b * np + p] - tgt[b * np + p]) * scale;
    if (threadIdx.x == 0) loss[b] = nt > 0.f ? nd / nt : 0.f;
}

// dhpre[b][j][p] = w2[j] dproj[b][p] gelu'(hpre[b][j][p])
__global__ void k_proj2_back(const float* dproj, const float* w2, const float* hpre, float* dhpre, int nb, int hdim,
                             size_t np)
{
    FNO_LOOP((size_t)nb * hdim * np)
    {
        const size_t p = i % np, j = (i / np) % hdim, b = i / (np * hdim);
        dhpre[i] = w2[j] * dproj[b * np + p] * gelu_grad(hpre[i]);
    }
}

// Workgroup j < hdim: dw2[j] += sum_{b,p} h[b][j][p] dproj[b][p]; workgroup hdim: db2 += sum dproj.
__global__ __launch_bounds__(256) void k_proj2_wgrad(const float* h, const float* dproj, float* dw2, float* db2,
                                                     int nb, int hdim, size_t np)
{
    __shared__ float sh[256];
    const int j = blockIdx.x;
    float acc = 0.f;
    for (size_t q = threadIdx.x; q < (size_t)nb * np; q += 256) {
        const size_t b = q / np, p = q % np;
        acc += (j < hdim ? h[(b * hdim + j) * np + p] : 1.f) * dproj[q];
    }
    const float s = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        if (j < hdim)
            dw2[j] += s;
        else
            db2[0] += s;
    }
}

// dbias[c] += sum_{b,p} g[b][c][p]  (one workgroup per channel)
__global__ __launch_bounds__(256) void k_bias_grad(const float* g, float* db, int nb, int c, size_t np)
{
    __shared__ float sh[256];
    const int ch = blockIdx.x;
    float acc = 0.f;
    for (size_t q = threadIdx.x; q < (size_t)nb * np; q += 256) acc += g[((q / np) * c + ch) * np + q % np];
    const float s = block_sum(acc, sh);
    if (threadIdx.x == 0) db[ch] += s;
}

__global__ void k_act_back(float* dv, const float* pre, size_t n)
{
    FNO_LOOP(n) dv[i] *= gelu_grad(pre[i]);
}

// Lift: dWl[c] += sum dv0 om / sig, dbl[c] += sum dv0  (one workgroup per channel)
__global__ __launch_bounds__(256) void k_lift_grad(const float* dv, const float* om, const float* sig, float* dwl,
                                                   float* dbl, int nb, int c, size_t np)
{
    __shared__ float sh[256];
    const int ch = blockIdx.x;
    float a0 = 0.f, a1 = 0.f;
    for (size_t q = threadIdx.x; q < (size_t)nb * np; q += 256) {
        const size_t b = q / np, p = q % np;
        const float g = dv[(b * c + ch) * np + p];
        a0 += g * (om[q] / sig[b]);
        a1 += g;
    }
    const float s0 = block_sum(a0, sh), s1 = block_sum(a1, sh);
    if (threadIdx.x == 0) {
        dwl[ch] += s0;
        dbl[ch] += s1;
    }
}

// out[i] += sum_{p < nparts} part[p size + i]
__global__ void k_sum_parts(const float* part, int nparts, size_t size, float* out)
{
    FNO_LOOP(size)
    {
        float s = 0.f;
        for (int p = 0; p < nparts; ++p) s += part[(size_t)p * size + i];
        out[i] += s;
    }
}

__global__ void k_adam(float* w, const float* g, float* m1, float* m2, size_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
    FNO_LOOP(n)
    {
        const float gi = g[i];
        m1[i] = b1 * m1[i] + (1.f - b1) * gi;
        m2[i] = b2 * m2[i] + (1.f - b2) * gi * gi;
        w[i] -= lr * (m1[i] / c1) / (sqrtf(m2[i] / c2) + eps);
    }
}

// Checker errors of a rollout: workgroup j*nb + b compares the surrogate state
// after step (j+1) E, traj[(j+1) E][b], with the solver's chk[j][b].
__global__ __launch_bounds__(256) void k_check_err(const float* traj, const float* chk, float* err, int nb, int every,
                                                   size_t np)
{
    __shared__ float sh[256];
    const int j = blockIdx.x / nb, b = blockIdx.x % nb;
    const float* a = traj + ((size_t)(j + 1) * every * nb + b) * np;
    const float* r = chk + ((size_t)j * nb + b) * np;
    float d2 = 0.f, r2 = 0.f;
    for (size_t p = threadIdx.x; p < np; p += 256) {
        const float d = a[p] - r[p];
        d2 += d * d;
        r2 += r[p] * r[p];
    }
    const float nd = block_sum(d2, sh), nr = block_sum(r2, sh);
    if (threadIdx.x == 0) err[blockIdx.x] = nr > 0.f ? sqrtf(nd / nr) : sqrtf(nd);
}

// Relative L2 distance per sample: err[b] = |a_b - r_b| / |r_b| (one workgroup per sample)
__global__ __launch_bounds__(256) void k_rel_err(const float* a, const float* r, float* err, size_t np)
{
    __shared__ float sh[256];
    const size_t o = blockIdx.x * np;
    float d2 = 0.f, r2 = 0.f;
    for (size_t p = threadIdx.x; p < np; p += 256) {
        const float d = a[o + p] - r[o + p];
        d2 += d * d;
        r2 += r[o + p] * r[o + p];
    }
    const float nd = block_sum(d2, sh), nr = block_sum(r2, sh);
    if (threadIdx.x == 0) err[blockIdx.x] = nr > 0.f ? sqrtf(nd / nr) : sqrtf(nd);
}

unsigned fno_grid(size_t n)
{
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b ? b : 1));
}

}  // namespace
}  // namespace tara

using namespace tara;

struct tara_fno {
    int n = 0, c = 0, mx = 0, my = 0, layers = 0, hdim = 0, maxb = 0, prec = 0;
    size_t np = 0, modes = 0, field = 0;
    size_t o_wl = 0, o_bl = 0, o_r = 0, o_w = 0, o_b = 0, o_w1 = 0, o_b1 = 0, o_w2 = 0, o_b2 = 0, nparam = 0;
    size_t r_sz = 0;  // spectral weights per layer
    DevBuf par, grad, m1, m2, wb;  // parameters, gradients, Adam moments, bfloat16 copy of the parameters
    DevBuf rs, rsb;                // stacked per-mode matrices of every layer (float, bfloat16)
    DevBuf fxc, fxs, fyre, fyim, gc, gs, hx;  // DFT matrices (float)
    DevBuf fxcb, fxsb, fyreb, fyimb, gcb, gsb;  // the same in bfloat16
    // activations, sized for maxb samples
    DevBuf sig, v, vb, pre, x, xb, t1, y, zf, kv, hpre, hb, out, dproj, loss;
    DevBuf dva, dvb, dz, dy, dx, dt1, drs, dhpre, part;
    int step = 0;
    hipStream_t st = nullptr;
    // surrogate rollouts: a captured inference step (gin -> gout), the checker's
    // results and per-check errors, events ordering the checker streams
    DevBuf gin, gout, chk, cerr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int gbatch = 0;
    std::vector<hipEvent_t> ev_in, ev_done;
};

namespace {

uint16_t host_bf16(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

int upload(DevBuf& d, const std::vector<float>& h)
{
    int rc = d.alloc(h.size() * 4);
    if (rc) return rc;
    TARA_HIP(hipMemcpy(d.p, h.data(), h.size() * 4, hipMemcpyHostToDevice));
    return TARA_OK;
}

int upload_bf(DevBuf& d, const std::vector<float>& h)
{
    std::vector<uint16_t> b(h.size());
    for (size_t i = 0; i < h.size(); ++i) b[i] = host_bf16(h[i]);
    int rc = d.alloc(b.size() * 2);
    if (rc) return rc;
    TARA_HIP(hipMemcpy(d.p, b.data(), b.size() * 2, hipMemcpyHostToDevice));
    return TARA_OK;
}

template <typename T>
T* P(const DevBuf& d, size_t off = 0)
{
    return d.as<T>() + off;
}

// out(m x n) += sum_b sum_k A_b(m, k) B_b(k, n) in FP32, split over k into
// chunks whose partial products are summed by a second kernel (a weight
// gradient: a small output reduced over every pixel of every sample).
int wgrad(tara_fno* f, long m, long n, long k, const float* A, long sam, long sak, long sab, const float* B, long sbk,
          long sbn, long sbb, int nb, float* out)
{
    long kc = k;
    while (kc > 64 && kc % 2 == 0 && (long)nb * (k / kc) < 256) kc /= 2;
    const long nch = k / kc;
    const size_t size = (size_t)m * n;
    TARA_CHECK(k % kc == 0 && (size_t)nb * nch * size * 4 <= f->part.bytes, TARA_ERR_STATE,
               "fno: gradient workspace too small");
    GemmArgs g;
    g.m = m, g.n = n, g.k = kc;
    g.a = A, g.sam = sam, g.sak = sak;
    g.b = B, g.sbk = sbk, g.sbn = sbn;
    g.c = f->part.p, g.scm = n, g.scn = 1;
    g.nb1 = (int)nch, g.sa1 = kc * sak, g.sb1 = kc * sbk, g.sc1 = (long)size;
    g.nb2 = nb, g.sa2 = sab, g.sb2 = sbb, g.sc2 = (long)(nch * size);
    int rc = gemm(g, GEMM_F32, GEMM_STORE_F32, f->st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum_parts, dim3(fno_grid(size)), dim3(256), 0, f->st, f->part.as<float>(),
                       (int)(nch * nb), size, out);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

// ------------------------------------------------------------------ forward
// out: prediction (device, [nb][n][n]).  train: keep what backward() needs (FP32
// activations, pre-activations, spectra of every layer); without it only the
// operands of the next product are written.
int forward(tara_fno* f, int nb, const float* om, float* out, bool train)
{
    const int n = f->n, c = f->c, c2 = 2 * c, mx = f->mx, my = f->my, M = (int)f->modes;
    const size_t np = f->np, field = (size_t)nb * c * np, xs = (size_t)M * nb * c2;
    const bool bf = f->prec == 0;
    const int dt = bf ? GEMM_BF16 : GEMM_F32;
    float* par = f->par.as<float>();
    int rc;
    hipLaunchKernelGGL(k_rms, dim3(nb), dim3(256), 0, f->st, om, P<float>(f->sig), np);
    const bool keep32 = train || !bf;  // FP32 activations needed
    hipLaunchKernelGGL(k_lift, dim3(fno_grid(field)), dim3(256), 0, f->st, om, P<float>(f->sig), par + f->o_wl,
                       par + f->o_bl, keep32 ? P<float>(f->v) : nullptr, bf ? P<uint16_t>(f->vb) : nullptr, nb, c,
                       np);
    for (int l = 0; l < f->layers; ++l) {
        // without training, layer activations alternate between two slots
        const size_t vo = (size_t)(train ? l : l % 2) * f->field, vn = (size_t)(train ? l + 1 : (l + 1) % 2) * f->field;
        const void* vin = bf ? (const void*)P<uint16_t>(f->vb, vo) : (const void*)P<float>(f->v, vo);
        // 1. along x: T1[b][c][part][y][kx] = sum_x v[b][c][y][x] Fx_part[x][kx]
        for (int part = 0; part < 2; ++part) {
            GemmArgs g;
            g.m = n, g.n = mx, g.k = n;
            g.a = vin, g.sam = n, g.sak = 1;
            g.b = bf ? (part ? f->fxsb.p : f->fxcb.p) : (part ? f->fxs.p : f->fxc.p), g.sbk = mx, g.sbn = 1;
            g.c = bf ? (void*)(P<uint16_t>(f->t1) + (size_t)part * n * mx)
                     : (void*)(P<float>(f->t1) + (size_t)part * n * mx);
            g.scm = mx, g.scn = 1;
            g.nb1 = c, g.sa1 = (long)np, g.sc1 = 2L * n * mx;
            g.nb2 = nb, g.sa2 = (long)c * np, g.sc2 = (long)c * 2 * n * mx;
            if ((rc = gemm(g, dt, bf ? GEMM_STORE_BF16 : GEMM_STORE_F32, f->st))) return rc;
        }
        // 2. along y: X[(ky_r, kx)][b][part C + c] = sum_y' FyP[ky_r][y'] T1[b][c][y'][kx], y' over (re, im) rows
        float* X = P<float>(f->x, (size_t)(train ? l : 0) * M * f->maxb * c2);
        for (int part = 0; part < 2; ++part) {
            GemmArgs g;
            g.m = my, g.n = mx, g.k = 2L * n;
            g.a = bf ? (part ? f->fyimb.p : f->fyreb.p) : (part ? f->fyim.p : f->fyre.p), g.sam = 2L * n, g.sak = 1;
            g.b = f->t1.p, g.sbk = mx, g.sbn = 1;
            g.c = X + (size_t)part * c, g.scm = (long)mx * nb * c2, g.scn = (long)nb * c2;
            g.nb1 = c, g.sb1 = 2L * n * mx, g.sc1 = 1;
            g.nb2 = nb, g.sb2 = (long)c * 2 * n * mx, g.sc2 = c2;
            if ((rc = gemm(g, dt, GEMM_STORE_F32, f->st))) return rc;
        }
        if (bf) hipLaunchKernelGGL(k_to_bf16, dim3(fno_grid(xs)), dim3(256), 0, f->st, X, P<uint16_t>(f->xb), xs);
        // 3. per-mode contraction: Y[m] = X[m] Rs[m]
        {
            GemmArgs g;
            g.m = nb, g.n = c2, g.k = c2;
            g.a = bf ? (const void*)P<uint16_t>(f->xb) : (const void*)X, g.sam = c2, g.sak = 1;
            g.b = bf ? (const void*)P<uint16_t>(f->rsb, (size_t)l * M * c2 * c2)
                     : (const void*)P<float>(f->rs, (size_t)l * M * c2 * c2);
            g.sbk = c2, g.sbn = 1;
            g.c = f->y.p, g.scm = c2, g.scn = 1;
            g.nb1 = M, g.sa1 = (long)nb * c2, g.sb1 = (long)c2 * c2, g.sc1 = (long)nb * c2;
            if ((rc = gemm(g, dt, bf ? GEMM_STORE_BF16 : GEMM_STORE_F32, f->st))) return rc;
        }
        // 4. inverse along y: Z[b][c][y][q MX + kx]; Zre = Gc Yre - Gs Yim, Zim = Gs Yre + Gc Yim
        float* Z = P<float>(f->zf);
        for (int q = 0; q < 2; ++q)
            for (int part = 0; part < 2; ++part) {
                const bool use_c = (q == part);
                GemmArgs g;
                g.m = n, g.n = mx, g.k = my;
                g.a = bf ? (use_c ? f->gcb.p : f->gsb.p) : (use_c ? f->gc.p : f->gs.p), g.sam = my, g.sak = 1;
                g.b = bf ? (const void*)(P<uint16_t>(f->y) + (size_t)part * c)
                         : (const void*)(P<float>(f->y) + (size_t)part * c);
                g.sbk = (long)mx * nb * c2, g.sbn = (long)nb * c2;
                g.c = Z + (size_t)q * mx, g.scm = 2L * mx, g.scn = 1;
                g.nb1 = c, g.sb1 = 1, g.sc1 = (long)n * 2 * mx;
                g.nb2 = nb, g.sb2 = c2, g.sc2 = (long)c * n * 2 * mx;
                g.alpha = (q == 0 && part == 1) ? -1.f : 1.f;
                if ((rc = gemm(g, dt, part == 0 ? GEMM_STORE_F32 : GEMM_ACCUM_F32, f->st))) return rc;
            }
        // 5. inverse along x: Kv[b][c][y][x] = sum_j Z[b][c][y][j] Hx[j][x]  (FP32)
        float* kv = P<float>(f->kv);
        {
            GemmArgs g;
            g.m = (long)nb * c * n, g.n = n, g.k = 2L * mx;
            g.a = Z, g.sam = 2L * mx, g.sak = 1;
            g.b = f->hx.p, g.sbk = n, g.sbn = 1;
            g.c = kv, g.scm = n, g.scn = 1;
            if ((rc = gemm(g, GEMM_F32, GEMM_STORE_F32, f->st))) return rc;
        }
        // 6. channel mixing: Kv[b][o][p] += sum_i W[o][i] v[b][i][p]
        {
            GemmArgs g;
            g.m = c, g.n = (long)np, g.k = c;
            g.a = bf ? (const void*)(P<uint16_t>(f->wb) + f->o_w + (size_t)l * c * c)
                     : (const void*)(par + f->o_w + (size_t)l * c * c);
            g.sam = c, g.sak = 1;
            g.b = vin, g.sbk = (long)np, g.sbn = 1;
            g.c = kv, g.scm = (long)np, g.scn = 1;
            g.nb1 = nb, g.sb1 = (long)c * np, g.sc1 = (long)c * np;
            if ((rc = gemm(g, dt, GEMM_ACCUM_F32, f->st))) return rc;
        }
        hipLaunchKernelGGL(k_bias_act, dim3(fno_grid(field)), dim3(256), 0, f->st, kv, par + f->o_b + (size_t)l * c,
                           train ? P<float>(f->pre, (size_t)l * f->field) : nullptr,
                           keep32 ? P<float>(f->v, vn) : nullptr, bf ? P<uint16_t>(f->vb, vn) : nullptr, c, np, field,
                           l < f->layers - 1 ? 1 : 0);
    }
    // projection: hpre[b][j][p] = W1[j][c] v[b][c][p] + b1[j]; out = omega + s (w2 . gelu(hpre) + b2)
    const size_t vl = (size_t)(train ? f->layers : f->layers % 2) * f->field;
    const int H = f->hdim;
    {
        GemmArgs g;
        g.m = H, g.n = (long)np, g.k = c;
        g.a = bf ? (const void*)(P<uint16_t>(f->wb) + f->o_w1) : (const void*)(par + f->o_w1), g.sam = c, g.sak = 1;
        g.b = bf ? (const void*)P<uint16_t>(f->vb, vl) : (const void*)P<float>(f->v, vl), g.sbk = (long)np,
        g.sbn = 1;
        g.c = f->kv.p, g.scm = (long)np, g.scn = 1;
        g.nb1 = nb, g.sb1 = (long)c * np, g.sc1 = (long)H * np;
        if ((rc = gemm(g, dt, GEMM_STORE_F32, f->st))) return rc;
    }
    const size_t hs = (size_t)nb * H * np;
    hipLaunchKernelGGL(k_bias_act, dim3(fno_grid(hs)), dim3(256), 0, f->st, P<float>(f->kv), par + f->o_b1,
                       train ? P<float>(f->hpre) : nullptr, P<float>(f->hb), (uint16_t*)nullptr, H, np, hs, 1);
    hipLaunchKernelGGL(k_proj2, dim3(fno_grid((size_t)nb * np)), dim3(256), 0, f->st, P<float>(f->hb),
                       par + f->o_w2, par + f->o_b2, om, P<float>(f->sig), out, nb, H, np);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

// ------------------------------------------------------------------ backward
// Gradients of the mean relative L2 loss into f->grad, after forward() on the same input.
int backward(tara_fno* f, int nb, const float* om, const float* tgt)
{
    const int n = f->n, c = f->c, c2 = 2 * c, mx = f->mx, my = f->my, M = (int)f->modes, H = f->hdim;
    const size_t np = f->np, field = (size_t)nb * c * np;
    float* par = f->par.as<float>();
    float* gr = f->grad.as<float>();
    const float* sig = P<float>(f->sig);
    int rc;
    TARA_HIP(hipMemsetAsync(f->grad.p, 0, f->grad.bytes, f->st));
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, f->st, P<float>(f->out), tgt, sig, P<float>(f->dproj),
                       P<float>(f->loss), nb, np);
    hipLaunchKernelGGL(k_proj2_wgrad, dim3(H + 1), dim3(256), 0, f->st, P<float>(f->hb), P<float>(f->dproj),
                       gr + f->o_w2, gr + f->o_b2, nb, H, np);
    hipLaunchKernelGGL(k_proj2_back, dim3(fno_grid((size_t)nb * H * np)), dim3(256), 0, f->st, P<float>(f->dproj),
                       par + f->o_w2, P<float>(f->hpre), P<float>(f->dhpre), nb, H, np);
    hipLaunchKernelGGL(k_bias_grad, dim3(H), dim3(256), 0, f->st, P<float>(f->dhpre), gr + f->o_b1, nb, H, np);
    const size_t vl = (size_t)f->layers * f->field;
    // dW1[j][c] += sum_{b,p} dhpre[b][j][p] v[b][c][p]
    if ((rc = wgrad(f, H, c, (long)np, P<float>(f->dhpre), (long)np, 1, (long)H * np, P<float>(f->v, vl), 1,
                    (long)np, (long)c * np, nb, gr + f->o_w1)))
        return rc;
    float* dv = P<float>(f->dva);
    float* dv2 = P<float>(f->dvb);
    {  // dv = W1^T dhpre
        GemmArgs g;
        g.m = c, g.n = (long)np, g.k = H;
        g.a = par + f->o_w1, g.sam = 1, g.sak = c;
        g.b = P<float>(f->dhpre), g.sbk = (long)np, g.sbn = 1;
        g.c = dv, g.scm = (long)np, g.scn = 1;
        g.nb1 = nb, g.sb1 = (long)H * np, g.sc1 = (long)c * np;
        if ((rc = gemm(g, GEMM_F32, GEMM_STORE_F32, f->st))) return rc;
    }
    for (int l = f->layers - 1; l >= 0; --l) {
        const size_t vo = (size_t)l * f->field;
        if (l < f->layers - 1)
            hipLaunchKernelGGL(k_act_back, dim3(fno_grid(field)), dim3(256), 0, f->st, dv,
                               P<float>(f->pre, (size_t)l * f->field), field);
        // dv now holds d(pre-activation of layer l)
        hipLaunchKernelGGL(k_bias_grad, dim3(c), dim3(256), 0, f->st, dv, gr + f->o_b + (size_t)l * c, nb, c, np);
        // channel mixing: dW[o][i] += sum dpre[b][o][p] v[b][i][p]; dv_in = W^T dpre
        if ((rc = wgrad(f, c, c, (long)np, dv, (long)np, 1, (long)c * np, P<float>(f->v, vo), 1, (long)np,
                        (long)c * np, nb, gr + f->o_w + (size_t)l * c * c)))
            return rc;
        {
            GemmArgs g;
            g.m = c, g.n = (long)np, g.k = c;
            g.a = par + f->o_w + (size_t)l * c * c, g.sam = 1, g.sak = c;
            g.b = dv, g.sbk = (long)np, g.sbn = 1;
            g.c = dv2, g.scm = (long)np, g.scn = 1;
            g.nb1 = nb, g.sb1 = (long)c * np, g.sc1 = (long)c * np;
            if ((rc = gemm(g, GEMM_F32, GEMM_STORE_F32, f->st))) return rc;
        }
        // spectral branch, backwards.  dZ = dpre Hx^T
        float* dz = P<float>(f->dz);
        {
            GemmArgs g;
            g.m = (long)nb * c * n, g.n = 2L * mx, g.k = n;
            g.a = dv, g.sam = n, g.sak = 1;
            g.b = f->hx.p, g.sbk = 1, g.sbn = n;
            g.c = dz, g.scm = 2L * mx, g.scn = 1;
            if ((rc = gemm(g, GEMM_F32, GEMM_STORE_F32, f->st))) return rc;
        }
        // dYre = Gc^T dZre + Gs^T dZim, dYim = -Gs^T dZre + Gc^T dZim  (layout [mode][b][2C])
        float* dy = P<float>(f->dy);
        for (int part = 0; part < 2; ++part)
            for (int q = 0; q < 2; ++q) {
                const bool use_c = (q == part);
                GemmArgs g;
                g.m = my, g.n = mx, g.k = n;
                g.a = use_c ? f->gc.p : f->gs.p, g.sam = 1, g.sak = my;
                g.b = dz + (size_t)q * mx, g.sbk = 2L * mx, g.sbn = 1;
                g.c = dy + (size_t)part * c, g.scm = (long)mx * nb * c2, g.scn = (long)nb * c2;
                g.nb1 = c, g.sb1 = (long)n * 2 * mx, g.sc1 = 1;
                g.nb2 = nb, g.sb2 = (long)c * n * 2 * mx, g.sc2 = c2;
                g.alpha = (part == 1 && q == 0) ? -1.f : 1.f;
                if ((rc = gemm(g, GEMM_F32, q == 0 ? GEMM_STORE_F32 : GEMM_ACCUM_F32, f->st))) return rc;
            }
        // dRs[m] = X[m]^T dY[m] ; dX[m] = dY[m] Rs[m]^T
        const float* X = P<float>(f->x, (size_t)l * M * f->maxb * c2);
        float* drs = P<float>(f->drs);
        {
            GemmArgs g;
            g.m = c2, g.n = c2, g.k = nb;
            g.a = X, g.sam = 1, g.sak = c2;
            g.b = dy, g.sbk = c2, g.sbn = 1;
            g.c = drs, g.scm = c2, g.scn = 1;
            g.nb1 = M, g.sa1 = (long)nb * c2, g.sb1 = (long)nb * c2, g.sc1 = (long)c2 * c2;
            if ((rc = gemm(g, GEMM_F32, GEMM_STORE_F32, f->st))) return rc;
        }
        hipLaunchKernelGGL(k_fold_drs, dim3(fno_grid((size_t)M * c * c)), dim3(256), 0, f->st, drs,
                           gr + f->o_r + (size_t)l * f->r_sz, M, c);
        float* dx = P<float>(f->dx);
        {
            GemmArgs g;
            g.m = nb, g.n = c2, g.k = c2;
            g.a = dy, g.sam = c2, g.sak = 1;
            g.b = P<float>(f->rs, (size_t)l * M * c2 * c2), g.sbk = 1, g.sbn = c2;
            g.c = dx, g.scm = c2, g.scn = 1;
            g.nb1 = M, g.sa1 = (long)nb * c2, g.sb1 = (long)c2 * c2, g.sc1 = (long)nb * c2;
            if ((rc = gemm(g, GEMM_F32, GEMM_STORE_F32, f->st))) return rc;
        }
        // dT1[b][c][y'][kx] = sum_part sum_ky_r FyP[ky_r][y'] dX_part[(ky_r, kx)][b][c]
        float* dt1 = P<float>(f->dt1);
        for (int part = 0; part < 2; ++part) {
            GemmArgs g;
            g.m = 2L * n, g.n = mx, g.k = my;
            g.a = part ? f->fyim.p : f->fyre.p, g.sam = 1, g.sak = 2L * n;
            g.b = dx + (size_t)part * c, g.sbk = (long)mx * nb * c2, g.sbn = (long)nb * c2;
            g.c = dt1, g.scm = mx, g.scn = 1;
            g.nb1 = c, g.sb1 = 1, g.sc1 = 2L * n * mx;
            g.nb2 = nb, g.sb2 = c2, g.sc2 = (long)c * 2 * n * mx;
            if ((rc = gemm(g, GEMM_F32, part == 0 ? GEMM_STORE_F32 : GEMM_ACCUM_F32, f->st))) return rc;
        }
        // dv_in += dT1_part Fx_part^T
        for (int part = 0; part < 2; ++part) {
            GemmArgs g;
            g.m = n, g.n = n, g.k = mx;
            g.a = dt1 + (size_t)part * n * mx, g.sam = mx, g.sak = 1;
            g.b = part ? f->fxs.p : f->fxc.p, g.sbk = 1, g.sbn = mx;
            g.c = dv2, g.scm = n, g.scn = 1;
            g.nb1 = c, g.sa1 = 2L * n * mx, g.sc1 = (long)np;
            g.nb2 = nb, g.sa2 = (long)c * 2 * n * mx, g.sc2 = (long)c * np;
            if ((rc = gemm(g, GEMM_F32, GEMM_ACCUM_F32, f->st))) return rc;
        }
        std::swap(dv, dv2);
    }
    hipLaunchKernelGGL(k_lift_grad, dim3(c), dim3(256), 0, f->st, dv, om, sig, gr + f->o_wl, gr + f->o_bl, nb, c, np);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

// bfloat16 copy of the parameters and the stacked per-mode matrices.
int refresh_weights(tara_fno* f)
{
    hipLaunchKernelGGL(k_to_bf16, dim3(fno_grid(f->nparam)), dim3(256), 0, f->st, f->par.as<float>(),
                       f->wb.as<uint16_t>(), f->nparam);
    const int c = f->c, c2 = 2 * c;
    const size_t M = f->modes;
    for (int l = 0; l < f->layers; ++l)
        hipLaunchKernelGGL(k_build_rs, dim3(fno_grid(M * c2 * c2)), dim3(256), 0, f->st,
                           f->par.as<float>() + f->o_r + (size_t)l * f->r_sz, P<float>(f->rs, l * M * c2 * c2),
                           P<uint16_t>(f->rsb, l * M * c2 * c2), (int)M, c);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

}  // namespace

extern "C" int tara_fno_create(int n, int width, int mx, int my, int layers, int max_batch, int precision,
                               unsigned seed, tara_fno** out)
{
    TARA_CHECK(out && n >= 8 && n % 8 == 0 && width >= 2 && mx >= 1 && my >= 2 && my % 2 == 0 && 2 * mx <= n &&
                   my <= n && layers >= 1 && max_batch >= 1 && (precision == 0 || precision == 1),
               TARA_ERR_ARG, "fno: bad shape (n %% 8 == 0, 2 mx <= n, even my <= n)");
    TARA_CHECK((long)max_batch * width <= 65535 && (long)mx * my <= 65535, TARA_ERR_ARG, "fno: batch too large");
    *out = nullptr;
    auto f = std::make_unique<tara_fno>();
    f->n = n, f->c = width, f->mx = mx, f->my = my, f->layers = layers, f->hdim = 2 * width, f->maxb = max_batch;
    f->prec = precision;
    f->np = (size_t)n * n;
    f->modes = (size_t)mx * my;
    const size_t c = width, c2 = 2 * c, M = f->modes, H = f->hdim, L = layers;
    f->r_sz = M * 2 * c * c;
    size_t o = 0;
    f->o_wl = o, o += c;
    f->o_bl = o, o += c;
    f->o_r = o, o += L * f->r_sz;
    f->o_w = o, o += L * c * c;
    f->o_b = o, o += L * c;
    f->o_w1 = o, o += H * c;
    f->o_b1 = o, o += H;
    f->o_w2 = o, o += H;
    f->o_b2 = o, o += 1;
    f->nparam = o;
    // Initial parameters, uniform: lift and mixing 1/sqrt(fan-in), spectral
    // weights 1/C^2, projection small so that the untrained network is near the
    // identity.
    std::vector<float> p(o, 0.f);
    uint64_t st = 0x9e3779b97f4a7c15ull ^ ((uint64_t)seed << 1);
    auto uni = [&]() {  // splitmix64 -> [-1, 1)
        st += 0x9e3779b97f4a7c15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        return (float)((double)(z >> 11) / 4503599627370496.0 - 1.0);
    };
    for (size_t i = 0; i < c; ++i) p[f->o_wl + i] = uni();
    for (size_t i = 0; i < L * f->r_sz; ++i) p[f->o_r + i] = uni() / (float)(c * c);
    for (size_t i = 0; i < L * c * c; ++i) p[f->o_w + i] = uni() / std::sqrt((float)c);
    for (size_t i = 0; i < H * c; ++i) p[f->o_w1 + i] = uni() / std::sqrt((float)c);
    for (size_t i = 0; i < H; ++i) p[f->o_w2 + i] = 0.1f * uni() / std::sqrt((float)H);
    int rc;
    if ((rc = upload(f->par, p))) return rc;
    if ((rc = f->grad.alloc(o * 4)) || (rc = f->m1.alloc(o * 4)) || (rc = f->m2.alloc(o * 4)) ||
        (rc = f->wb.alloc(o * 2)))
        return rc;
    TARA_HIP(hipMemset(f->m1.p, 0, o * 4));
    TARA_HIP(hipMemset(f->m2.p, 0, o * 4));
    if ((rc = f->rs.alloc(L * M * c2 * c2 * 4)) || (rc = f->rsb.alloc(L * M * c2 * c2 * 2))) return rc;
    // DFT matrices (angles reduced mod n in integers before scaling)
    const double tp = 2.0 * 3.14159265358979323846;
    auto kyv = [&](int r) { return r < my / 2 ? r : r - my; };
    std::vector<float> fxc(n * mx), fxs(n * mx), fyre(my * 2 * n), fyim(my * 2 * n), gc(n * my), gs(n * my),
        hx(2 * mx * n);
    for (int x = 0; x < n; ++x)
        for (int k = 0; k < mx; ++k) {
            const double a = tp * (double)(((long)k * x) % n) / n;
            fxc[x * mx + k] = (float)std::cos(a);   // T1re = sum_x v cos
            fxs[x * mx + k] = (float)-std::sin(a);  // T1im = -sum_x v sin
            const double w = (k == 0 ? 1.0 : 2.0) / ((double)n * n);
            hx[k * n + x] = (float)(w * std::cos(a));
            hx[(mx + k) * n + x] = (float)(-w * std::sin(a));
        }
    for (int r = 0; r < my; ++r)
        for (int y = 0; y < n; ++y) {
            const long ph = (((long)kyv(r) * y) % n + n) % n;
            const double a = tp * (double)ph / n, cs = std::cos(a), sn = std::sin(a);
            // e^{-i a} (T1re + i T1im): re = cos T1re + sin T1im, im = -sin T1re + cos T1im
            fyre[r * 2 * n + y] = (float)cs;
            fyre[r * 2 * n + n + y] = (float)sn;
            fyim[r * 2 * n + y] = (float)-sn;
            fyim[r * 2 * n + n + y] = (float)cs;
            gc[y * my + r] = (float)cs;  // e^{+i a}
            gs[y * my + r] = (float)sn;
        }
    if ((rc = upload(f->fxc, fxc)) || (rc = upload(f->fxs, fxs)) || (rc = upload(f->fyre, fyre)) ||
        (rc = upload(f->fyim, fyim)) || (rc = upload(f->gc, gc)) || (rc = upload(f->gs, gs)) ||
        (rc = upload(f->hx, hx)) || (rc = upload_bf(f->fxcb, fxc)) || (rc = upload_bf(f->fxsb, fxs)) ||
        (rc = upload_bf(f->fyreb, fyre)) || (rc = upload_bf(f->fyimb, fyim)) || (rc = upload_bf(f->gcb, gc)) ||
        (rc = upload_bf(f->gsb, gs)))
        return rc;
    // activations
    const size_t B = max_batch, np = f->np;
    f->field = B * c * np;
    const size_t field = f->field, hfield = B * H * np, spec = M * B * c2, t1 = B * c * 2 * n * mx;
    // weight-gradient partial products: up to max(256 chunks, one per sample) x (H x C)
    const size_t parts = std::max<size_t>(512, 2 * B) * H * c;
    if ((rc = f->sig.alloc(B * 4)) || (rc = f->v.alloc((L + 1) * field * 4)) ||
        (rc = f->vb.alloc((L + 1) * field * 2)) || (rc = f->pre.alloc(L * field * 4)) ||
        (rc = f->x.alloc(L * spec * 4)) || (rc = f->xb.alloc(spec * 2)) || (rc = f->t1.alloc(t1 * 4)) ||
        (rc = f->y.alloc(spec * 4)) || (rc = f->zf.alloc(t1 * 4)) ||
        (rc = f->kv.alloc(std::max(field, hfield) * 4)) || (rc = f->hpre.alloc(hfield * 4)) ||
        (rc = f->hb.alloc(hfield * 4)) || (rc = f->out.alloc(B * np * 4)) || (rc = f->dproj.alloc(B * np * 4)) ||
        (rc = f->loss.alloc(B * 4)) || (rc = f->dva.alloc(field * 4)) || (rc = f->dvb.alloc(field * 4)) ||
        (rc = f->dz.alloc(t1 * 4)) || (rc = f->dy.alloc(spec * 4)) || (rc = f->dx.alloc(spec * 4)) ||
        (rc = f->dt1.alloc(t1 * 4)) || (rc = f->drs.alloc(M * c2 * c2 * 4)) || (rc = f->dhpre.alloc(hfield * 4)) ||
        (rc = f->part.alloc(parts * 4)))
        return rc;
    TARA_HIP(hipStreamCreate(&f->st));
    if ((rc = refresh_weights(f.get()))) return rc;
    TARA_HIP(hipStreamSynchronize(f->st));
    *out = f.release();
    return TARA_OK;
}

extern "C" int tara_fno_num_params(tara_fno* f, long* n)
{
    TARA_CHECK(f && n, TARA_ERR_ARG, "null argument");
    *n = (long)f->nparam;
    return TARA_OK;
}

extern "C" int tara_fno_get_params(tara_fno* f, float* host)
{
    TARA_CHECK(f && host, TARA_ERR_ARG, "null argument");
    TARA_HIP(hipStreamSynchronize(f->st));
    TARA_HIP(hipMemcpy(host, f->par.p, f->nparam * 4, hipMemcpyDeviceToHost));
    return TARA_OK;
}

extern "C" int tara_fno_set_params(tara_fno* f, const float* host)
{
    TARA_CHECK(f && host, TARA_ERR_ARG, "null argument");
    TARA_HIP(hipStreamSynchronize(f->st));
    TARA_HIP(hipMemcpy(f->par.p, host, f->nparam * 4, hipMemcpyHostToDevice));
    int rc = refresh_weights(f);
    if (rc) return rc;
    TARA_HIP(hipStreamSynchronize(f->st));
    return TARA_OK;
}

extern "C" int tara_fno_forward(tara_fno* f, int batch, const float* omega, float* out)
{
    TARA_CHECK(f && omega && out && batch >= 1 && batch <= f->maxb, TARA_ERR_ARG, "fno: bad argument");
    int rc = forward(f, batch, omega, out, false);
    if (rc) return rc;
    TARA_HIP(hipStreamSynchronize(f->st));
    return TARA_OK;
}

extern "C" int tara_fno_grads(tara_fno* f, int batch, const float* omega, const float* target, float* loss,
                              float* grad_host)
{
    TARA_CHECK(f && omega && target && batch >= 1 && batch <= f->maxb, TARA_ERR_ARG, "fno: bad argument");
    int rc = forward(f, batch, omega, P<float>(f->out), true);
    if (rc) return rc;
    if ((rc = backward(f, batch, omega, target))) return rc;
    TARA_HIP(hipStreamSynchronize(f->st));
    if (loss) {
        std::vector<float> l(batch);
        TARA_HIP(hipMemcpy(l.data(), f->loss.p, batch * 4, hipMemcpyDeviceToHost));
        double s = 0;
        for (float v : l) s += v;
        *loss = (float)(s / batch);
    }
    if (grad_host) TARA_HIP(hipMemcpy(grad_host, f->grad.p, f->nparam * 4, hipMemcpyDeviceToHost));
    return TARA_OK;
}

extern "C" int tara_fno_train_step(tara_fno* f, int batch, const float* omega, const float* target, float lr,
                                   float* loss)
{
    int rc = tara_fno_grads(f, batch, omega, target, loss, nullptr);
    if (rc) return rc;
    f->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)f->step), c2 = 1.f - std::pow(b2, (float)f->step);
    hipLaunchKernelGGL(k_adam, dim3(fno_grid(f->nparam)), dim3(256), 0, f->st, f->par.as<float>(),
                       f->grad.as<float>(), f->m1.as<float>(), f->m2.as<float>(), f->nparam, lr, b1, b2, 1e-8f, c1,
                       c2);
    TARA_HIP(hipGetLastError());
    return refresh_weights(f);
}

extern "C" int tara_fno_rel_err(tara_fno* f, int batch, const float* a, const float* ref, float* err_host)
{
    TARA_CHECK(f && a && ref && err_host && batch >= 1 && batch <= f->maxb, TARA_ERR_ARG, "fno: bad argument");
    hipLaunchKernelGGL(k_rel_err, dim3(batch), dim3(256), 0, f->st, a, ref, P<float>(f->loss), f->np);
    TARA_HIP(hipGetLastError());
    TARA_HIP(hipMemcpyAsync(err_host, f->loss.p, batch * 4, hipMemcpyDeviceToHost, f->st));
    TARA_HIP(hipStreamSynchronize(f->st));
    return TARA_OK;
}

namespace {

void drop_fno_graph(tara_fno* f)
{
    if (f->gexec) (void)hipGraphExecDestroy(f->gexec);
    if (f->graph) (void)hipGraphDestroy(f->graph);
    f->gexec = nullptr;
    f->graph = nullptr;
    f->gbatch = 0;
}

// One inference step gin -> gout of `nb` samples, captured once per batch size.
int capture_step(tara_fno* f, int nb)
{
    drop_fno_graph(f);
    const size_t bytes = (size_t)f->maxb * f->np * 4;
    int rc;
    if ((!f->gin.p && (rc = f->gin.alloc(bytes))) || (!f->gout.p && (rc = f->gout.alloc(bytes)))) return rc;
    TARA_HIP(hipStreamBeginCapture(f->st, hipStreamCaptureModeThreadLocal));
    rc = forward(f, nb, f->gin.as<float>(), f->gout.as<float>(), false);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(f->st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    TARA_HIP(ce);
    f->graph = g;
    TARA_HIP(hipGraphInstantiate(&f->gexec, f->graph, nullptr, nullptr, 0));
    f->gbatch = nb;
    return TARA_OK;
}

// traj[t] -> traj[t+1] for nb samples on the context's stream.
int rollout_step(tara_fno* f, int nb, const float* in, float* out, bool graph)
{
    if (!graph) return forward(f, nb, in, out, false);
    const size_t bytes = (size_t)nb * f->np * 4;
    TARA_HIP(hipMemcpyAsync(f->gin.p, in, bytes, hipMemcpyDeviceToDevice, f->st));
    TARA_HIP(hipGraphLaunch(f->gexec, f->st));
    TARA_HIP(hipMemcpyAsync(out, f->gout.p, bytes, hipMemcpyDeviceToDevice, f->st));
    return TARA_OK;
}

}  // namespace

extern "C" int tara_fno_rollout(tara_fno* f, int batch, float* traj, int nsteps, int every, int nchk,
                                tara_ens* const* checkers, int solver_steps, float tol, int use_graph,
                                float* check_err, int* fallbacks)
{
    TARA_CHECK(f && traj && batch >= 1 && batch <= f->maxb && nsteps >= 1 && every >= 1 && nchk == nsteps / every &&
                   (nchk == 0 || (checkers && solver_steps >= 0)),
               TARA_ERR_ARG, "fno rollout: bad argument");
    for (int j = 0; j < nchk; ++j)
        for (int b = 0; b < batch; ++b)
            TARA_CHECK(checkers[j] && ens_grid_of(checkers[j], b) == f->n, TARA_ERR_ARG,
                       "fno rollout: checker %d needs %d runs of grid %d", j, batch, f->n);
    const size_t np = f->np, per = (size_t)batch * np;
    int rc;
    if (nchk > 0) {
        if ((!f->chk.p || f->chk.bytes < (size_t)nchk * per * 4) && (rc = f->chk.alloc((size_t)nchk * per * 4)))
            return rc;
        if ((!f->cerr.p || f->cerr.bytes < (size_t)nchk * batch * 4) &&
            (rc = f->cerr.alloc((size_t)nchk * batch * 4)))
            return rc;
        while ((int)f->ev_in.size() < nchk) {
            hipEvent_t a, b;
            TARA_HIP(hipEventCreateWithFlags(&a, hipEventDisableTiming));
            TARA_HIP(hipEventCreateWithFlags(&b, hipEventDisableTiming));
            f->ev_in.push_back(a);
            f->ev_done.push_back(b);
        }
        // capture the checkers' step graphs before any of them is in flight
        for (int j = 0; j < nchk; ++j)
            if ((rc = tara_ens_advance(checkers[j], 0))) return rc;
    }
    const bool graph = use_graph != 0;
    if (graph && f->gbatch != batch && (rc = capture_step(f, batch))) return rc;
    std::vector<float> err((size_t)nchk * batch, 0.f), fresh(err.size());
    int nfall = 0, s = 0;
    while (s < nsteps) {
        int first_chk = nchk;  // checks launched in this pass: first_chk .. nchk-1
        for (int t = s; t < nsteps; ++t) {
            if ((t + 1) % every == 0 && (t + 1) / every <= nchk) {
                // step t+1 is checked: the solver advances traj[t] by solver_steps
                const int j = (t + 1) / every - 1;
                first_chk = std::min(first_chk, j);
                hipStream_t es = ens_stream(checkers[j]);
                TARA_HIP(hipEventRecord(f->ev_in[j], f->st));
                TARA_HIP(hipStreamWaitEvent(es, f->ev_in[j], 0));
                for (int b = 0; b < batch; ++b)
                    if ((rc = tara_ens_set_omega(checkers[j], b, traj + (size_t)t * per + b * np))) return rc;
                if ((rc = tara_ens_advance(checkers[j], solver_steps))) return rc;
                for (int b = 0; b < batch; ++b)
                    if ((rc = tara_ens_get_omega(checkers[j], b, f->chk.as<float>() + (size_t)j * per + b * np)))
                        return rc;
                TARA_HIP(hipEventRecord(f->ev_done[j], es));
            }
            if ((rc = rollout_step(f, batch, traj + (size_t)t * per, traj + (size_t)(t + 1) * per, graph)))
                return rc;
        }
        if (first_chk == nchk) break;
        for (int j = first_chk; j < nchk; ++j) TARA_HIP(hipStreamWaitEvent(f->st, f->ev_done[j], 0));
        hipLaunchKernelGGL(k_check_err, dim3(nchk * batch), dim3(256), 0, f->st, traj, f->chk.as<float>(),
                           f->cerr.as<float>(), batch, every, np);
        TARA_HIP(hipGetLastError());
        TARA_HIP(hipMemcpyAsync(fresh.data(), f->cerr.p, fresh.size() * 4, hipMemcpyDeviceToHost, f->st));
        TARA_HIP(hipStreamSynchronize(f->st));
        for (size_t q = (size_t)first_chk * batch; q < err.size(); ++q) err[q] = fresh[q];
        // the earliest failing check hands its step to the solver; the rest is redone from there
        int fail = -1;
        for (int j = first_chk; j < nchk && fail < 0; ++j)
            for (int b = 0; b < batch; ++b)
                if (!(err[(size_t)j * batch + b] <= tol)) fail = j;
        if (fail < 0) break;
        ++nfall;
        const int tf = (fail + 1) * every;
        for (int b = 0; b < batch; ++b)
            if (!(err[(size_t)fail * batch + b] <= tol))
                TARA_HIP(hipMemcpyAsync(traj + (size_t)tf * per + b * np,
                                        f->chk.as<float>() + (size_t)fail * per + b * np, np * 4,
                                        hipMemcpyDeviceToDevice, f->st));
        s = tf;
    }
    TARA_HIP(hipStreamSynchronize(f->st));
    if (check_err) std::memcpy(check_err, err.data(), err.size() * 4);
    if (fallbacks) *fallbacks = nfall;
    return TARA_OK;
}

extern "C" int tara_fno_destroy(tara_fno* f)
{
    if (!f) return TARA_OK;
    if (f->st) {
        (void)hipStreamSynchronize(f->st);
        (void)hipStreamDestroy(f->st);
    }
    drop_fno_graph(f);
    for (auto e : f->ev_in) (void)hipEventDestroy(e);
    for (auto e : f->ev_done) (void)hipEventDestroy(e);
    delete f;
    return TARA_OK;
}
