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
// derivative of gelu(x) = x Phi(x) (erf form, as applied in the GEMM epilogues)
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

// v0[b][c][p] = Wl[c] om[b][p] / sig[b] + bl[c]  (null outputs are skipped);
// grid (pixel blocks, nb * c)
__global__ void k_lift(const float* om, const float* sig, const float* wl, const float* bl, float* v, uint16_t* vb,
                       int c, size_t np)
{
    const int row = blockIdx.y, b = row / c, ch = row % c;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const float r = wl[ch] * (om[(size_t)b * np + p] / sig[b]) + bl[ch];
    const size_t i = (size_t)row * np + p;
    if (v) v[i] = r;
    if (vb) vb[i] = fno_f2bf(r);
}

// out[b][p] = om[b][p] + sig[b] (out[b][p] + b2); grid (pixel blocks, nb)
__global__ void k_finalize(float* out, const float* om, const float* sig, const float* b2, size_t np)
{
    const int b = blockIdx.y;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const size_t i = (size_t)b * np + p;
    out[i] = om[i] + sig[b] * (out[i] + b2[0]);
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

__global__ void k_act_back(float* dv, const float* pre, size_t n)
{
    FNO_LOOP(n) dv[i] *= gelu_grad(pre[i]);
}

// Reductions over every pixel of every sample run as RED_SPLIT workgroups per
// output, each over one contiguous slice, and a second pass over the partials.
constexpr int RED_SPLIT = 64;

// part[s][ch] = sum over slice s of q = b np + p < nb np of g[b][ch][p] w[q] / sig[b]
// (w, sig optional); grid (c, RED_SPLIT).
__global__ __launch_bounds__(256) void k_chan_dot(const float* g, const float* w, int c, size_t np, int nb,
                                                  const float* sig, float* part)
{
    __shared__ float sh[256];
    const int ch = blockIdx.x, sl = blockIdx.y;
    const size_t total = (size_t)nb * np, chunk = (total + RED_SPLIT - 1) / RED_SPLIT;
    const size_t q0 = sl * chunk, q1 = q0 + chunk < total ? q0 + chunk : total;
    float acc = 0.f;
    for (size_t q = q0 + threadIdx.x; q < q1; q += 256) {
        const size_t b = q / np, p = q - b * np;
        float v = g[(b * c + ch) * np + p];
        if (w) v *= w[q];
        if (sig) v /= sig[b];
        acc += v;
    }
    const float r = block_sum(acc, sh);
    if (threadIdx.x == 0) part[(size_t)sl * c + ch] = r;
}

// part[b][s] = sum of om[b][p]^2 over slice s of the pixels; grid (nb, RED_SPLIT).
__global__ __launch_bounds__(256) void k_rms_part(const float* om, size_t np, float* part)
{
    __shared__ float sh[256];
    const int b = blockIdx.x, sl = blockIdx.y;
    const size_t chunk = (np + RED_SPLIT - 1) / RED_SPLIT, p0 = sl * chunk, p1 = p0 + chunk < np ? p0 + chunk : np;
    const float* o = om + (size_t)b * np;
    float acc = 0.f;
    for (size_t p = p0 + threadIdx.x; p < p1; p += 256) acc += o[p] * o[p];
    const float r = block_sum(acc, sh);
    if (threadIdx.x == 0) part[(size_t)b * RED_SPLIT + sl] = r;
}

// sig[b] = sqrt(sum_s part[b][s] / np)
__global__ void k_rms_final(const float* part, int nb, size_t np, float* sig)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    float s = 0.f;
    for (int k = 0; k < RED_SPLIT; ++k) s += part[(size_t)b * RED_SPLIT + k];
    sig[b] = sqrtf(s / (float)np) + 1e-20f;
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
    DevBuf fxt, fy2, g2, hx;     // truncated DFT matrices (float)
    DevBuf fxtb, fy2b, g2b, hxb;  // the same in bfloat16
    // activations, sized for maxb samples
    DevBuf sig, v, vb, pre, x, xb, t1, y, yb, zf, zb, kv, hpre, hb, out, dproj, loss;
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
    hipEvent_t ev_anchor = nullptr;  // the anchor solver's take-over is done
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

// sig[b] = rms(omega_b): partial sums of squares over RED_SPLIT slices, then a
// second pass.
int rms(tara_fno* f, int nb, const float* om)
{
    hipLaunchKernelGGL(k_rms_part, dim3(nb, RED_SPLIT), dim3(256), 0, f->st, om, f->np, P<float>(f->part));
    hipLaunchKernelGGL(k_rms_final, dim3((nb + 255) / 256), dim3(256), 0, f->st, P<float>(f->part), nb, f->np,
                       P<float>(f->sig));
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

// out[ch] += sum_{b,p} g[b][ch][p] w[b][p] / sig[b]   (w, sig optional), for ch < c.
int chan_reduce(tara_fno* f, const float* g, int c, int nb, const float* w, const float* sig, float* out)
{
    hipLaunchKernelGGL(k_chan_dot, dim3(c, RED_SPLIT), dim3(256), 0, f->st, g, w, c, f->np, nb, sig,
                       P<float>(f->part));
    hipLaunchKernelGGL(k_sum_parts, dim3(fno_grid((size_t)c)), dim3(256), 0, f->st, P<float>(f->part), RED_SPLIT,
                       (size_t)c, out);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

// ------------------------------------------------------------------ forward
// out: prediction (device, [nb][n][n]).  train: keep what backward() needs (FP32
// activations, pre-activations, spectra of every layer); without it only the
// operands of the next product are written.
//
// Per layer, with F = nb C and bc = b C + c:
//   T1[y][q][bc]     q = p MX + kx (p: 0 real, 1 imaginary part)   x transform
//   X[p][mode][bc]   mode = ky_r MX + kx                            y transform
//   Y[p][mode][bc]                                                  per-mode contraction
//   Z[bc][y][q]                                                     inverse y transform
//   Kv[bc][y][x]                                                    inverse x transform
int forward(tara_fno* f, int nb, const float* om, float* out, bool train)
{
    const int n = f->n, c = f->c, mx = f->mx, my = f->my, M = (int)f->modes, q2 = 2 * mx;
    const long F = (long)nb * c, c4 = 4L * c * c;
    const size_t np = f->np;
    const bool bf = f->prec == 0;
    const int dt = bf ? GEMM_BF16 : GEMM_F32;
    float* par = f->par.as<float>();
    int rc;
    if ((rc = rms(f, nb, om))) return rc;
    const bool keep32 = train || !bf;  // FP32 activations needed
    hipLaunchKernelGGL(k_lift, dim3((unsigned)((np + 255) / 256), (unsigned)F), dim3(256), 0, f->st, om,
                       P<float>(f->sig), par + f->o_wl, par + f->o_bl, keep32 ? P<float>(f->v) : nullptr,
                       bf ? P<uint16_t>(f->vb) : nullptr, c, np);
    float* Y = P<float>(f->y);
    float* Z = P<float>(f->zf);
    float* kv = P<float>(f->kv);
    for (int l = 0; l < f->layers; ++l) {
        // without training, layer activations alternate between two slots
        const size_t vo = (size_t)(train ? l : l % 2) * f->field;
        const size_t vn = (size_t)(train ? l + 1 : (l + 1) % 2) * f->field;
        const void* vin = bf ? (const void*)P<uint16_t>(f->vb, vo) : (const void*)P<float>(f->v, vo);
        {  // 1. along x, per row y: T1[y][q][bc] = sum_x FxT[q][x] v[bc][y][x]
            GemmArgs g;
            g.m = q2, g.n = F, g.k = n;
            g.a = bf ? f->fxtb.p : f->fxt.p, g.sam = n, g.sak = 1;
            g.b = vin, g.sbk = 1, g.sbn = (long)np;
            g.c = f->t1.p, g.scm = F, g.scn = 1;
            g.nb1 = n, g.sb1 = n, g.sc1 = (long)q2 * F;
            if ((rc = gemm(g, dt, bf ? GEMM_STORE_BF16 : GEMM_STORE_F32, f->st))) return rc;
        }
        float* X = P<float>(f->x, (size_t)(train ? l : 0) * 2 * M * f->maxb * c);
        {  // 2. along y, per kx: X[p'][ky_r][kx][bc] = sum_{y,p} Fy2[p' MY + ky_r][2y + p] T1[y][p MX + kx][bc]
            GemmArgs g;
            g.m = 2 * my, g.n = F, g.k = 2L * n;
            g.a = bf ? f->fy2b.p : f->fy2.p, g.sam = 2L * n, g.sak = 1;
            g.b = f->t1.p, g.sbk = (long)mx * F, g.sbn = 1;
            g.scm = (long)mx * F, g.scn = 1;
            g.nb1 = mx, g.sb1 = F, g.sc1 = F;
            int epi = GEMM_STORE_F32;
            g.c = X;
            if (bf && train) {
                g.c_bf16 = f->xb.p;  // FP32 spectrum kept for the weight gradient, bf16 copy for the product
            } else if (bf) {
                g.c = f->xb.p;
                epi = GEMM_STORE_BF16;
            }
            if ((rc = gemm(g, dt, epi, f->st))) return rc;
        }
        // 3. per mode: Y[p'][mode][b][o] = sum_p sum_c X[p][mode][b][c] Rs[mode][p C + c][p' C + o]
        for (int p = 0; p < 2; ++p) {
            GemmArgs g;
            g.m = nb, g.n = c, g.k = c;
            g.a = bf ? (const void*)(P<uint16_t>(f->xb) + (size_t)p * M * F) : (const void*)(X + (size_t)p * M * F);
            g.sam = c, g.sak = 1;
            g.b = bf ? (const void*)P<uint16_t>(f->rsb, (size_t)l * M * c4 + (size_t)p * c * 2 * c)
                     : (const void*)P<float>(f->rs, (size_t)l * M * c4 + (size_t)p * c * 2 * c);
            g.sbk = 2L * c, g.sbn = 1;
            g.c = Y, g.scm = c, g.scn = 1;
            g.nb1 = M, g.sa1 = F, g.sb1 = c4, g.sc1 = F;
            g.nb2 = 2, g.sb2 = c, g.sc2 = (long)M * F;
            if (bf && p == 1) g.c_bf16 = f->yb.p;  // the finished sum, rounded for the next product
            if ((rc = gemm(g, dt, p == 0 ? GEMM_STORE_F32 : GEMM_ACCUM_F32, f->st))) return rc;
        }
        {  // 4. inverse along y, per kx: Z[bc][y][p' MX + kx] = sum_{p,ky_r} G2[2y + p'][p MY + ky_r] Y[p][ky_r][kx][bc]
            GemmArgs g;
            g.m = 2L * n, g.n = F, g.k = 2 * my;
            g.a = bf ? f->g2b.p : f->g2.p, g.sam = 2 * my, g.sak = 1;
            g.b = bf ? (const void*)P<uint16_t>(f->yb) : (const void*)Y, g.sbk = (long)mx * F, g.sbn = 1;
            g.c = bf ? f->zb.p : (void*)Z, g.scm = mx, g.scn = (long)n * q2;
            g.nb1 = mx, g.sb1 = F, g.sc1 = 1;
            if ((rc = gemm(g, dt, bf ? GEMM_STORE_BF16 : GEMM_STORE_F32, f->st))) return rc;
        }
        {  // 5. inverse along x: Kv[bc][y][x] = sum_j Z[bc][y][j] Hx[j][x]
            GemmArgs g;
            g.m = F * n, g.n = n, g.k = q2;
            g.a = bf ? f->zb.p : (const void*)Z, g.sam = q2, g.sak = 1;
            g.b = bf ? f->hxb.p : f->hx.p, g.sbk = n, g.sbn = 1;
            g.c = kv, g.scm = n, g.scn = 1;
            if ((rc = gemm(g, dt, GEMM_STORE_F32, f->st))) return rc;
        }
        {  // 6. channel mixing with bias and activation in the epilogue:
           //    v_next[b][o][p] = act(Kv[b][o][p] + sum_i W[o][i] v[b][i][p] + b[o])
            GemmArgs g;
            g.m = c, g.n = (long)np, g.k = c;
            g.a = bf ? (const void*)(P<uint16_t>(f->wb) + f->o_w + (size_t)l * c * c)
                     : (const void*)(par + f->o_w + (size_t)l * c * c);
            g.sam = c, g.sak = 1;
            g.b = vin, g.sbk = (long)np, g.sbn = 1;
            g.c = kv, g.scm = (long)np, g.scn = 1;
            g.nb1 = nb, g.sb1 = (long)c * np, g.sc1 = (long)c * np;
            g.bias = par + f->o_b + (size_t)l * c;
            g.act = l < f->layers - 1 ? 1 : 0;
            g.pre = train ? P<float>(f->pre, (size_t)l * f->field) : nullptr;
            g.c_f32 = keep32 ? P<float>(f->v, vn) : nullptr;
            g.c_bf16 = bf ? (void*)P<uint16_t>(f->vb, vn) : nullptr;
            if ((rc = gemm(g, dt, GEMM_ADD_TO, f->st))) return rc;
        }
    }
    // projection: hpre[b][j][p] = W1[j][c] v[b][c][p] + b1[j]; out = omega + s (w2 . gelu(hpre) + b2)
    const size_t vl = (size_t)(train ? f->layers : f->layers % 2) * f->field;
    const int H = f->hdim;
    {  // hidden layer, bias and GELU in the epilogue: h[b][j][p] = gelu(W1[j][c] v[b][c][p] + b1[j])
        GemmArgs g;
        g.m = H, g.n = (long)np, g.k = c;
        g.a = bf ? (const void*)(P<uint16_t>(f->wb) + f->o_w1) : (const void*)(par + f->o_w1), g.sam = c, g.sak = 1;
        g.b = bf ? (const void*)P<uint16_t>(f->vb, vl) : (const void*)P<float>(f->v, vl), g.sbk = (long)np,
        g.sbn = 1;
        g.c = f->hb.p, g.scm = (long)np, g.scn = 1;
        g.nb1 = nb, g.sb1 = (long)c * np, g.sc1 = (long)H * np;
        g.bias = par + f->o_b1;
        g.act = 1;
        g.pre = train ? P<float>(f->hpre) : nullptr;
        if ((rc = gemm(g, dt, GEMM_STORE_F32, f->st))) return rc;
    }
    {  // out[b][p] = sum_j w2[j] h[b][j][p]  (FP32)
        GemmArgs g;
        g.m = 1, g.n = (long)np, g.k = H;
        g.a = par + f->o_w2, g.sam = H, g.sak = 1;
        g.b = f->hb.p, g.sbk = (long)np, g.sbn = 1;
        g.c = out, g.scm = (long)np, g.scn = 1;
        g.nb1 = nb, g.sb1 = (long)H * np, g.sc1 = (long)np;
        if ((rc = gemm(g, GEMM_F32, GEMM_STORE_F32, f->st))) return rc;
    }
    hipLaunchKernelGGL(k_finalize, dim3((unsigned)((np + 255) / 256), (unsigned)nb), dim3(256), 0, f->st, out, om,
                       P<float>(f->sig), par + f->o_b2, np);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

// ------------------------------------------------------------------ backward
// Gradients of the mean relative L2 loss into f->grad, after forward(train) on the same input.
int backward(tara_fno* f, int nb, const float* om, const float* tgt)
{
    const int n = f->n, c = f->c, mx = f->mx, my = f->my, M = (int)f->modes, H = f->hdim, q2 = 2 * mx;
    const long F = (long)nb * c, c4 = 4L * c * c;
    const size_t np = f->np, field = (size_t)F * np;
    float* par = f->par.as<float>();
    float* gr = f->grad.as<float>();
    const float* sig = P<float>(f->sig);
    int rc;
    TARA_HIP(hipMemsetAsync(f->grad.p, 0, f->grad.bytes, f->st));
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, f->st, P<float>(f->out), tgt, sig, P<float>(f->dproj),
                       P<float>(f->loss), nb, np);
    TARA_HIP(hipGetLastError());
    // projection head: dw2[j] = sum h[b][j] dproj[b], db2 = sum dproj
    if ((rc = chan_reduce(f, P<float>(f->hb), H, nb, P<float>(f->dproj), nullptr, gr + f->o_w2)) ||
        (rc = chan_reduce(f, P<float>(f->dproj), 1, nb, nullptr, nullptr, gr + f->o_b2)))
        return rc;
    hipLaunchKernelGGL(k_proj2_back, dim3(fno_grid((size_t)nb * H * np)), dim3(256), 0, f->st, P<float>(f->dproj),
                       par + f->o_w2, P<float>(f->hpre), P<float>(f->dhpre), nb, H, np);
    if ((rc = chan_reduce(f, P<float>(f->dhpre), H, nb, nullptr, nullptr, gr + f->o_b1))) return rc;
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
    float* dz = P<float>(f->dz);
    float* dy = P<float>(f->dy);
    float* dx = P<float>(f->dx);
    float* dt1 = P<float>(f->dt1);
    for (int l = f->layers - 1; l >= 0; --l) {
        const size_t vo = (size_t)l * f->field;
        if (l < f->layers - 1)
            hipLaunchKernelGGL(k_act_back, dim3(fno_grid(field)), dim3(256), 0, f->st, dv,
                               P<float>(f->pre, (size_t)l * f->field), field);
        // dv now holds d(pre-activation of layer l)
        if ((rc = chan_reduce(f, dv, c, nb, nullptr, nullptr, gr + f->o_b + (size_t)l * c))) return rc;
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
        {  // spectral branch, backwards.  dZ[y][j][bc] = sum_x Hx[j][x] dpre[bc][y][x]
            GemmArgs g;
            g.m = q2, g.n = F, g.k = n;
            g.a = f->hx.p, g.sam = n, g.sak = 1;
            g.b = dv, g.sbk = 1, g.sbn = (long)np;
            g.c = dz, g.scm = F, g.scn = 1;
            g.nb1 = n, g.sb1 = n, g.sc1 = (long)q2 * F;
            if ((rc = gemm(g, GEMM_F32, GEMM_STORE_F32, f->st))) return rc;
        }
        {  // dY[p][ky_r][kx][bc] = sum_{y,p'} G2[2y + p'][p MY + ky_r] dZ[y][p' MX + kx][bc]
            GemmArgs g;
            g.m = 2 * my, g.n = F, g.k = 2L * n;
            g.a = f->g2.p, g.sam = 1, g.sak = 2 * my;
            g.b = dz, g.sbk = (long)mx * F, g.sbn = 1;
            g.c = dy, g.scm = (long)mx * F, g.scn = 1;
            g.nb1 = mx, g.sb1 = F, g.sc1 = F;
            if ((rc = gemm(g, GEMM_F32, GEMM_STORE_F32, f->st))) return rc;
        }
        const float* X = P<float>(f->x, (size_t)l * 2 * M * f->maxb * c);
        float* drs = P<float>(f->drs);
        for (int p = 0; p < 2; ++p) {  // dRs[mode][p C + c][p' C + o] = sum_b X[p][mode][b][c] dY[p'][mode][b][o]
            GemmArgs g;
            g.m = c, g.n = c, g.k = nb;
            g.a = X + (size_t)p * M * F, g.sam = 1, g.sak = c;
            g.b = dy, g.sbk = c, g.sbn = 1;
            g.c = drs + (size_t)p * c * 2 * c, g.scm = 2L * c, g.scn = 1;
            g.nb1 = M, g.sa1 = F, g.sb1 = F, g.sc1 = c4;
            g.nb2 = 2, g.sb2 = (long)M * F, g.sc2 = c;
            if ((rc = gemm(g, GEMM_F32, GEMM_STORE_F32, f->st))) return rc;
        }
        hipLaunchKernelGGL(k_fold_drs, dim3(fno_grid((size_t)M * c * c)), dim3(256), 0, f->st, drs,
                           gr + f->o_r + (size_t)l * f->r_sz, M, c);
        for (int pp = 0; pp < 2; ++pp) {  // dX[p][mode][b][c] = sum_p' sum_o dY[p'][mode][b][o] Rs[mode][p C + c][p' C + o]
            GemmArgs g;
            g.m = nb, g.n = c, g.k = c;
            g.a = dy + (size_t)pp * M * F, g.sam = c, g.sak = 1;
            g.b = P<float>(f->rs, (size_t)l * M * c4 + (size_t)pp * c), g.sbk = 1, g.sbn = 2L * c;
            g.c = dx, g.scm = c, g.scn = 1;
            g.nb1 = M, g.sa1 = F, g.sb1 = c4, g.sc1 = F;
            g.nb2 = 2, g.sb2 = 2L * c * c, g.sc2 = (long)M * F;
            if ((rc = gemm(g, GEMM_F32, pp == 0 ? GEMM_STORE_F32 : GEMM_ACCUM_F32, f->st))) return rc;
        }
        {  // dT1[y][p MX + kx][bc] = sum_{p',ky_r} Fy2[p' MY + ky_r][2y + p] dX[p'][ky_r][kx][bc]
            GemmArgs g;
            g.m = 2L * n, g.n = F, g.k = 2 * my;
            g.a = f->fy2.p, g.sam = 1, g.sak = 2L * n;
            g.b = dx, g.sbk = (long)mx * F, g.sbn = 1;
            g.c = dt1, g.scm = (long)mx * F, g.scn = 1;
            g.nb1 = mx, g.sb1 = F, g.sc1 = F;
            if ((rc = gemm(g, GEMM_F32, GEMM_STORE_F32, f->st))) return rc;
        }
        {  // dv_in[bc][y][x] += sum_q dT1[y][q][bc] FxT[q][x], per row y
            GemmArgs g;
            g.m = F, g.n = n, g.k = q2;
            g.a = dt1, g.sam = 1, g.sak = F;
            g.b = f->fxt.p, g.sbk = n, g.sbn = 1;
            g.c = dv2, g.scm = (long)np, g.scn = 1;
            g.nb1 = n, g.sa1 = (long)q2 * F, g.sc1 = n;
            if ((rc = gemm(g, GEMM_F32, GEMM_ACCUM_F32, f->st))) return rc;
        }
        std::swap(dv, dv2);
    }
    // lift: dWl[c] = sum dv0 omega / sig, dbl[c] = sum dv0
    if ((rc = chan_reduce(f, dv, c, nb, om, sig, gr + f->o_wl)) ||
        (rc = chan_reduce(f, dv, c, nb, nullptr, nullptr, gr + f->o_bl)))
        return rc;
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
    // Truncated DFT matrices (angles reduced mod n in integers before scaling):
    //   FxT[q][x]            q = p MX + kx: cos(a) (p = 0), -sin(a) (p = 1), a = 2 pi kx x / n
    //   Fy2[p' MY + r][2y+p] e^{-i b} applied to (re, im) = (p = 0, 1), output part p', b = 2 pi ky_r y / n
    //   G2[2y + p'][p MY + r] e^{+i b} applied to (Yre, Yim), output part p'
    //   Hx[q][x]             inverse along x with the half-spectrum weights 1 (kx = 0), 2 and 1/n^2
    const double tp = 2.0 * 3.14159265358979323846;
    auto kyv = [&](int r) { return r < my / 2 ? r : r - my; };
    std::vector<float> fxt(2 * mx * n), fy2(2 * my * 2 * n), g2(2 * n * 2 * my), hx(2 * mx * n);
    for (int x = 0; x < n; ++x)
        for (int k = 0; k < mx; ++k) {
            const double a = tp * (double)(((long)k * x) % n) / n;
            fxt[k * n + x] = (float)std::cos(a);
            fxt[(mx + k) * n + x] = (float)-std::sin(a);
            const double w = (k == 0 ? 1.0 : 2.0) / ((double)n * n);
            hx[k * n + x] = (float)(w * std::cos(a));
            hx[(mx + k) * n + x] = (float)(-w * std::sin(a));
        }
    for (int r = 0; r < my; ++r)
        for (int y = 0; y < n; ++y) {
            const long ph = (((long)kyv(r) * y) % n + n) % n;
            const double a = tp * (double)ph / n;
            const float cs = (float)std::cos(a), sn = (float)std::sin(a);
            // forward: re' = cos re + sin im, im' = -sin re + cos im
            fy2[(size_t)r * 2 * n + 2 * y] = cs;
            fy2[(size_t)r * 2 * n + 2 * y + 1] = sn;
            fy2[(size_t)(my + r) * 2 * n + 2 * y] = -sn;
            fy2[(size_t)(my + r) * 2 * n + 2 * y + 1] = cs;
            // inverse: re' = cos re - sin im, im' = sin re + cos im
            g2[(size_t)(2 * y) * 2 * my + r] = cs;
            g2[(size_t)(2 * y) * 2 * my + my + r] = -sn;
            g2[(size_t)(2 * y + 1) * 2 * my + r] = sn;
            g2[(size_t)(2 * y + 1) * 2 * my + my + r] = cs;
        }
    if ((rc = upload(f->fxt, fxt)) || (rc = upload(f->fy2, fy2)) || (rc = upload(f->g2, g2)) ||
        (rc = upload(f->hx, hx)) || (rc = upload_bf(f->fxtb, fxt)) || (rc = upload_bf(f->fy2b, fy2)) ||
        (rc = upload_bf(f->g2b, g2)) || (rc = upload_bf(f->hxb, hx)))
        return rc;
    // activations
    const size_t B = max_batch, np = f->np;
    f->field = B * c * np;
    const size_t field = f->field, hfield = B * H * np, spec = M * B * c2, t1 = B * c * 2 * n * mx;
    // partial sums: weight gradients, up to max(512 chunks, two per sample) x (H x C);
    // reductions, RED_SPLIT x max(H, B)
    const size_t parts = std::max(std::max<size_t>(512, 2 * B) * H * c, (size_t)RED_SPLIT * std::max<size_t>(H, B));
    if ((rc = f->sig.alloc(B * 4)) || (rc = f->v.alloc((L + 1) * field * 4)) ||
        (rc = f->vb.alloc((L + 1) * field * 2)) || (rc = f->pre.alloc(L * field * 4)) ||
        (rc = f->x.alloc(L * spec * 4)) || (rc = f->xb.alloc(spec * 2)) || (rc = f->t1.alloc(t1 * 4)) ||
        (rc = f->y.alloc(spec * 4)) || (rc = f->yb.alloc(spec * 2)) || (rc = f->zf.alloc(t1 * 4)) ||
        (rc = f->zb.alloc(t1 * 2)) ||
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

namespace {
// Training-batch indices passed by value (at most 64 samples per batch).
struct PairIdx {
    long i[64];
};

// xb[b] = inp[idx[b]], yb[b] = tgt[idx[b]] for b < batch; per floats per sample.
__global__ void k_gather_pairs(const float* __restrict__ inp, const float* __restrict__ tgt, long per, PairIdx idx,
                               int batch, float* __restrict__ xb, float* __restrict__ yb)
{
    const long total = (long)batch * per;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
        const long b = q / per, e = q - b * per;
        const long src = idx.i[b] * per + e;
        xb[q] = inp[src];
        yb[q] = tgt[src];
    }
}
}  // namespace

extern "C" int tara_fno_gather_pairs(tara_fno* f, int batch, const long* idx, long npairs, const float* inp,
                                     const float* tgt, float* xb, float* yb)
{
    TARA_CHECK(f && idx && inp && tgt && xb && yb && batch >= 1 && batch <= f->maxb && batch <= 64 && npairs >= 1,
               TARA_ERR_ARG, "fno: bad argument");
    PairIdx pi{};
    for (int b = 0; b < batch; ++b) {
        TARA_CHECK(idx[b] >= 0 && idx[b] < npairs, TARA_ERR_ARG, "fno: pair index %ld out of [0, %ld)", idx[b],
                   npairs);
        pi.i[b] = idx[b];
    }
    const long per = (long)f->n * f->n;
    hipLaunchKernelGGL(k_gather_pairs, dim3(fno_grid((size_t)batch * per)), dim3(256), 0, f->st, inp, tgt, per, pi,
                       batch, xb, yb);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
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

extern "C" int tara_fno_rollout(tara_fno* f, int batch, float* traj, int nsteps, int every, int nens,
                                tara_ens* const* checkers, tara_ens* anchor, int solver_steps, float tol,
                                int use_graph, float* check_err, int* fallbacks)
{
    const int nchk = every >= 1 ? nsteps / every : 0;
    // one ensemble per check (started as soon as its input exists), or one
    // ensemble holding every check (all advanced together after the rollout)
    const bool batched = nchk > 1 && nens == 1;
    TARA_CHECK(f && traj && batch >= 1 && batch <= f->maxb && nsteps >= 1 && every >= 1 &&
                   (nens == nchk || batched) && (nchk == 0 || (checkers && solver_steps >= 0)),
               TARA_ERR_ARG, "fno rollout: bad argument (one checker per check, or one for all)");
    for (int j = 0; j < nens; ++j) {
        const int runs = batched ? nchk * batch : batch;
        TARA_CHECK(checkers[j] && ens_grid_of(checkers[j], runs - 1) == f->n, TARA_ERR_ARG,
                   "fno rollout: checker %d needs %d runs of grid %d", j, runs, f->n);
    }
    TARA_CHECK(nchk == 0 || (anchor && ens_grid_of(anchor, batch - 1) == f->n), TARA_ERR_ARG,
               "fno rollout: the anchor solver needs %d runs of grid %d", batch, f->n);
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
        if (!f->ev_anchor) TARA_HIP(hipEventCreateWithFlags(&f->ev_anchor, hipEventDisableTiming));
        // capture the checkers' and the anchor's step graphs before any of them is in flight
        for (int j = 0; j < nens; ++j)
            if ((rc = tara_ens_advance(checkers[j], 0))) return rc;
        if ((rc = tara_ens_advance(anchor, 0))) return rc;
        // the anchor holds the last state known to be the FP64 solution (traj[0]
        // here) in FP64, at step ta of the rollout
        TARA_HIP(hipEventRecord(f->ev_anchor, f->st));
        TARA_HIP(hipStreamWaitEvent(ens_stream(anchor), f->ev_anchor, 0));
        for (int b = 0; b < batch; ++b)
            if ((rc = tara_ens_set_omega(anchor, b, traj + b * np))) return rc;
    }
    int ta = 0;
    const bool graph = use_graph != 0;
    if (graph && f->gbatch != batch && (rc = capture_step(f, batch))) return rc;
    std::vector<float> err((size_t)nchk * batch, 0.f), fresh(err.size());
    int nfall = 0, s = 0;
    while (s < nsteps) {
        int first_chk = nchk;  // checks launched in this pass: first_chk .. nchk-1
        for (int t = s; t < nsteps; ++t) {
            if (!batched && (t + 1) % every == 0 && (t + 1) / every <= nchk) {
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
        if (batched) {
            // every check whose step was (re)computed in this pass, in one ensemble advance
            first_chk = s / every;
            if (first_chk < nchk) {
                tara_ens* e = checkers[0];
                hipStream_t es = ens_stream(e);
                TARA_HIP(hipEventRecord(f->ev_in[0], f->st));
                TARA_HIP(hipStreamWaitEvent(es, f->ev_in[0], 0));
                for (int j = first_chk; j < nchk; ++j)
                    for (int b = 0; b < batch; ++b)
                        if ((rc = tara_ens_set_omega(e, j * batch + b,
                                                     traj + (size_t)((j + 1) * every - 1) * per + b * np)))
                            return rc;
                if ((rc = tara_ens_advance(e, solver_steps))) return rc;
                for (int j = first_chk; j < nchk; ++j)
                    for (int b = 0; b < batch; ++b)
                        if ((rc = tara_ens_get_omega(e, j * batch + b, f->chk.as<float>() + (size_t)j * per + b * np)))
                            return rc;
                for (int j = first_chk; j < nchk; ++j) TARA_HIP(hipEventRecord(f->ev_done[j], es));
            }
        }
        if (first_chk == nchk) break;
        for (int j = first_chk; j < nchk; ++j) TARA_HIP(hipStreamWaitEvent(f->st, f->ev_done[j], 0));
        hipLaunchKernelGGL(k_check_err, dim3(nchk * batch), dim3(256), 0, f->st, traj, f->chk.as<float>(),
                           f->cerr.as<float>(), batch, every, np);
        TARA_HIP(hipGetLastError());
        TARA_HIP(hipMemcpyAsync(fresh.data(), f->cerr.p, fresh.size() * 4, hipMemcpyDeviceToHost, f->st));
        TARA_HIP(hipStreamSynchronize(f->st));
        for (size_t q = (size_t)first_chk * batch; q < err.size(); ++q) err[q] = fresh[q];
        // At the earliest failing check the FP64 solver takes the trajectory
        // over: the anchor advances from its step ta to the failing step tf and
        // its fields replace traj[tf] for every sample of the batch, so every
        // state after a take-over is the FP64 solution from traj[0], not a
        // correction of an already drifted surrogate state.  The rollout is
        // then redone from tf, and tf becomes the anchor's step.
        int fail = -1;
        for (int j = first_chk; j < nchk && fail < 0; ++j)
            for (int b = 0; b < batch; ++b)
                if (!(err[(size_t)j * batch + b] <= tol)) fail = j;
        if (fail < 0) break;
        ++nfall;
        const int tf = (fail + 1) * every;
        if ((rc = tara_ens_advance(anchor, (tf - ta) * solver_steps))) return rc;
        for (int b = 0; b < batch; ++b)
            if ((rc = tara_ens_get_omega(anchor, b, traj + (size_t)tf * per + b * np))) return rc;
        TARA_HIP(hipEventRecord(f->ev_anchor, ens_stream(anchor)));
        TARA_HIP(hipStreamWaitEvent(f->st, f->ev_anchor, 0));
        ta = tf;
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
    if (f->ev_anchor) (void)hipEventDestroy(f->ev_anchor);
    delete f;
    return TARA_OK;
}
