// Snapshot and data formats around the solver (host code):
//  - FP64 checkpoints of half spectra (a 64-byte header, then the fields in the
//    reference's (Nh, Ny[, Nz]) complex(8) layout);
//  - the reference's text snapshot, one line "x y omega ux uy" per grid point,
//    j fastest (src/main2csh.f95:170-178, x(i) = (i-1) dx, :99-101);
//  - a trajectory-pair store for surrogate training: real-space omega before and
//    after a fixed interval, stored as FP32 or BF16 (round to nearest even).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.h"

namespace tara {
namespace {

struct CkptHeader {
    char magic[8];      // "TARACKPT"
    uint32_t version;   // 1
    uint32_t kind;      // 1: complex(8) half spectra
    uint32_t nx, ny, nz, nfields;
    int64_t step;
    double time;
    uint8_t reserved[16];
};
static_assert(sizeof(CkptHeader) == 64, "checkpoint header is 64 bytes");

struct PairHeader {
    char magic[8];      // "TARAPAIR"
    uint32_t version;   // 1
    uint32_t dtype;     // 1: float32, 2: bfloat16
    uint32_t nx, ny;
    uint64_t npairs;
    double dt_pair;     // simulated time between the two fields of a pair
    double nu;
    uint8_t reserved[16];
};
static_assert(sizeof(PairHeader) == 64, "pair header is 64 bytes");

uint16_t to_bf16(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    const uint32_t bias = 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)((u + bias) >> 16);
}

}  // namespace
}  // namespace tara

using namespace tara;

struct tara_pairs {
    FILE* f = nullptr;
    PairHeader h{};
    std::vector<uint8_t> buf;
};

extern "C" int tara_ckpt_write(const char* path, int nx, int ny, int nz, int nfields, long step, double time,
                               const double* const fields[])
{
    TARA_CHECK(path && fields && nx > 0 && ny > 0 && nz > 0 && nfields > 0, TARA_ERR_ARG, "ckpt: bad argument");
    CkptHeader h{};
    std::memcpy(h.magic, "TARACKPT", 8);
    h.version = 1;
    h.kind = 1;
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    h.nfields = nfields;
    h.step = step;
    h.time = time;
    FILE* f = std::fopen(path, "wb");
    TARA_CHECK(f, TARA_ERR_ARG, "ckpt: cannot open %s", path);
    const size_t n = (size_t)(nx / 2 + 1) * ny * nz * 2;
    bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
    for (int k = 0; ok && k < nfields; ++k) ok = fields[k] && std::fwrite(fields[k], 8, n, f) == n;
    ok = (std::fclose(f) == 0) && ok;
    TARA_CHECK(ok, TARA_ERR_ARG, "ckpt: write to %s failed", path);
    return TARA_OK;
}

extern "C" int tara_ckpt_read_header(const char* path, int* nx, int* ny, int* nz, int* nfields, long* step,
                                     double* time)
{
    TARA_CHECK(path, TARA_ERR_ARG, "ckpt: null path");
    FILE* f = std::fopen(path, "rb");
    TARA_CHECK(f, TARA_ERR_ARG, "ckpt: cannot open %s", path);
    CkptHeader h{};
    const bool ok = std::fread(&h, sizeof h, 1, f) == 1;
    std::fclose(f);
    TARA_CHECK(ok && std::memcmp(h.magic, "TARACKPT", 8) == 0 && h.version == 1, TARA_ERR_ARG,
               "ckpt: %s is not a checkpoint", path);
    if (nx) *nx = (int)h.nx;
    if (ny) *ny = (int)h.ny;
    if (nz) *nz = (int)h.nz;
    if (nfields) *nfields = (int)h.nfields;
    if (step) *step = (long)h.step;
    if (time) *time = h.time;
    return TARA_OK;
}

extern "C" int tara_ckpt_read(const char* path, double* const fields[])
{
    TARA_CHECK(path && fields, TARA_ERR_ARG, "ckpt: bad argument");
    int nx, ny, nz, nf;
    int rc = tara_ckpt_read_header(path, &nx, &ny, &nz, &nf, nullptr, nullptr);
    if (rc) return rc;
    FILE* f = std::fopen(path, "rb");
    TARA_CHECK(f, TARA_ERR_ARG, "ckpt: cannot open %s", path);
    const size_t n = (size_t)(nx / 2 + 1) * ny * nz * 2;
    bool ok = std::fseek(f, (long)sizeof(CkptHeader), SEEK_SET) == 0;
    for (int k = 0; ok && k < nf; ++k) ok = fields[k] && std::fread(fields[k], 8, n, f) == n;
    std::fclose(f);
    TARA_CHECK(ok, TARA_ERR_ARG, "ckpt: %s is truncated", path);
    return TARA_OK;
}

extern "C" int tara_fort_write(const char* path, int nx, int ny, double lx, double ly, const double* omega,
                               const double* ux, const double* uy)
{
    TARA_CHECK(path && omega && ux && uy && nx > 0 && ny > 0, TARA_ERR_ARG, "fort: bad argument");
    FILE* f = std::fopen(path, "w");
    TARA_CHECK(f, TARA_ERR_ARG, "fort: cannot open %s", path);
    const double dx = lx / nx, dy = ly / ny;
    bool ok = true;
    for (int i = 0; i < nx && ok; ++i)
        for (int j = 0; j < ny && ok; ++j) {
            const size_t q = (size_t)j * nx + i;
            ok = std::fprintf(f, " %24.16E %24.16E %24.16E %24.16E %24.16E\n", (double)i * dx, (double)j * dy,
                              omega[q], ux[q], uy[q]) > 0;
        }
    ok = (std::fclose(f) == 0) && ok;
    TARA_CHECK(ok, TARA_ERR_ARG, "fort: write to %s failed", path);
    return TARA_OK;
}

extern "C" int tara_pairs_create(const char* path, int nx, int ny, int dtype, double dt_pair, double nu,
                                 tara_pairs** out)
{
    TARA_CHECK(path && out && nx > 0 && ny > 0 && (dtype == 1 || dtype == 2), TARA_ERR_ARG, "pairs: bad argument");
    *out = nullptr;
    auto* p = new tara_pairs;
    p->f = std::fopen(path, "wb");
    if (!p->f) {
        delete p;
        set_error("pairs: cannot open %s", path);
        return TARA_ERR_ARG;
    }
    std::memcpy(p->h.magic, "TARAPAIR", 8);
    p->h.version = 1;
    p->h.dtype = dtype;
    p->h.nx = nx;
    p->h.ny = ny;
    p->h.dt_pair = dt_pair;
    p->h.nu = nu;
    if (std::fwrite(&p->h, sizeof p->h, 1, p->f) != 1) {
        std::fclose(p->f);
        delete p;
        set_error("pairs: write to %s failed", path);
        return TARA_ERR_ARG;
    }
    *out = p;
    return TARA_OK;
}

extern "C" int tara_pairs_append(tara_pairs* p, const double* w_in, const double* w_out)
{
    TARA_CHECK(p && p->f && w_in && w_out, TARA_ERR_ARG, "pairs: bad argument");
    const size_t n = (size_t)p->h.nx * p->h.ny, es = p->h.dtype == 1 ? 4 : 2;
    p->buf.resize(2 * n * es);
    for (int s = 0; s < 2; ++s) {
        const double* src = s ? w_out : w_in;
        uint8_t* dst = p->buf.data() + s * n * es;
        for (size_t q = 0; q < n; ++q) {
            const float v = (float)src[q];
            if (es == 4) {
                std::memcpy(dst + 4 * q, &v, 4);
            } else {
                const uint16_t b = to_bf16(v);
                std::memcpy(dst + 2 * q, &b, 2);
            }
        }
    }
    TARA_CHECK(std::fwrite(p->buf.data(), 1, p->buf.size(), p->f) == p->buf.size(), TARA_ERR_ARG,
               "pairs: write failed");
    p->h.npairs += 1;
    return TARA_OK;
}

extern "C" int tara_pairs_close(tara_pairs* p)
{
    if (!p) return TARA_OK;
    bool ok = true;
    if (p->f) {
        ok = std::fseek(p->f, 0, SEEK_SET) == 0 && std::fwrite(&p->h, sizeof p->h, 1, p->f) == 1;
        ok = (std::fclose(p->f) == 0) && ok;
    }
    delete p;
    TARA_CHECK(ok, TARA_ERR_ARG, "pairs: finishing the file failed");
    return TARA_OK;
}
