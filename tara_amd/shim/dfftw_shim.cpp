// libtara_dfftw.so: the FFTW 3 Fortran entry points the reference's programs
// call (plan / execute / destroy for r2c and c2r in 1, 2 and 3 dimensions,
// and the thread set-up calls), implemented over libtara_hip.so's transforms,
// so that a kept main program links without FFTW.
//
// Replaces lib/fftw/api/f77funcs.h:26-34 (execute, destroy_plan), :200-218 and
// :294-312 (plan_dft_r2c / c2r _1d, _2d, _3d), :269 and :363 (execute_dft_r2c,
// execute_dft_c2r) and lib/fftw/threads/api.c's init_threads / plan_with_nthreads.
// Semantics: FFTW's (unnormalised, forward e^{-i}, halved axis = the Fortran
// first index); plans remember their arrays; flags and thread counts are
// accepted and ignored (every transform runs on the GPU).  Unlike FFTW, c2r
// leaves its input intact.  A failing transform prints the library's message and
// stops the program, as the reference has no error path to take.
//
// Every entry point is exported under both Fortran manglings the reference's
// sources produce: names written with a trailing underscore in the source
// (dfftw_execute_) get a second one from the compiler (dfftw_execute__), names
// written without get one (dfftw_plan_dft_r2c_1d_).
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "tara_hip.h"

namespace {

struct Plan {
    int r2c;         // 1: real -> half spectrum, 0: half spectrum -> real
    int rank;        // 1, 2 or 3
    int n[3];        // Fortran order: n[0] is the halved (first) dimension
    void* in;
    void* out;
};

void fail(const char* what)
{
    std::fprintf(stderr, "%s: libtara_hip: %s\n", what, tara_last_error());
    std::exit(1);
}

void run(const Plan* p, void* in, void* out)
{
    int rc;
    const double* i = static_cast<const double*>(in);
    double* o = static_cast<double*>(out);
    if (p->rank == 1)
        rc = p->r2c ? tara_fft_r2c_1d(p->n[0], i, o) : tara_fft_c2r_1d(p->n[0], i, o);
    else if (p->rank == 2)
        rc = p->r2c ? tara_fft_r2c_2d(p->n[0], p->n[1], i, o) : tara_fft_c2r_2d(p->n[0], p->n[1], i, o);
    else
        rc = p->r2c ? tara_fft_r2c_3d(p->n[0], p->n[1], p->n[2], i, o)
                    : tara_fft_c2r_3d(p->n[0], p->n[1], p->n[2], i, o);
    if (rc) fail("dfftw_execute");
}

void make(int64_t* plan, int r2c, int rank, int n0, int n1, int n2, void* in, void* out)
{
    Plan* p = new Plan{r2c, rank, {n0, n1, n2}, in, out};
    *plan = reinterpret_cast<int64_t>(p);
}

Plan* get(const int64_t* plan)
{
    if (!plan || !*plan) {
        std::fprintf(stderr, "dfftw: null plan\n");
        std::exit(1);
    }
    return reinterpret_cast<Plan*>(*plan);
}

}  // namespace

#define TARA_F77(name, args)       \
    extern "C" void name##_ args;  \
    extern "C" void name##__ args; \
    extern "C" void name##_ args

TARA_F77(dfftw_plan_dft_r2c_1d, (int64_t* p, const int* n, double* in, double* out, const int* flags))
{
    (void)flags;
    make(p, 1, 1, *n, 1, 1, in, out);
}
TARA_F77(dfftw_plan_dft_c2r_1d, (int64_t* p, const int* n, double* in, double* out, const int* flags))
{
    (void)flags;
    make(p, 0, 1, *n, 1, 1, in, out);
}
TARA_F77(dfftw_plan_dft_r2c_2d, (int64_t* p, const int* nx, const int* ny, double* in, double* out, const int* fl))
{
    (void)fl;
    make(p, 1, 2, *nx, *ny, 1, in, out);
}
TARA_F77(dfftw_plan_dft_c2r_2d, (int64_t* p, const int* nx, const int* ny, double* in, double* out, const int* fl))
{
    (void)fl;
    make(p, 0, 2, *nx, *ny, 1, in, out);
}
TARA_F77(dfftw_plan_dft_r2c_3d,
         (int64_t* p, const int* nx, const int* ny, const int* nz, double* in, double* out, const int* fl))
{
    (void)fl;
    make(p, 1, 3, *nx, *ny, *nz, in, out);
}
TARA_F77(dfftw_plan_dft_c2r_3d,
         (int64_t* p, const int* nx, const int* ny, const int* nz, double* in, double* out, const int* fl))
{
    (void)fl;
    make(p, 0, 3, *nx, *ny, *nz, in, out);
}
TARA_F77(dfftw_execute, (const int64_t* p))
{
    const Plan* q = get(p);
    run(q, q->in, q->out);
}
TARA_F77(dfftw_execute_dft_r2c, (const int64_t* p, double* in, double* out)) { run(get(p), in, out); }
TARA_F77(dfftw_execute_dft_c2r, (const int64_t* p, double* in, double* out)) { run(get(p), in, out); }
TARA_F77(dfftw_destroy_plan, (int64_t* p))
{
    if (p && *p) delete reinterpret_cast<Plan*>(*p);
    if (p) *p = 0;
}
TARA_F77(dfftw_init_threads, (int* iret))
{
    if (iret) *iret = 1;  // FFTW: nonzero on success
}
TARA_F77(dfftw_plan_with_nthreads, (const int* n)) { (void)n; }
TARA_F77(dfftw_cleanup_threads, (void)) {}
TARA_F77(dfftw_cleanup, (void)) {}

// The double-underscore manglings forward to the single-underscore bodies.
#define TARA_F77_ALIAS(name, params, call) \
    extern "C" void name##__ params { name##_ call; }
TARA_F77_ALIAS(dfftw_plan_dft_r2c_1d, (int64_t* p, const int* n, double* in, double* out, const int* f), (p, n, in, out, f))
TARA_F77_ALIAS(dfftw_plan_dft_c2r_1d, (int64_t* p, const int* n, double* in, double* out, const int* f), (p, n, in, out, f))
TARA_F77_ALIAS(dfftw_plan_dft_r2c_2d, (int64_t* p, const int* a, const int* b, double* in, double* out, const int* f),
               (p, a, b, in, out, f))
TARA_F77_ALIAS(dfftw_plan_dft_c2r_2d, (int64_t* p, const int* a, const int* b, double* in, double* out, const int* f),
               (p, a, b, in, out, f))
TARA_F77_ALIAS(dfftw_plan_dft_r2c_3d,
               (int64_t* p, const int* a, const int* b, const int* c, double* in, double* out, const int* f),
               (p, a, b, c, in, out, f))
TARA_F77_ALIAS(dfftw_plan_dft_c2r_3d,
               (int64_t* p, const int* a, const int* b, const int* c, double* in, double* out, const int* f),
               (p, a, b, c, in, out, f))
TARA_F77_ALIAS(dfftw_execute, (const int64_t* p), (p))
TARA_F77_ALIAS(dfftw_execute_dft_r2c, (const int64_t* p, double* in, double* out), (p, in, out))
TARA_F77_ALIAS(dfftw_execute_dft_c2r, (const int64_t* p, double* in, double* out), (p, in, out))
TARA_F77_ALIAS(dfftw_destroy_plan, (int64_t* p), (p))
TARA_F77_ALIAS(dfftw_init_threads, (int* iret), (iret))
TARA_F77_ALIAS(dfftw_plan_with_nthreads, (const int* n), (n))
TARA_F77_ALIAS(dfftw_cleanup_threads, (void), ())
TARA_F77_ALIAS(dfftw_cleanup, (void), ())
