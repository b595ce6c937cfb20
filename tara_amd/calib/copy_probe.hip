// Device-to-device copy of a known size: the calibration kernel for the
// FETCH_SIZE / WRITE_SIZE counters (bytes the counter reports for a copy of
// exactly `bytes` read and `bytes` written) and a measured HBM copy bandwidth.
//
//   tara_amd/calib/copy_probe [MiB]        (default 128 MiB per buffer)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ __launch_bounds__(256) void copy_probe(const double2* __restrict__ a, double2* __restrict__ b, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

#define CHECK(x)                                                                       \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));              \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

int main(int argc, char** argv) {
    const size_t mib = argc > 1 ? std::strtoul(argv[1], nullptr, 10) : 128;
    if (mib < 1 || mib > 4096) {
        std::fprintf(stderr, "size must be 1..4096 MiB\n");
        return 2;
    }
    const size_t bytes = mib << 20, n = bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 0, bytes));
    CHECK(hipMemset(b, 0, bytes));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const int reps = 20;
    // two launch shapes: 8192 workgroups striding over the buffer, and one
    // element per thread (n / 256 workgroups), which keeps more loads in flight
    const unsigned grids[2] = {8192u, (unsigned)((n + 255) / 256)};
    double best = 0.0;
    for (unsigned g : grids) {
        float ms = 0;
        for (int w = 0; w < 2; ++w) {
            CHECK(hipEventRecord(e0));
            for (int k = 0; k < reps; ++k) hipLaunchKernelGGL(copy_probe, dim3(g), dim3(256), 0, 0, a, b, n);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            CHECK(hipEventElapsedTime(&ms, e0, e1));
        }
        const double gbs = 2.0 * bytes * reps / (ms * 1e-3) / 1e9;
        best = gbs > best ? gbs : best;
        std::printf("copy_probe %zu MiB, %u workgroups: %.3f ms/launch, %.1f GB/s (read + write)\n", mib, g,
                    ms / reps, gbs);
    }
    std::printf("best %.1f GB/s\n", best);
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
