// Tiled transposes through LDS (32 x 32 complex tiles, one padding column) for
// the kx-fastest <-> ky-major conversions at the host boundary.
#include "layout.h"

namespace tara {

// out[c*rows + r] = in[r*cols + c] within each plane (rows x cols, row-major in).
__global__ __launch_bounds__(256) void k_transpose(const double2* __restrict__ in, double2* __restrict__ out,
                                                   int rows, int cols)
{
    __shared__ double2 tile[32][33];
    const size_t plane = (size_t)blockIdx.z * rows * cols;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
    for (int k = ty; k < 32; k += 8) {
        const int r = r0 + k, c = c0 + tx;
        if (r < rows && c < cols) tile[k][tx] = in[plane + (size_t)r * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const int c = c0 + k, r = r0 + tx;
        if (r < rows && c < cols) out[plane + (size_t)c * rows + r] = tile[tx][k];
    }
}

static int transpose(const double2* in, double2* out, int rows, int cols, int nplanes, hipStream_t st)
{
    dim3 g((cols + 31) / 32, (rows + 31) / 32, nplanes);
    hipLaunchKernelGGL(k_transpose, g, dim3(256), 0, st, in, out, rows, cols);
    TARA_HIP(hipGetLastError());
    return TARA_OK;
}

int to_kymajor(const double2* src, double2* dst, int nh, int ny, int nplanes, hipStream_t st)
{
    return transpose(src, dst, ny, nh, nplanes, st);
}

int from_kymajor(const double2* src, double2* dst, int nh, int ny, int nplanes, hipStream_t st)
{
    return transpose(src, dst, nh, ny, nplanes, st);
}

}  // namespace tara
