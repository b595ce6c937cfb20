// Conversions between the reference's kx-fastest half spectra (Nh, Ny[, Nz])
// and the device's ky-major layout (element (i, j) at i*Ny + j per z plane).
#pragma once
#include "common.h"

namespace tara {

// dst[i*ny + j] = src[j*nh + i] for each of nplanes planes of nh*ny elements.
int to_kymajor(const double2* src, double2* dst, int nh, int ny, int nplanes, hipStream_t st);
// dst[j*nh + i] = src[i*ny + j], the inverse.
int from_kymajor(const double2* src, double2* dst, int nh, int ny, int nplanes, hipStream_t st);

}  // namespace tara
