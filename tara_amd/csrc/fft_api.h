// Device-array transform launchers (stream-ordered, no synchronisation).
#pragma once
#include "common.h"

namespace tara {

// Padded leading dimension (complex elements) of device half spectra for nx.
int spec_ld(int nx);
// In-place complex FFT of length n along a strided axis: columns i = 0..ld-1
// of each of nplanes planes, element stride es, plane stride pstride (complex
// elements).  dir = -1 forward, +1 backward.
int col_fft(double2* a, int n, int ld, long es, long pstride, int nplanes, int dir, hipStream_t st);
// x pass over nrows rows: real (pitch nx) -> half spectrum (pitch ld).
int row_r2c(const double* x, double2* X, int nx, int ld, int nrows, hipStream_t st);
// x pass over nrows rows: half spectrum (pitch ld) -> real (pitch nx), times s.
int row_c2r(const double2* X, double* x, int nx, int ld, int nrows, double s, hipStream_t st);

}  // namespace tara
