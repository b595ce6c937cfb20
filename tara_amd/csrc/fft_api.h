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
// The three above take any lengths and shapes: powers of two 8..4096 (x:
// 16..8192) whose batch fits the fast kernels' tiling run there, the rest on
// the mixed-radix transforms below (allocating, synchronising).
int any_col_fft(double2* a, int n, int ncols, long es, long pstride, int nplanes, int dir, hipStream_t st);
int any_row_r2c(const double* x, double2* X, int nx, int ld, int nrows, hipStream_t st);
int any_row_c2r(const double2* X, double* x, int nx, int ld, int nrows, double s, hipStream_t st);

}  // namespace tara
