// harness.h -- the backend seam of the reference's test harness
// (doit.h:2-4 for 1D, doit_2d.h:2-4 for 2D).
//
// A backend stages the caller's double-precision complex data (interleaved
// re/im, signal b at data + 2*n*b) on the device as fp16, runs `iter` forward
// transforms in place, and returns the device result as doubles.  Each mode
// (accuracy.cpp, speed.cpp, accuracy_2d.cpp, speed_2d.cpp) is one main over
// this seam; linking it with a different backend gives the reference's
// tcfft_* / cufft_* pairs, here tcfft_* (this library through the drop-in
// headers) and hipfft_* (the vendor's half-precision FFT).
//
// setup* returns false (after printing why) when the backend cannot run the
// size; the mains then exit with status 1 instead of printing a result.
#pragma once

bool setup(const double *data, int n, int n_batch);
void doit(int iter);
void finalize(double *result);

bool setup_2d(const double *data, int nx, int ny, int n_batch);
void doit_2d(int iter);
void finalize_2d(double *result);

// Name of the backend, for diagnostics on stderr.
const char *backend_name();
