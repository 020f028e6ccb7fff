// hipfft_doit_half.cpp -- harness backend: the vendor's half-precision FFT
// (hipFFT over rocFFT), the counterpart of the reference's cuFFT backend
// (cufft_doit_half.cpp:11-39): natural-order fp16 data, one in-place
// HIP_C_16F plan of n_batch transforms (hipfftXtMakePlanMany), HIPFFT_FORWARD.
#include "harness.h"
#include "hipfft_common.h"

namespace {
HipfftHalf g_be;
}

const char *backend_name() { return "hipFFT (rocFFT, HIP_C_16F)"; }

bool setup(const double *data, int n, int n_batch) {
    long long dims[1] = {n};
    return g_be.setup(data, (size_t)n * n_batch, 1, dims, n_batch);
}

void doit(int iter) { g_be.doit(iter); }

void finalize(double *result) { g_be.finalize(result); }
