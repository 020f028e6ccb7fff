// hipfft_doit_half_2d.cpp -- harness backend: hipFFT's half-precision 2D
// transform, the counterpart of the reference's cufft_doit_half_2d.cpp
// (rank-2 HIP_C_16F plan of [nx][ny] images, in place, forward).
#include "harness.h"
#include "hipfft_common.h"

namespace {
HipfftHalf g_be;
}

const char *backend_name() { return "hipFFT (rocFFT, HIP_C_16F, 2D)"; }

bool setup_2d(const double *data, int nx, int ny, int n_batch) {
    long long dims[2] = {nx, ny};
    return g_be.setup(data, (size_t)nx * ny * n_batch, 2, dims, n_batch);
}

void doit_2d(int iter) { g_be.doit(iter); }

void finalize_2d(double *result) { g_be.finalize(result); }
