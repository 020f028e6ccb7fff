// tcfft_doit_half.cpp -- harness backend: this library through the drop-in
// header tcfft_half.h, used the way the reference's backend uses its own
// (tcfft_doit_half.cpp:31-68 of the reference): tcfftCreate, the caller-side
// digit reversal by plan.radices on the host, double -> half staging, one
// device buffer transformed in place by tcfftExec, synchronise, half ->
// double.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "harness.h"
#include "tcfft_half.h"

namespace {
tcfftHandle g_plan{};
half *g_dev = nullptr;
std::vector<half> g_host;
}  // namespace

const char *backend_name() { return "tcfft (MI355X, drop-in tcfft_half.h)"; }

bool setup(const double *data, int n, int n_batch) {
    tcfftCreate(&g_plan, n, n_batch);
    if (!g_plan.impl) return false;  // tcfftCreate printed the status
    std::vector<int> rev(n);
    if (tcfft_gen_rev(n, g_plan.radices, g_plan.n_radices, rev.data()) != TCFFT_SUCCESS) return false;
    g_host.resize((size_t)2 * n * n_batch);
    for (size_t b = 0; b < (size_t)n_batch; ++b) {
        const double *src = data + 2 * b * n;
        half *dst = g_host.data() + 2 * b * n;
        for (int p = 0; p < n; ++p) {
            dst[2 * p] = __float2half((float)src[2 * (size_t)rev[p]]);
            dst[2 * p + 1] = __float2half((float)src[2 * (size_t)rev[p] + 1]);
        }
    }
    if (hipMalloc(&g_dev, g_host.size() * sizeof(half)) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc of %zu bytes failed\n", g_host.size() * sizeof(half));
        return false;
    }
    return hipMemcpy(g_dev, g_host.data(), g_host.size() * sizeof(half), hipMemcpyHostToDevice) == hipSuccess;
}

void doit(int iter) {
    for (int i = 0; i < iter; ++i) tcfftExec(g_plan, g_dev);
    (void)hipDeviceSynchronize();
}

void finalize(double *result) {
    (void)hipMemcpy(g_host.data(), g_dev, g_host.size() * sizeof(half), hipMemcpyDeviceToHost);
    for (size_t i = 0; i < g_host.size(); ++i) result[i] = __half2float(g_host[i]);
    (void)hipFree(g_dev);
    g_dev = nullptr;
    tcfftDestroy(&g_plan);
}
