// tcfft_doit_half_2d.cpp -- harness backend: this library through the
// drop-in header tcfft_half_2d.h, as the reference's 2D backend uses its own
// (tcfft_doit_half_2d.cpp:32-71 of the reference): digit reversal of both
// axes on the host by plan.radices_x / radices_y, half staging, tcfftExec in
// place, synchronise, half -> double.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "harness.h"
#include "tcfft_half_2d.h"

namespace {
tcfftHandle g_plan{};
half *g_dev = nullptr;
std::vector<half> g_host;
}  // namespace

const char *backend_name() { return "tcfft (MI355X, drop-in tcfft_half_2d.h)"; }

bool setup_2d(const double *data, int nx, int ny, int n_batch) {
    tcfftCreate(&g_plan, nx, ny, n_batch);
    if (!g_plan.impl) return false;
    std::vector<int> rx(nx), ry(ny);
    if (tcfft_gen_rev(nx, g_plan.radices_x, g_plan.n_radices_x, rx.data()) != TCFFT_SUCCESS ||
        tcfft_gen_rev(ny, g_plan.radices_y, g_plan.n_radices_y, ry.data()) != TCFFT_SUCCESS)
        return false;
    g_host.resize((size_t)2 * nx * ny * n_batch);
    for (size_t b = 0; b < (size_t)n_batch; ++b)
        for (int x = 0; x < nx; ++x) {
            const double *src = data + 2 * ((b * nx + rx[x]) * ny);
            half *dst = g_host.data() + 2 * ((b * nx + x) * ny);
            for (int y = 0; y < ny; ++y) {
                dst[2 * y] = __float2half((float)src[2 * (size_t)ry[y]]);
                dst[2 * y + 1] = __float2half((float)src[2 * (size_t)ry[y] + 1]);
            }
        }
    if (hipMalloc(&g_dev, g_host.size() * sizeof(half)) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc of %zu bytes failed\n", g_host.size() * sizeof(half));
        return false;
    }
    return hipMemcpy(g_dev, g_host.data(), g_host.size() * sizeof(half), hipMemcpyHostToDevice) == hipSuccess;
}

void doit_2d(int iter) {
    for (int i = 0; i < iter; ++i) tcfftExec(g_plan, g_dev);
    (void)hipDeviceSynchronize();
}

void finalize_2d(double *result) {
    (void)hipMemcpy(g_host.data(), g_dev, g_host.size() * sizeof(half), hipMemcpyDeviceToHost);
    for (size_t i = 0; i < g_host.size(); ++i) result[i] = __half2float(g_host[i]);
    (void)hipFree(g_dev);
    g_dev = nullptr;
    tcfftDestroy(&g_plan);
}
