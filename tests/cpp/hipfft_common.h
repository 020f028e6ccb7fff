// hipfft_common.h -- the hipFFT half-precision backend shared by the 1D and
// 2D harness binaries: fp16 staging of natural-order data, one in-place
// HIP_C_16F plan (hipfftXtMakePlanMany, dense layout, n_batch transforms),
// hipfftXtExec(HIPFFT_FORWARD) on the null stream.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <hipfft/hipfftXt.h>

#include <cstdio>
#include <vector>

struct HipfftHalf {
    hipfftHandle plan = nullptr;
    half *dev = nullptr;
    std::vector<half> host;

    bool setup(const double *data, size_t points, int rank, long long *dims, int n_batch) {
        host.resize(2 * points);
        for (size_t i = 0; i < 2 * points; ++i) host[i] = __float2half((float)data[i]);
        if (hipMalloc(&dev, host.size() * sizeof(half)) != hipSuccess ||
            hipMemcpy(dev, host.data(), host.size() * sizeof(half), hipMemcpyHostToDevice) != hipSuccess) {
            std::fprintf(stderr, "hipFFT backend: device buffer of %zu bytes failed\n", host.size() * sizeof(half));
            return false;
        }
        if (hipfftCreate(&plan) != HIPFFT_SUCCESS) {
            std::fprintf(stderr, "hipfftCreate failed\n");
            return false;
        }
        size_t work = 0;
        const hipfftResult r = hipfftXtMakePlanMany(plan, rank, dims, nullptr, 1, 0, HIP_C_16F, nullptr, 1, 0,
                                                    HIP_C_16F, n_batch, &work, HIP_C_16F);
        if (r != HIPFFT_SUCCESS) {
            std::fprintf(stderr, "hipfftXtMakePlanMany(HIP_C_16F) returned %d\n", (int)r);
            return false;
        }
        return true;
    }
    void doit(int iter) {
        for (int i = 0; i < iter; ++i) {
            const hipfftResult r = hipfftXtExec(plan, dev, dev, HIPFFT_FORWARD);
            if (r != HIPFFT_SUCCESS) {
                std::fprintf(stderr, "hipfftXtExec returned %d\n", (int)r);
                break;
            }
        }
        (void)hipDeviceSynchronize();
    }
    void finalize(double *result) {
        (void)hipMemcpy(host.data(), dev, host.size() * sizeof(half), hipMemcpyDeviceToHost);
        for (size_t i = 0; i < host.size(); ++i) result[i] = __half2float(host[i]);
    }
    ~HipfftHalf() {
        if (plan) (void)hipfftDestroy(plan);
        if (dev) (void)hipFree(dev);
    }
};
