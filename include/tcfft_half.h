// tcfft_half.h -- drop-in replacement for tcFFT's 1D header (tcfft_half.h:1-21).
//
// Same type name, public fields and function signatures as the reference:
//   struct tcfftHandle { N, N_batch, radices[9], n_radices, mergings[3], n_mergings, ... }
//   void tcfftCreate(tcfftHandle *plan, int n, int n_batch);
//   void tcfftExec(tcfftHandle plan, half *data);
// implemented as inline shims over the C ABI in tcfft.h (libtcfft.so).
//
// Contract kept from the reference: `data` is a device buffer of n*n_batch
// interleaved fp16 complex values, transformed in place; its input must be
// digit-reversed by plan.radices[0..n_radices) (the caller's gen_rev,
// tcfft_doit_half.cpp:6-46); output is natural order; forward sign
// exp(-2*pi*i*nk/N), unnormalised; asynchronous on the null stream.
//
// Differences: any power of two 256 <= n <= 2^27 is supported (the
// reference's table lists ten sizes, tcfft_half.cu:1250-1330).  Where the
// reference silently ran a wrong plan (unlisted n, batch not a multiple of
// its tile), tcfftCreate prints the status to stderr and leaves impl NULL,
// and tcfftExec then refuses to run.  layer_0/layer_1/F_* become the opaque
// `impl`; mergings/n_mergings are kept for layout compatibility and hold
// the number of passes (n_mergings) the plan launches.  tcfftDestroy is an
// addition (the reference leaks its plan buffers).
#ifndef TCFFT_HALF_H_DROPIN
#define TCFFT_HALF_H_DROPIN

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdio>

#include "tcfft.h"

struct tcfftHandle {
    int N, N_batch;
    int radices[9] = {16, 16, 16, 16, 16, 16, 16, 16, 16};
    int n_radices;
    int mergings[3] = {0, 0, 0};
    int n_mergings;
    tcfft_plan impl = nullptr;
};

inline void tcfftCreate(tcfftHandle *plan, int n, int n_batch) {
    plan->N = n;
    plan->N_batch = n_batch;
    plan->n_radices = 0;
    plan->n_mergings = 0;
    plan->impl = nullptr;
    int rc = tcfft_plan_create_1d(&plan->impl, n, n_batch);
    if (rc != TCFFT_SUCCESS) {
        std::fprintf(stderr, "tcfftCreate(n=%d, n_batch=%d): %s\n", n, n_batch, tcfft_status_string(rc));
        plan->impl = nullptr;
        return;
    }
    rc = tcfft_plan_radices(plan->impl, 0, plan->radices, 9, &plan->n_radices);
    if (rc == TCFFT_SUCCESS) rc = tcfft_plan_num_passes(plan->impl, &plan->n_mergings);
    if (rc != TCFFT_SUCCESS) {
        std::fprintf(stderr, "tcfftCreate(n=%d): %s\n", n, tcfft_status_string(rc));
        tcfft_plan_destroy(plan->impl);
        plan->impl = nullptr;
    }
}

inline void tcfftExec(tcfftHandle plan, half *data) {
    if (!plan.impl) {
        std::fprintf(stderr, "tcfftExec: plan was not created successfully\n");
        return;
    }
    int rc = tcfft_exec(plan.impl, data, TCFFT_FORWARD, nullptr);
    if (rc != TCFFT_SUCCESS) std::fprintf(stderr, "tcfftExec: %s\n", tcfft_status_string(rc));
}

// Addition: adjoint (conjugate-twiddle, unnormalised) transform, same contract.
inline void tcfftExecInverse(tcfftHandle plan, half *data) {
    if (!plan.impl) {
        std::fprintf(stderr, "tcfftExecInverse: plan was not created successfully\n");
        return;
    }
    int rc = tcfft_exec(plan.impl, data, TCFFT_INVERSE, nullptr);
    if (rc != TCFFT_SUCCESS) std::fprintf(stderr, "tcfftExecInverse: %s\n", tcfft_status_string(rc));
}

// Addition: release the plan's device tables.
inline void tcfftDestroy(tcfftHandle *plan) {
    if (plan->impl) tcfft_plan_destroy(plan->impl);
    plan->impl = nullptr;
}

#endif
