// tcfft_half_2d.h -- drop-in replacement for tcFFT's 2D header (tcfft_half_2d.h:1-21).
//
// Same type name, public fields and function signatures as the reference:
//   struct tcfftHandle { Nx, Ny, N_batch, radices_x[3], radices_y[3],
//                        n_radices_x, n_radices_y, mergings[2], ... }
//   void tcfftCreate(tcfftHandle *plan, int nx, int ny, int n_batch);
//   void tcfftExec(tcfftHandle plan, half *data);
// as inline shims over the C ABI in tcfft.h.  As in the reference, this
// header and tcfft_half.h declare the same names and cannot be included in
// one translation unit.
//
// Contract kept from the reference (tcfft_half_2d.cu:1128-1149,
// tcfft_doit_half_2d.cpp:36-50): device buffer [n_batch][nx][ny] of fp16
// complex, ny contiguous, in place; input digit-reversed along x by
// radices_x and along y by radices_y; natural-order output; forward sign.
//
// Differences: nx and ny are any powers of two with 2 <= nx, 256 <= ny and
// per-axis radix lists of at most three entries (nx, ny <= 4096 here; the C
// ABI goes to 2^16).  The reference supports nx in {256, 512} and
// ny in {256, 512, 1024} (nx = 1024 selects an unassigned kernel pointer,
// tcfft_half_2d.cu:1164-1165).  Errors are reported on stderr and leave
// impl NULL, which tcfftExec refuses.
#ifndef TCFFT_HALF_2D_H_DROPIN
#define TCFFT_HALF_2D_H_DROPIN

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdio>

#include "tcfft.h"

struct tcfftHandle {
    int Nx, Ny, N_batch;
    int radices_x[3] = {16, 16, 2};
    int radices_y[3] = {16, 16, 2};
    int n_radices_x, n_radices_y;
    int mergings[2] = {0, 0};
    tcfft_plan impl = nullptr;
};

inline void tcfftCreate(tcfftHandle *plan, int nx, int ny, int n_batch) {
    plan->Nx = nx;
    plan->Ny = ny;
    plan->N_batch = n_batch;
    plan->n_radices_x = plan->n_radices_y = 0;
    plan->impl = nullptr;
    int rc = tcfft_plan_create_2d(&plan->impl, nx, ny, n_batch);
    if (rc == TCFFT_SUCCESS) rc = tcfft_plan_radices(plan->impl, 0, plan->radices_x, 3, &plan->n_radices_x);
    if (rc == TCFFT_SUCCESS) rc = tcfft_plan_radices(plan->impl, 1, plan->radices_y, 3, &plan->n_radices_y);
    if (rc != TCFFT_SUCCESS) {
        std::fprintf(stderr, "tcfftCreate(nx=%d, ny=%d, n_batch=%d): %s\n", nx, ny, n_batch,
                     tcfft_status_string(rc));
        if (plan->impl) tcfft_plan_destroy(plan->impl);
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

inline void tcfftExecInverse(tcfftHandle plan, half *data) {
    if (!plan.impl) {
        std::fprintf(stderr, "tcfftExecInverse: plan was not created successfully\n");
        return;
    }
    int rc = tcfft_exec(plan.impl, data, TCFFT_INVERSE, nullptr);
    if (rc != TCFFT_SUCCESS) std::fprintf(stderr, "tcfftExecInverse: %s\n", tcfft_status_string(rc));
}

inline void tcfftDestroy(tcfftHandle *plan) {
    if (plan->impl) tcfft_plan_destroy(plan->impl);
    plan->impl = nullptr;
}

#endif
