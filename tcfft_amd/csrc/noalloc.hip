// noalloc.hip -- tcfft_exec_natural_noalloc: natural-order execution that
// allocates nothing, for callers that cannot call hipMalloc (code under
// hipGraph capture, serving loops, harnesses that forbid allocation).
//
// tcfft_exec_natural stages the input of its in-place cases (2D plans, 1D
// plans whose first pass does not read natural order) in a workspace of the
// data's size that the plan allocates at creation.  This variant stages into
// `scratch`, one fixed array per GPU in the library's own code object, and
// otherwise runs the same kernels with the same results.  Calls whose
// staging exceeds TCFFT_NOALLOC_SCRATCH_BYTES are rejected
// (TCFFT_INVALID_SIZE).  Every call on one GPU shares that GPU's array:
// calls that overlap on one GPU (two streams, two host threads) race on it.
// No memset is needed: each call writes the whole staged range before its
// kernels read it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/tcfft.h"
#include "plan_internal.hpp"

namespace {

constexpr size_t kScratchBytes = TCFFT_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "scratch is a whole number of 256-byte blocks");

// Each GPU holds its own copy of this array until the process ends.  In host
// code `scratch` names a host-side stand-in, not device memory: the device
// address of a GPU's copy comes from hipGetSymbolAddress with that GPU
// current, and only that address is ever passed to the device.
static __device__ __align__(256) unsigned int scratch[kScratchBytes / 4];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};  // device address of each GPU's scratch, once resolved

// Device address of `scratch` on device dev: resolved on the first call for
// that device (with it current), then cached.
int scratch_on(int dev, void** out) {
    if (dev < 0 || dev >= kMaxDevices) return TCFFT_HIP_ERROR;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        int prev = 0;
        if (hipGetDevice(&prev) != hipSuccess) return TCFFT_HIP_ERROR;
        if (prev != dev && hipSetDevice(dev) != hipSuccess) return TCFFT_HIP_ERROR;
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (prev != dev) (void)hipSetDevice(prev);
        if (e != hipSuccess || !ptr) return TCFFT_HIP_ERROR;
        g_scratch[dev] = ptr;
    }
    *out = g_scratch[dev];
    return TCFFT_SUCCESS;
}

}  // namespace

extern "C" int tcfft_exec_natural_noalloc(tcfft_plan plan, const void* in, void* out, int direction, void* stream) {
    const int dev = tcfft::plan_device(plan);
    if (dev < 0) return TCFFT_INVALID_PLAN;
    const size_t need = tcfft::natural_stage_bytes(plan, in, out);
    if (need > kScratchBytes) return TCFFT_INVALID_SIZE;
    void* buf = nullptr;
    if (need) {
        const int rc = scratch_on(dev, &buf);
        if (rc) return rc;
    }
    return tcfft::exec_natural_staged(plan, in, out, direction, stream, buf, need ? kScratchBytes : 0);
}
