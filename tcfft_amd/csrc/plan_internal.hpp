// plan_internal.hpp -- plan-level helpers shared between plan.cpp and the
// entry points defined in other translation units (not part of the C ABI).
#pragma once
#include <stddef.h>

#include "../../include/tcfft.h"

namespace tcfft {

// tcfft_exec_natural with the staging copy of the in-place cases in
// stage_buf (at least stage_cap bytes; TCFFT_INVALID_PLAN when a stage is
// needed and the buffer is missing or too small).
int exec_natural_staged(tcfft_plan plan, const void* in, void* out, int direction, void* stream, void* stage_buf,
                        size_t stage_cap);

// Bytes that tcfft_exec_natural(plan, in, out, ...) stages (0: none; in-place
// calls on 2D plans and on 1D plans whose first pass does not read natural
// order stage the whole buffer).
size_t natural_stage_bytes(tcfft_plan plan, const void* in, const void* out);

// The device a plan was created on (-1 for an invalid plan).
int plan_device(tcfft_plan plan);

}  // namespace tcfft
