// Host-side definition of the C-ABI context (include/srcwa.h), shared by the sources that implement
// entry points, and the internal hooks between them.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/srcwa.h"

struct srcwa_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t back = nullptr;     // second stream, created for the first plan with several chunks
    srcwa_pool* pool = nullptr;     // plans created on this context draw device memory from it
    int plan_flags = 0;             // SRCWA_PLAN_* flags of plans created on this context
};

// Message of the calling thread's last failed entry point (srcwa_last_error); runtime.cpp.
void srcwa_set_last_error(const std::string& msg);

// Per-device set-up of the no-allocation entry point's scratch array (dense.hip), run once per device
// by srcwa_context_create with that device current.  Returns a hipError_t.
int dense_scratch_setup(int device);
