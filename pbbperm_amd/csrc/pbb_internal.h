// pbb_internal.h -- host-side helpers shared by the engine (pbb_host.cpp) and
// the multi-instance pool (pbb_pool.cpp).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/pbb_engine.h"

namespace pbb {

// record a message for pbb_last_error() and return code
int fail(int code, const char *fmt, ...);
int validate(const pbb_config *c);
int check_shape(const int32_t *ptm, int n, int m);
int check_device();
// Stream replays report on stderr every 500 served requests when the
// environment sets PBB_REPLAY_PROGRESS (a long stream runs for minutes).
void replay_progress(const char *what, int before, int after, int count);

}  // namespace pbb

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            return pbb::fail(PBB_E_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                \
    } while (0)
