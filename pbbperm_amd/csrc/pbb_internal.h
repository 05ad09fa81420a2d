// pbb_internal.h -- host-side helpers shared by the engine (pbb_host.cpp) and
// the multi-instance pool (pbb_pool.cpp).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/pbb_engine.h"

#include <vector>

namespace pbb {

// Host copy of one instance's bound tables (bound_fsp_weak::init,
// fillMinTimeArrDep, bound_fsp_strong::remplirMachine/remplirLag/
// remplirTabJohnson): machine-major and job-major times, min heads/tails,
// the machine pairs a | b << 8 (all m(m-1)/2, lexicographic) and per pair the
// Johnson order {job, p_a, p_b, lag}.
struct HostTables {
    int n, m, mp, npairs;
    std::vector<int> ptm, ptmT, min_arr, min_dep, pair_ab;
    std::vector<int4> jtab;
};
void build_tables(const int32_t *ptm, int n, int m, HostTables &T);
// every job id 0..n-1 exactly once
bool is_perm(const int *p, int n);

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
