// pbb_internal.h -- host-side helpers shared by the engine (pbb_host.cpp) and
// the multi-instance pool (pbb_pool.cpp).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/pbb_engine.h"
#include "pbb_util.h"

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
int check_device();

}  // namespace pbb

#define HIPCHK(x)                                                                      \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess)                                                          \
            return pbb::fail(PBB_E_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                \
    } while (0)
