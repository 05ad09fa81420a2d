// pbb_dispatch.cpp -- runtime dispatch over the machine count: each M in
// [2, 20] has its own instantiation of the explorer and bound kernels.
#include "pbb_kernels.h"

namespace pbb {

#define PBB_FOR_M(CALL, m)                \
    switch (m) {                          \
    case 2: return CALL(2);               \
    case 3: return CALL(3);               \
    case 4: return CALL(4);               \
    case 5: return CALL(5);               \
    case 6: return CALL(6);               \
    case 7: return CALL(7);               \
    case 8: return CALL(8);               \
    case 9: return CALL(9);               \
    case 10: return CALL(10);             \
    case 11: return CALL(11);             \
    case 12: return CALL(12);             \
    case 13: return CALL(13);             \
    case 14: return CALL(14);             \
    case 15: return CALL(15);             \
    case 16: return CALL(16);             \
    case 17: return CALL(17);             \
    case 18: return CALL(18);             \
    case 19: return CALL(19);             \
    case 20: return CALL(20);             \
    default: return hipErrorInvalidValue; \
    }

hipError_t launch_explore(const KArgs &a, int blocks, hipStream_t s)
{
#define C_(M) launch_explore_m<M>(a, blocks, s)
    PBB_FOR_M(C_, a.I.m)
#undef C_
}

hipError_t launch_team(const KArgs &a, hipStream_t s)
{
#define C_(M) launch_team_m<M>(a, s)
    PBB_FOR_M(C_, a.I.m)
#undef C_
}

hipError_t explore_occupancy(const KArgs &a, int *blocks_per_cu)
{
#define C_(M) occupancy_m<M>(a, blocks_per_cu)
    PBB_FOR_M(C_, a.I.m)
#undef C_
}

hipError_t launch_root(const KArgs &a, const RootPerm &root_perm, hipStream_t s)
{
#define C_(M) launch_root_m<M>(a, root_perm, s)
    PBB_FOR_M(C_, a.I.m)
#undef C_
}

hipError_t launch_bounds(int which, const KArgs &a, int count, const int *perms, const int *l1,
                         const int *l2, int arg, int *o1, int *o2, hipStream_t s)
{
#define C_(M) launch_bounds_m<M>(which, a, count, perms, l1, l2, arg, o1, o2, s)
    PBB_FOR_M(C_, a.I.m)
#undef C_
}

hipError_t launch_pool(int m, const KArgs &a, const PoolArgs &p, int groups, int init_count, hipStream_t s)
{
#define C_(M) launch_pool_m<M>(a, p, groups, init_count, s)
    PBB_FOR_M(C_, m)
#undef C_
}

hipError_t pool_occupancy(int m, const KArgs &a, const PoolArgs &p, int *groups_per_cu)
{
#define C_(M) pool_occupancy_m<M>(a, p, groups_per_cu)
    PBB_FOR_M(C_, m)
#undef C_
}

hipError_t launch_neh(int m, const int *ptms, const long long *off, const int *ns, const int *ms, int count,
                      int nmax, int *perms, int *costs, hipStream_t s)
{
#define C_(M) launch_neh_m<M>(ptms, off, ns, ms, count, nmax, perms, costs, s)
    PBB_FOR_M(C_, m)
#undef C_
}

hipError_t launch_ig(int m, const int *ptms, const long long *off, const int *ns, const int *ms, int count,
                     int chains, int iters, unsigned long long seed, const unsigned long long *thr, int *perms,
                     int *costs, hipStream_t s)
{
#define C_(M) launch_ig_m<M>(ptms, off, ns, ms, count, chains, iters, seed, thr, perms, costs, s)
    PBB_FOR_M(C_, m)
#undef C_
}

}  // namespace pbb
