// pbb_kernels.hip -- gfx950 kernels of the IVM branch-and-bound engine.
//
//   k_explore<M,REF> persistent explorer kernel: one IVM per 64-lane
//                  wavefront, 4 explorers per 256-thread workgroup.  The IVM
//                  matrix (int8 rows) and the front/back completion-time
//                  stacks of each explorer live in LDS; the interval vectors
//                  and the current row live in VGPRs.  Implements
//                  sequentialbb::next / unfold (multicore/src/sequentialbb.cpp:110-264)
//                  with ivm.cpp navigation and ivm_bound.cpp bounding,
//                  branching and pruning.  Work is balanced on the device: an
//                  explorer whose interval is exhausted registers in an idle
//                  ring; busy explorers split their interval as
//                  bbthread::shareWork (multicore/src/bbthread.cpp:39-72) and
//                  hand the upper part to an idle explorer.  No host round
//                  trip until the run ends (all idle, node budget, deadline
//                  or interrupt).  REF selects the bounding variant at
//                  compile time: 0 LB1 only (boundingMode 0), 1 LB1 + LB1'
//                  refinement (modes 1/2 as shipped), 2 LB1 + Johnson LB2.
//   k_bound_root<M> LB1 root bounding, ivm_bound::boundRoot (ivm_bound.cpp:273-330).
//   k_bounds<M>    batched bound_abstract calls over device buffers, one
//                  partial schedule per wavefront, on the explorer's own
//                  routines (LB1 children, LB1', LB2, makespan).
//
// Translation units: this file is compiled once per machine count M
// (-DPBB_M=<m>) so the 19 instantiations build in parallel.
#include "pbb_explorer.h"

namespace pbb {

// ------------------------------------------------------------ kernels
// Minimum resident waves per SIMD the compiler must allow (register
// budget): 4 for the LB1 explorer (128 VGPRs), free for the refinement
// variants.
#ifndef PBB_EXPLORE_MIN_WAVES
#define PBB_EXPLORE_MIN_WAVES 4
#endif

// The explorer loop of one wavefront (explorer e); SC is the memory scope of
// the explorers it shares work with.
template <int M, int REF, int SC>
__device__ __forceinline__ void explore_loop(const KArgs &a, const Lds &lds, int e)
{
    Explorer<M, REF, SC> X(a, lds, e, engine_inst(a), a.ctl, a.I.n, a.I.mp);
    X.local_best = uni((int)(unsigned)(ld_rlx<SC>(a.best) >> 32));
    // what the host or the previous launch left is visible at launch start
    int st = uni(ld_rlx<SC>(a.state + e));
    if (st == ST_UNFOLD) {
        if (X.start_unfold()) {
            st = ST_ACTIVE;
        } else {
            if (X.go_idle() && X.lane == 0) st_rel<SC>(&a.ctl->done, 1);
            st = ST_IDLE;
        }
    } else if (st == ST_ACTIVE) {
        X.load_state();
    } else {
        // An idle explorer enters the idle ring only once it runs (the host
        // leaves the ring empty at every launch), so donors only ever hand
        // work to wavefronts that are resident: one still waiting for a CU
        // (another kernel or process on the device, an occupancy estimate
        // that was too high) holds no work that the others wait for.
        X.register_idle();
    }

    unsigned iter = 0;
    bool hungry = true;
    while (true) {
        if (st == ST_IDLE) {
            // wait for a donation, global termination or a stop; the host
            // interrupt word (a PCIe round trip) is read by every 16th
            // explorer, at every 16th pass once its back-off is at the
            // longest, and the first to see it raises the device stop word
            int nap = 1;
            while (true) {
                // relaxed poll (an acquire would invalidate the L2 at every
                // pass); the fence pairs with the donor's release of the word
                int s = uni(ld_rlx<SC>(a.state + e));
                if (s == ST_ACTIVE) {
                    fence_acq<SC>();
                    st = ST_ACTIVE;
                    X.load_state();
                    // a donated interval is explored under the current incumbent,
                    // not the one this explorer saw before it went idle
                    X.take_best(ld_rlx<SC>(a.best));
                    break;
                }
                if (uni(ld_rlx<SC>(&a.ctl->done)) || uni(ld_rlx<SC>(&a.ctl->stop)) ||
                    (X.t_end && wall_clock64() >= X.t_end)) {
                    X.flush_counters();
                    return;
                }
                if (a.host_flag && (e & 15) == 0 && nap == 16 && X.lane == 0 &&
                    __hip_atomic_load(a.host_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
                    st_rlx<SC>(&a.ctl->stop, 1);
                if constexpr (SC == __HIP_MEMORY_SCOPE_WORKGROUP) __builtin_amdgcn_s_sleep(2);
                else X.idle_pause(nap);
            }
        }
        // idle-ring poll issued before the node's work so its latency is
        // hidden; none while the ramp holds this explorer's work back.  In
        // the workgroup team (at most 16 explorers, workgroup-scope atomics)
        // a donor that saw explorers waiting polls again at the next node; at
        // device scope every busy explorer racing for the same ring ticket at
        // every node made the ring words a hot spot, so there it polls every
        // PBB_POLL_EVERY nodes only.
        constexpr bool team = SC == __HIP_MEMORY_SCOPE_WORKGROUP;
        const bool poll = X.since >= a.ramp && ((team && hungry) || (iter & (PBB_POLL_EVERY - 1)) == 0u);
        unsigned rh = 0, rt = 0;
        if (poll) {
            rh = ld_rlx<SC>(&a.ctl->idle_head);
            rt = ld_rlx<SC>(&a.ctl->idle_tail);
        }
        const bool refresh = (iter & (PBB_BEST_EVERY - 1)) == 0u;
        unsigned long long bw = ~0ull;
        if (refresh) bw = ld_rlx<SC>(a.best);
        if (!X.step()) {
            if (X.go_idle() && X.lane == 0) st_rel<SC>(&a.ctl->done, 1);
            st = ST_IDLE;
            continue;
        }
        if (refresh) X.take_best(bw);
        if ((++iter & 63u) == 0u && X.check_stop()) {
            X.save_state(ST_ACTIVE);
            X.flush_counters();
            return;
        }
        if (poll) hungry = X.try_donate(rh, rt);
    }
}

template <int M, int REF>
__global__ __launch_bounds__(64 * PBB_EXPLORE_WAVES, REF == 0 ? PBB_EXPLORE_MIN_WAVES : 1) void k_explore(KArgs a)
{
    extern __shared__ __align__(16) char smem[];
    Lds lds = lds_setup(a.I, smem, REF == 2, REF == 2);
    const int e = blockIdx.x * PBB_EXPLORE_WAVES + (threadIdx.x >> 6);
    if (e >= a.E) return;
    explore_loop<M, REF, __HIP_MEMORY_SCOPE_AGENT>(a, lds, e);
}

// The team launch: explorers 0..a.E-1 as the wavefronts of one workgroup
// (up to team_waves_max(REF)), sharing work with workgroup-scope ordering only.
// A fresh search starts here: a small tree (a few thousand nodes) is solved
// by the team alone, without the cross-XCD writebacks and invalidates that
// every hand-over costs in a device-wide launch; a larger one stops at the
// node budget and the device-wide launch resumes from the saved intervals.
template <int M, int REF>
__global__ __launch_bounds__(64 * team_waves_max(REF)) void k_team(KArgs a)
{
    extern __shared__ __align__(16) char smem[];
    Lds lds = lds_setup(a.I, smem, REF == 2, REF == 2);
    const int e = threadIdx.x >> 6;
    if (e >= a.E) return;
    explore_loop<M, REF, __HIP_MEMORY_SCOPE_WORKGROUP>(a, lds, e);
}

template <int M>
__global__ __launch_bounds__(64) void k_bound_root(KArgs a, RootPerm root_perm)
{
    extern __shared__ __align__(16) char smem[];
    Lds lds = lds_setup(a.I, smem, false, false);
    Explorer<M, 0> X(a, lds, 0, engine_inst(a), a.ctl, a.I.n, a.I.mp);
    X.local_best = a.initial_ub;
    X.rowv = X.lane < X.n ? root_perm.p[X.lane] : 0;
    X.reset_stacks();
    X.bound_row(true);
    int d0 = rdlane(X.dirv, 0);
    if (X.lane < X.n) a.root_row_out[X.lane] = X.rowv;
    if (X.lane == 0) a.root_dir_out[0] = d0;
}

// ------------------------------------------------- batched bound_abstract
// Load an explicit partial schedule (perm, l1, l2) into an explorer's
// stacks: F <- perm[0..l1] (scheduleFront), B <- perm[l2..n-1] in reverse
// (scheduleBack), remain <- the middle.  Then the explorer's own routines
// produce the bounds.  Requires (l1+1) + (n-l2) < n.
template <int M>
__device__ __forceinline__ void load_partial(Explorer<M, 2> &X, const int *perm, int l1, int l2)
{
    X.reset_stacks();
    for (int i = 0; i <= l1; i++) X.push_job(uni(perm[i]), 0);
    for (int i = X.n - 1; i >= l2; i--) X.push_job(uni(perm[i]), 1);
}

// makespan of perm with perm[mid] the only job between the stacks
// (front = perm[0..mid-1], back = perm[mid+1..n-1])
template <int M>
__device__ __forceinline__ int makespan_mid(Explorer<M, 2> &X, const int *perm, int mid)
{
    load_partial<M>(X, perm, mid - 1, mid + 1);
    return X.makespan_with(uni(perm[mid]));
}

template <int M>
__global__ __launch_bounds__(256) void k_bounds(KArgs a, int which, int count, const int *perms,
                                                const int *l1s, const int *l2s, int arg, int *o1, int *o2)
{
    extern __shared__ __align__(16) char smem[];
    Lds lds = lds_setup(a.I, smem, which == 2, true);
    const int w = threadIdx.x >> 6;
    const int s = blockIdx.x * 4 + w;
    if (s >= count) return;
    Explorer<M, 2> X(a, lds, 0, engine_inst(a), a.ctl, a.I.n, a.I.mp);
    const int n = X.n, lane = X.lane;
    const int *perm = perms + (size_t)s * n;
    // inputs must be permutations of 0..n-1 with -1 <= l1 < l2 <= n;
    // anything else yields -1 outputs instead of out-of-range table reads
    const int pv = lane < n ? perm[lane] : 0;
    bool bad = ballot(lane < n && (pv < 0 || pv >= n)) != 0ull;
    int l1 = -1, l2 = n;
    if (which != 3) {
        l1 = uni(l1s[s]);
        l2 = uni(l2s[s]);
        bad = bad || l1 < -1 || l2 > n || l1 >= l2;
    }
    if (!bad) {
        if (lane < n) lds.scr[lane] = 0;
        __builtin_amdgcn_wave_barrier();
        if (lane < n) atomicAdd(&lds.scr[pv], 1);
        __builtin_amdgcn_wave_barrier();
        bad = ballot(lane < n && lds.scr[lane] != 1) != 0ull;  // repeated job ids
        __builtin_amdgcn_wave_barrier();
    }
    if (bad) {
        if (which == 0) {
            if (lane < n) o1[(size_t)s * n + lane] = -1;
            if (arg && lane < n) o2[(size_t)s * n + lane] = -1;
        } else if (lane == 0) {
            o1[s] = -1;
        }
        return;
    }
    if (which == 3) {
        // bound_fsp_weak::evalSolution
        int v = makespan_mid<M>(X, perm, n - 1);
        if (lane == 0) o1[s] = v;
        return;
    }
    const int k = l2 - l1 - 1;
    if (which == 0 && k == 0) {
        if (lane < n) {
            o1[(size_t)s * n + lane] = 0;
            if (arg) o2[(size_t)s * n + lane] = 0;
        }
        return;
    }
    if ((which == 1 && k == 0) || (which == 2 && k <= 1)) {
        // nothing (LB1') or at most one job (LB2, bound_fsp_strong.cpp:543-546)
        // unscheduled: both reduce to the makespan of the completed permutation
        int v = makespan_mid<M>(X, perm, k == 1 ? l1 + 1 : (l1 >= 0 ? l1 : 0));
        if (lane == 0) o1[s] = v;
        return;
    }
    load_partial<M>(X, perm, l1, l2);
    const bool valid = lane < k;
    const int job = valid ? perm[l1 + 1 + lane] : 0;
    if (which == 0) {
        // bound_fsp_weak::boundChildren (bound_fsp_weak.cpp:193-221)
        X.rowv = job;
        X.L = n - k;
        int cb, ce;
        X.lb1_children(k, cb, ce);
        int *inv = lds.scr;
        if (lane < n) inv[pv] = lane;
        __builtin_amdgcn_wave_barrier();
        int ci = lane < n ? inv[lane] - (l1 + 1) : -1;
        bool child = ci >= 0 && ci < k;
        int vF = shfl(cb, child ? ci : 0);
        int vB = shfl(ce, child ? ci : 0);
        if (lane < n) {
            o1[(size_t)s * n + lane] = child ? vF : 0;
            if (arg) o2[(size_t)s * n + lane] = child ? vB : 0;
        }
    } else if (which == 1) {
        // bound_fsp_weak::bornes_calculer (bound_fsp_weak.cpp:288-311)
        int t0 = 0, lb = 0;
#pragma unroll
        for (int q = 0; q < M; q++) {
            int f = X.fval(q), b = X.bval(q), r = X.rval(q);
            if (q == 0) {
                t0 = f + r;
                lb = t0 + b;
            } else {
                int t1 = max(f + r, t0);
                lb = max(lb, t1 + b);
                t0 = t1;
            }
        }
        if (lane == 0) o1[s] = lb;
    } else {
        // bound_fsp_strong::bornes_calculer (bound_fsp_strong.cpp:540-560)
        uint64_t um = X.row_jobs(valid, job);
        int *fs = lds.scr, *bs = lds.scr + 32;
        if (lane < M) {
            fs[lane] = X.frow(X.nf)[lane];
            bs[lane] = X.brow(X.nb)[lane];
        }
        __builtin_amdgcn_wave_barrier();
        int pinfo[Explorer<M, 2>::PCH];
        pair_lanes<Explorer<M, 2>::PCH>(a.I.pair_ab, a.I.npairs, a.I.m, arg, pinfo);
        int v = lb2_pairs_any<Explorer<M, 2>::PCH>(pinfo, lds.jtl, lds.jtabL, lds.jtabG, a.I.n, fs, bs, um);
        if (lane == 0) o1[s] = v;
    }
}

// ---------------------------------------------------------- host launchers
template <int M>
static const void *explore_kernel(const KArgs &a)
{
    if (a.bound_mode == 0) return (const void *)k_explore<M, 0>;
    return a.refine == 1 ? (const void *)k_explore<M, 2> : (const void *)k_explore<M, 1>;
}

template <int M>
hipError_t launch_explore_m(const KArgs &a, int blocks, hipStream_t s)
{
    const bool lb2 = a.bound_mode != 0 && a.refine == 1;
    const size_t lds = explore_lds_bytes(a.I, lb2);
    if (const hipError_t err =
            hipFuncSetAttribute(explore_kernel<M>(a), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds))
        return err;
    if (a.bound_mode == 0)
        hipLaunchKernelGGL((k_explore<M, 0>), dim3(blocks), dim3(64 * PBB_EXPLORE_WAVES), lds, s, a);
    else if (a.refine == 1)
        hipLaunchKernelGGL((k_explore<M, 2>), dim3(blocks), dim3(64 * PBB_EXPLORE_WAVES), lds, s, a);
    else
        hipLaunchKernelGGL((k_explore<M, 1>), dim3(blocks), dim3(64 * PBB_EXPLORE_WAVES), lds, s, a);
    return hipGetLastError();
}

template <int M>
hipError_t launch_team_m(const KArgs &a, hipStream_t s)
{
    const bool lb2 = a.bound_mode != 0 && a.refine == 1;
    const size_t lds = team_lds_bytes(a.I, lb2, a.E);
    const void *k = a.bound_mode == 0 ? (const void *)k_team<M, 0>
                    : a.refine == 1   ? (const void *)k_team<M, 2>
                                      : (const void *)k_team<M, 1>;
    if (const hipError_t err = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds))
        return err;
    if (a.bound_mode == 0)
        hipLaunchKernelGGL((k_team<M, 0>), dim3(1), dim3(64 * a.E), lds, s, a);
    else if (a.refine == 1)
        hipLaunchKernelGGL((k_team<M, 2>), dim3(1), dim3(64 * a.E), lds, s, a);
    else
        hipLaunchKernelGGL((k_team<M, 1>), dim3(1), dim3(64 * a.E), lds, s, a);
    return hipGetLastError();
}

template <int M>
hipError_t occupancy_m(const KArgs &a, int *blocks_per_cu)
{
    const bool lb2 = a.bound_mode != 0 && a.refine == 1;
    const size_t lds = explore_lds_bytes(a.I, lb2);
    if (lds > 160 * 1024) {
        *blocks_per_cu = 0;
        return hipSuccess;
    }
    const size_t q = residency_lds_bytes(lds);
    if (const hipError_t err =
            hipFuncSetAttribute(explore_kernel<M>(a), hipFuncAttributeMaxDynamicSharedMemorySize, (int)q))
        return err;
    if (a.bound_mode == 0)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_explore<M, 0>, 64 * PBB_EXPLORE_WAVES, q);
    if (a.refine == 1)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_explore<M, 2>, 64 * PBB_EXPLORE_WAVES, q);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_explore<M, 1>, 64 * PBB_EXPLORE_WAVES, q);
}

template <int M>
hipError_t launch_root_m(const KArgs &a, const RootPerm &root_perm, hipStream_t s)
{
    const size_t lds = root_lds_bytes(a.I);
    if (const hipError_t err =
            hipFuncSetAttribute((const void *)k_bound_root<M>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds))
        return err;
    hipLaunchKernelGGL(k_bound_root<M>, dim3(1), dim3(64), lds, s, a, root_perm);
    return hipGetLastError();
}

template <int M>
hipError_t launch_bounds_m(int which, const KArgs &a, int count, const int *perms, const int *l1,
                           const int *l2, int arg, int *o1, int *o2, hipStream_t s)
{
    const bool lb2 = which == 2 && jtab_in_lds(a.I.n, a.I.npairs);
    const size_t lds = lds_table_bytes(a.I.n, a.I.m, a.I.npairs, lb2) + 4 * lds_wave_bytes(a.I.n, a.I.m, true);
    if (const hipError_t err =
            hipFuncSetAttribute((const void *)k_bounds<M>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds))
        return err;
    hipLaunchKernelGGL(k_bounds<M>, dim3((count + 3) / 4), dim3(256), lds, s, a, which, count, perms, l1, l2,
                       arg, o1, o2);
    return hipGetLastError();
}

#ifdef PBB_M
template hipError_t launch_explore_m<PBB_M>(const KArgs &, int, hipStream_t);
template hipError_t launch_team_m<PBB_M>(const KArgs &, hipStream_t);
template hipError_t occupancy_m<PBB_M>(const KArgs &, int *);
template hipError_t launch_root_m<PBB_M>(const KArgs &, const RootPerm &, hipStream_t);
template hipError_t launch_bounds_m<PBB_M>(int, const KArgs &, int, const int *, const int *, const int *,
                                          int, int *, int *, hipStream_t);
#endif

}  // namespace pbb
