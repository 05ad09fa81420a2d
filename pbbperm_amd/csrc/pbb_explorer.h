// pbb_explorer.h -- the explorer: one IVM branch-and-bound search per
// 64-lane wavefront (navigation, bounding, branching, pruning, interval
// sharing), shared by the one-instance engine kernels (pbb_kernels.hip) and
// the multi-instance pool kernel (pbb_pool.hip).  Device code only; every
// function is force-inlined into the kernels that use it.
#pragma once

#include <utility>

#include "pbb_device.h"
#include "pbb_kernels.h"

namespace pbb {

// ------------------------------------------------------------------ atomics
// Every helper takes the memory scope of the explorers that share the
// words: the agent for a launch over the device, the workgroup for a team
// of explorers in one workgroup (whose release/acquire then need no L2
// writeback or invalidate).
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ int ld_acq(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, S);
}
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ int ld_rlx(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, S);
}
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ unsigned ld_rlx(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, S);
}
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ unsigned long long ld_rlx(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, S);
}
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ void st_rel(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, S);
}
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ void st_rlx(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, S);
}

template <int S>
__device__ __forceinline__ void fence_acq()
{
    if constexpr (S == __HIP_MEMORY_SCOPE_WORKGROUP) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// lane-0 read-modify-writes, result broadcast to the wavefront
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ int w_add(int *p, int v)
{
    int r = 0;
    if (lane_id() == 0) r = __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, S);
    return rdlane(r, 0);
}
// Relaxed counters: on this multi-XCD device an agent-scope release writes
// back the wave's L2 and an acquire invalidates it, so the idle-ring tickets
// and the work counter, whose values carry no other data, take no fence.
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ int w_add_rlx(int *p, int v)
{
    int r = 0;
    if (lane_id() == 0) r = __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, S);
    return rdlane(r, 0);
}
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ unsigned w_addu_rlx(unsigned *p, unsigned v)
{
    unsigned r = 0;
    if (lane_id() == 0) r = __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, S);
    return (unsigned)rdlane((int)r, 0);
}
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ unsigned long long ld_rlx64(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, S);
}
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ void st_rlx64(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, S);
}
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ unsigned long long w_add64(unsigned long long *p, unsigned long long v)
{
    unsigned long long r = 0;
    if (lane_id() == 0) r = __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, S);
    unsigned lo = (unsigned)rdlane((int)(unsigned)r, 0);
    unsigned hi = (unsigned)rdlane((int)(unsigned)(r >> 32), 0);
    return ((unsigned long long)hi << 32) | lo;
}
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ bool w_cas(unsigned *p, unsigned expect, unsigned desired)
{
    int ok = 0;
    if (lane_id() == 0) {
        unsigned e = expect;
        ok = __hip_atomic_compare_exchange_strong(p, &e, desired, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                                  S);
    }
    return rdlane(ok, 0) != 0;
}
// Relaxed claim: no agent-scope release/acquire, so a failed attempt costs
// one atomic round trip and no L2 writeback or invalidate.
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ bool w_cas_rlx(unsigned *p, unsigned expect, unsigned desired)
{
    int ok = 0;
    if (lane_id() == 0) {
        unsigned e = expect;
        ok = __hip_atomic_compare_exchange_strong(p, &e, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                  S);
    }
    return rdlane(ok, 0) != 0;
}
template <int S = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ void w_min64(unsigned long long *p, unsigned long long v)
{
    if (lane_id() == 0) __hip_atomic_fetch_min(p, v, __ATOMIC_ACQ_REL, S);
}

// ------------------------------------------------------ lane scans (DPP)
// Inclusive prefix scans over lanes 0..31 (row_shr 1,2,4,8 inside each
// 16-lane row, then row_bcast:15 into row 1).  Lanes only combine with lanes
// to their left, so values in lanes >= the vector length are never read.
__device__ __forceinline__ int scan_add32(int x)
{
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);
    return x;
}
__device__ __forceinline__ int scan_max32(int x)
{
    const int lo = (int)0x80000000;
    x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x111, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x112, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x114, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x118, 0xF, 0xF, false));
    x = max(x, __builtin_amdgcn_update_dpp(lo, x, 0x142, 0xA, 0xF, false));
    return x;
}

// Max-plus chain of one job over the machines, lane q = machine q:
// c_q = max(base_q, c_{q-1}) + p_q  ==  P_q + max_{i<=q}(base_i - P_{i-1})
// with P the prefix sum of p (the recurrence of bound_fsp_weak.cpp:76-82).
__device__ __forceinline__ int chain_scan(int base, int p)
{
    const int P = scan_add32(p);
    return P + scan_max32(base - (P - p));
}

// ------------------------------------------------------------ LDS layout
struct Lds {
    int *pm;        // [m][PM_STRIDE] machine-major processing times
    int *pmR;       // the same rows, machines reversed: pmR[q] = pm[m-1-q]
    int ps;         // row stride of pm (PM_STRIDE)
    int *minA, *minD, *tot;
    int4 *jtabL;    // Johnson table staged in LDS (valid when jtl)
    const int4 *jtabG;  // the same table in global memory
    bool jtl;
    int8_t *mat;    // this wave's IVM matrix, row l at mat + lds_mat_row(l, n)
    int *stk;       // this wave's stacks: F entry s at row s, B entry s at row n-s ((n+1) x m)
    int2 *U;        // this wave's LB1 operand block [2][m]
    int *scr;       // this wave's scratch (64 ints; LB2 variant only)
    int2 *sink;     // this wave's 64 int2 where the lanes past a row store: the
                    // row stores of the node loop then need no exec-mask branch
                    // (2066 against 2100-2104 M nodes/s at 50x20)
};

__device__ __forceinline__ Lds lds_setup(const DevInst &I, char *smem, bool load_jtab, bool scratch)
{
    Lds L;
    const int n = I.n, mp = I.mp, m = I.m;
    L.ps = PM_STRIDE;
    L.pm = (int *)smem;
    L.pmR = L.pm + m * PM_STRIDE;
    L.minA = L.pmR + m * PM_STRIDE;
    L.minD = L.minA + mp;
    L.tot = L.minD + mp;
    size_t off = ((size_t)2 * m * PM_STRIDE + 3 * (size_t)mp) * 4;
    off = (off + 15) & ~(size_t)15;
    load_jtab = load_jtab && jtab_in_lds(n, I.npairs);
    L.jtabL = (int4 *)(smem + off);
    L.jtabG = I.jtab;
    L.jtl = load_jtab;
    char *wbase = smem + lds_table_bytes(n, m, I.npairs, load_jtab) +
                  (size_t)(threadIdx.x >> 6) * lds_wave_bytes(n, m, scratch);
    L.mat = (int8_t *)wbase;
    L.stk = (int *)(wbase + lds_mat_bytes(n));
    L.U = (int2 *)(wbase + lds_mat_bytes(n) + lds_stack_bytes(n, m));
    L.scr = scratch ? (int *)(wbase + lds_mat_bytes(n) + lds_stack_bytes(n, m) + (size_t)m * 16) : nullptr;
    L.sink = (int2 *)(wbase + lds_mat_bytes(n) + lds_stack_bytes(n, m) + (size_t)m * 16 + (scratch ? 64 * 4 : 0));
    for (int i = threadIdx.x; i < n * m; i += blockDim.x) {
        const int q = i / n, j = i - q * n;
        L.pm[q * PM_STRIDE + j] = I.ptm[i];
        L.pmR[(m - 1 - q) * PM_STRIDE + j] = I.ptm[i];
    }
    for (int i = threadIdx.x; i < mp; i += blockDim.x) {
        L.minA[i] = i < m ? I.min_arr[i] : 0;
        L.minD[i] = i < m ? I.min_dep[i] : 0;
        int s = 0;
        if (i < m)
            for (int j = 0; j < n; j++) s += I.ptm[i * n + j];
        L.tot[i] = s;
    }
    if (load_jtab)
        for (int i = threadIdx.x; i < I.npairs * n; i += blockDim.x) L.jtabL[i] = I.jtab[i];
    __syncthreads();
    return L;
}

// NEH (heuristic/src/fastNEH.cpp:151-243, with the intended partial
// makespan at every insertion position): one wavefront.  Jobs get stable
// ranks by decreasing total time; then job k is inserted into the first k
// at the first position of least partial makespan (lane = position, each
// lane evaluates its candidate sequence on M machines).  The first-minimum
// rule also settles the first two jobs as fastNEH.cpp:233-236 does (the
// original order only when strictly better).  pm: [M][ps] machine-major;
// seq: n ints of the wave's LDS, receives the order.  Returns the makespan.
template <int M>
__device__ __forceinline__ int neh_wave(const int *pm, int ps, int n, int *seq)
{
    const int lane = lane_id();
    int tot = 0;
    if (lane < n)
#pragma unroll
        for (int q = 0; q < M; q++) tot += pm[q * ps + lane];
    int rank = 0;
    for (int i = 0; i < n; i++) {
        const int ti = rdlane(tot, i);
        rank += (ti > tot || (ti == tot && i < lane)) ? 1 : 0;
    }
    if (lane < n) seq[rank] = lane;
    __builtin_amdgcn_wave_barrier();
    int cmin = 0;
    if (n == 1) {
        int c = 0;
#pragma unroll
        for (int q = 0; q < M; q++) c += pm[q * ps + seq[0]];
        return c;
    }
    for (int k = 1; k < n; k++) {
        const int job = seq[k];
        int cost = 0x7fffffff;
        if (lane <= k) {
            int cv[M];
#pragma unroll
            for (int q = 0; q < M; q++) cv[q] = 0;
            for (int i = 0; i <= k; i++) {
                const int j = i < lane ? seq[i] : (i == lane ? job : seq[i - 1]);
                cv[0] += pm[j];
#pragma unroll
                for (int q = 1; q < M; q++) cv[q] = max(cv[q - 1], cv[q]) + pm[q * ps + j];
            }
            cost = cv[M - 1];
        }
        cmin = wave_min(cost);
        const int pb = __ffsll((unsigned long long)ballot(lane <= k && cost == cmin)) - 1;
        const int v = lane < n ? seq[lane] : 0;
        const int vprev = shfl(v, lane > 0 ? lane - 1 : 0);
        __builtin_amdgcn_wave_barrier();
        if (lane == pb) seq[lane] = job;
        else if (lane > pb && lane <= k) seq[lane] = vprev;
        __builtin_amdgcn_wave_barrier();
    }
    return cmin;
}

// instance reference of the one-instance engine: everything from the
// kernel arguments
__device__ __forceinline__ InstRef engine_inst(const KArgs &a)
{
    InstRef r;
    r.ctl = a.ctl;
    r.ring = a.ring;
    r.ring_cap = a.ring_cap;
    r.ub0 = a.initial_ub;
    r.best = a.best;
    r.root_row = a.root_row;
    r.root_dir = a.root_dir;
    r.m = a.I.m;
    r.npairs = a.I.npairs;
    r.pairs_mode = a.pairs_mode;
    r.pair_ab = a.I.pair_ab;
    return r;
}

// ------------------------------------------------------------- explorer
// Stacks: F entry s (row s of stk) = completion times on every machine of
// the first s front jobs, F entry 0 = min_arr (the front of an empty prefix,
// bound_fsp_weak.cpp:67-70); B entry s (row n-s) = reverse completion times
// of the last s back jobs, B entry 0 = min_dep.  nf + nb < n, so the two
// never overlap.  A chain started from min_arr/min_dep equals one started
// from zero, since min_arr[q] never exceeds a single job's head (and min_dep
// its tail).
// REF: 0 LB1 only (boundingMode 0); 1 LB1 + LB1' refinement (boundingMode
// 1/2 as shipped); 2 LB1 + Johnson LB2 refinement.
template <int M, int REF, int SC = __HIP_MEMORY_SCOPE_AGENT>
struct Explorer {
    const KArgs &a;
    Lds lds;
    InstRef ir;   // the instance this explorer works on
    Ctl *run;     // stop / error words of the launch
    int n, mp, lane, e;
    int rows;     // row count of the HBM explorer state (>= n)
    int L;                            // current line
    int rowv, posv, endv, dirv, remv; // lane vectors (remv: lane q = remaining work on machine q)
    int nf, nb;
    int ftop;                         // lane q < M: top front entry f_q (frow(nf)), else 0
    static constexpr int PCH = (M * (M - 1) / 2 + PBB_WAVE - 1) / PBB_WAVE;  // pair chunks (M >= 2)
    int pinfo[PCH];                   // this lane's Johnson machine pairs (REF 2)
    int local_best;
    unsigned dec, leaves, sb, jb;     // deltas since the last flush
    unsigned nchk;                    // periodic checks so far (host-word phase)
    unsigned since;                   // nodes decomposed in this launch's own interval (ramp)
    bool end_dirty;                   // end moved since pos was last compared with it
    int hist;                         // lane d: nodes decomposed at depth d since the flush
    int jhist;                        // lane d: refinement bounds at depth d (REF != 0)
    unsigned long long t_end;

    __device__ __forceinline__ Explorer(const KArgs &args, const Lds &l, int eid, const InstRef &r, Ctl *runctl,
                                        int n_, int mp_)
        : a(args), lds(l), ir(r), run(runctl), n(n_), mp(mp_), lane(lane_id()), e(eid), rows(args.rows)
    {
        L = 0;
        rowv = posv = endv = dirv = remv = 0;
        nf = nb = 0;
        ftop = 0;
        if constexpr (REF == 2) pair_lanes<PCH>(r.pair_ab, r.npairs, r.m, r.pairs_mode, pinfo);
        dec = leaves = sb = jb = 0;
        since = 0;
        nchk = (unsigned)eid;  // staggers the host-word reads across explorers
        end_dirty = true;
        hist = jhist = 0;
        local_best = PBB_INF;
        t_end = args.deadline ? wall_clock64() + args.deadline : 0ull;
    }

    __device__ __forceinline__ bool gt_best(int v) const
    {
        return a.find_all ? v > local_best : v >= local_best;
    }

    __device__ __forceinline__ int *frow(int s) const { return lds.stk + s * M; }
    __device__ __forceinline__ int *brow(int s) const { return lds.stk + (n - s) * M; }
    // uniform stack values (LDS broadcast reads)
    __device__ __forceinline__ int fval(int q) const { return frow(nf)[q]; }
    __device__ __forceinline__ int bval(int q) const { return brow(nb)[q]; }
    __device__ __forceinline__ int rval(int q) const { return rdlane(remv, q); }

    // lane q: p[q][job].  Lanes past the machines read machine M-1 and drop
    // it: an unconditional load, no exec-mask branch around it (as for the
    // other per-machine and per-cell loads of the node loop; 2044-2046
    // against 2068-2070 M nodes/s at 50x20).
    __device__ __forceinline__ int pmach(int job) const
    {
        const int v = lds.pm[(lane < M ? lane : M - 1) * PM_STRIDE + job];
        return lane < M ? v : 0;
    }

    __device__ __forceinline__ void reset_stacks()
    {
        nf = nb = 0;
        ftop = lane < M ? lds.minA[lane] : 0;
        if (lane < M) {
            frow(0)[lane] = ftop;
            brow(0)[lane] = lds.minD[lane];
        }
        remv = lane < M ? lds.tot[lane] : 0;
        __builtin_amdgcn_wave_barrier();
    }

    // schedule job j in direction d: push a new F (d=0) or B (d=1) entry,
    // computed lane-parallel by chain_scan; remain -= p_j.
    __device__ __forceinline__ void push_job(int j, int d)
    {
        const int pk = pmach(j);
        if (d == 0) {
            int c = chain_scan(ftop, pk);
            *(lane < M ? frow(nf + 1) + lane : (int *)(lds.sink + lane)) = c;
            ftop = lane < M ? c : 0;
            nf++;
        } else {
            // back chain runs from machine M-1 down: lane q handles machine M-1-q
            const int qb = lane < M ? lane : M - 1;
            int base = brow(nb)[M - 1 - qb];
            int pr = lds.pmR[qb * PM_STRIDE + j];
            base = lane < M ? base : 0;
            pr = lane < M ? pr : 0;
            int c = chain_scan(base, pr);
            *(lane < M ? brow(nb + 1) + (M - 1 - lane) : (int *)(lds.sink + lane)) = c;
            nb++;
        }
        remv -= pk;
        __builtin_amdgcn_wave_barrier();
    }

    __device__ __forceinline__ void pop_job(int j, int d)
    {
        if (d == 0) {
            nf--;
            ftop = lane < M ? frow(nf)[lane] : 0;
        } else {
            nb--;
        }
        remv += pmach(j);
    }

    // ivm::beforeEnd (ivm.cpp:152-160): lexicographic pos <= end
    __device__ __forceinline__ bool before_end() const
    {
        const uint64_t in = lanes_below(n);
        uint64_t lt = in & ballot(posv < endv);
        uint64_t gt = in & ballot(posv > endv);
        uint64_t d = lt | gt;
        if (d == 0) return true;
        int f = __ffsll((unsigned long long)d) - 1;
        return (lt >> f) & 1ull;
    }

    // ------------------------------------------------------- LB1 children
    // bound_fsp_weak::addFrontAndBound / addBackAndBound (bound_fsp_weak.cpp:133-191)
    // for all children of the node, both directions in one pass: lane i < 32
    // bounds child base+i in the front direction, lane 32+i the same child in
    // the back direction.  The back recurrence run over machines M-1..0 has
    // the same form as the front one run over 0..M-1, so both halves execute
    // one loop:  t1 = max(x_q, t0); lb = max(lb, t1 + y_q); t0 = t1 + p_q,
    // with (x_q, y_q) = (front_q, remain_q + back_q) in the front half and
    // (back_{M-1-q}, front_{M-1-q} + remain_{M-1-q}) in the back half.
    // The per-wave operand block U holds the front half's pairs in machine
    // order, then the back half's.
    __device__ __forceinline__ void build_operands() const
    {
        const bool in = lane < M;
        const int f = ftop, b = brow(nb)[in ? lane : M - 1], r = remv;
        *(in ? lds.U + lane : lds.sink + lane) = make_int2(f, r + b);
        *(in ? lds.U + M + (M - 1 - lane) : lds.sink + lane) = make_int2(b, f + r);
        __builtin_amdgcn_wave_barrier();
    }

    // lane Q of every 16-lane row, to all lanes of the row (DPP row_newbcast,
    // folded into the instruction that uses it)
    template <int Q>
    __device__ __forceinline__ static int row_bcast_lane(int v)
    {
        return __builtin_amdgcn_update_dpp(0, v, 0x150 + (Q & 15), 0xF, 0xF, false);
    }

    // machine step Q of the fused pass; the step's operand pair is lane Q mod
    // 16 of the row's copy of U (A for Q < 16, A2 above)
    template <int Q>
    __device__ __forceinline__ static void lb1_step(const int2 &A, const int2 &A2, const int (&pj)[M], int &t0,
                                                    int &lb)
    {
        int x, y;
        if constexpr (Q < 16) {
            x = row_bcast_lane<Q>(A.x);
            y = row_bcast_lane<Q>(A.y);
        } else {
            x = row_bcast_lane<Q - 16>(A2.x);
            y = row_bcast_lane<Q - 16>(A2.y);
        }
        if constexpr (Q == 0) {
            t0 = x + pj[0];
            lb = x + y;
        } else {
            const int t1 = max(x, t0);
            lb = max(lb, t1 + y);
            t0 = t1 + pj[Q];
        }
    }

    template <int... Q>
    __device__ __forceinline__ static int lb1_steps(const int2 &A, const int2 &A2, const int (&pj)[M],
                                                    std::integer_sequence<int, Q...>)
    {
        int t0 = 0, lb = 0;
        (lb1_step<Q>(A, A2, pj, t0, lb), ...);
        return lb;
    }

    // one fused pass over children base..base+31; returns this lane's bound.
    // Each 16-lane row reads its half's operand pairs once (two vector LDS
    // reads, no bank conflicts) and every machine step takes its pair from
    // lane q of the row by DPP, instead of M broadcast reads of U.
    __device__ __forceinline__ int lb1_pass(int base, int k) const
    {
        const int half = lane >> 5;
        const int c = base + (lane & 31);
        const int job = base == 0 ? bcast_low_half(rowv) : bcast_high_half(rowv);
        const int jb = c < k ? job : 0;
        // machine-major rows: front lanes walk machines 0..M-1, back lanes M-1..0
        const int *pcol = (half ? lds.pmR : lds.pm) + jb;
        const int ql = lane & 15;
        const int2 *u = lds.U + half * M;
        const int2 A = u[min(ql, M - 1)];
        int2 A2 = make_int2(0, 0);
        if constexpr (M > 16) A2 = u[16 + min(ql, M - 17)];
        int pj[M];
#pragma unroll
        for (int q = 0; q < M; q++) pj[q] = pcol[q * PM_STRIDE];
        return lb1_steps(A, A2, pj, std::make_integer_sequence<int, M>{});
    }

    // row layout (lane = child): cb (front) and ce (back) for k <= 64 children
    __device__ __forceinline__ void lb1_children(int k, int &cb, int &ce) const
    {
        build_operands();
        int lb0 = lb1_pass(0, k);
        if (k <= 32) {
            cb = lb0;
            ce = bcast_high_half(lb0);
        } else {
            int lb1 = lb1_pass(32, k);
            int sh0 = bcast_high_half(lb0), sh1 = bcast_low_half(lb1);
            cb = lane < 32 ? lb0 : sh1;
            ce = lane < 32 ? sh0 : lb1;
        }
    }

    // chooseChildrenSet (ivm_bound.cpp:413-507): lanes beyond the row hold
    // the reference's zero entries and are excluded from every term.
    __device__ __forceinline__ int choose(bool valid, int cb, int ce) const
    {
        switch (a.branching) {
        case -3: return (L % 2) ? 1 : 0;
        case -2: return 1;
        case -1: return 0;
        case 1: {
            int s = wave_sum(valid ? cb - ce : 0);
            return s < 0 ? 1 : 0;
        }
        case 2: {
            int ub = a.single_node ? local_best : ir.ub0;
            const uint64_t vm = ballot(valid);
            int elim = __popcll(vm & ballot(cb >= ub)) - __popcll(vm & ballot(ce >= ub));
            int s = wave_sum(valid ? ((cb < ub ? cb : 0) - (ce < ub ? ce : 0)) : 0);
            if (elim > 0) return 0;
            if (elim < 0) return 1;
            return s < 0 ? 1 : 0;
        }
        default: {
            int ub = a.single_node ? local_best : ir.ub0;
            int mn = wave_min(valid ? min(cb > 0 ? cb : PBB_INF, ce > 0 ? ce : PBB_INF) : PBB_INF);
            const uint64_t vm = ballot(valid);
            int minc = __popcll(vm & ballot(cb == mn)) - __popcll(vm & ballot(ce == mn));
            int elimc = __popcll(vm & ballot(cb >= ub)) - __popcll(vm & ballot(ce >= ub));
            if (minc > 0) return 1;
            if (minc < 0) return 0;
            return elimc > 0 ? 0 : 1;
        }
        }
    }

    // ----------------------------------------------------------- bounding
    // Bound the children of the node at line L (row in rowv, unflagged),
    // choose the direction, refine, prune, reverse (ivm_bound.cpp:141-236,
    // :337-676).  root: boundRoot always uses weakBoundPrune.
    __device__ __forceinline__ void bound_row(bool root)
    {
        const int k = n - L;
        const bool valid = lane < k;
        const int job = valid ? rowv : 0;
        const bool both = a.branching > 0;
        int dir, c1, c2 = 0;
        bool refined = false;
        if constexpr (REF == 0) {
            int cb, ce;
            lb1_children(k, cb, ce);
            if (!both) ce = 0;  // costsEnd[WEAK] stays zero (boundChildren front only)
            sb += (unsigned)(both ? 2 * k : k);
            dir = choose(valid, cb, ce);
            c1 = dir ? ce : cb;
        } else {
            const int mode = root ? 0 : a.bound_mode;
            int cb, ce;
            lb1_children(k, cb, ce);
            if (mode == 1) {
                int csF = 0, csB = 0;
                refine_both(valid, job, csF, csB);
                dir = choose(valid, csF, csB);
                c1 = c2 = dir ? csB : csF;
                refined = true;
            } else {
                if (!both) ce = 0;
                sb += (unsigned)(both ? 2 * k : k);
                dir = choose(valid, cb, ce);
                c1 = dir ? ce : cb;
                if (mode == 2) {
                    c2 = refine_dir(dir, valid, job, c1);
                    refined = true;
                }
            }
        }
        bool pr = valid && (gt_best(c1) || (refined && gt_best(c2)));
        int v = pr ? flag_neg(job) : job;
        int prev = L > 0 ? rdlane(dirv, L - 1) : 0;
        if (prev != dir) v = shfl(v, k - 1 - lane);
        rowv = valid ? v : 0;
        dirv = wrlane(dirv, dir, L);
    }

    // ---------------------------------------------------------- refinement
    __device__ __forceinline__ uint64_t row_jobs(bool valid, int job) const
    {
        uint64_t bits = valid ? (1ull << job) : 0ull;
        unsigned lo = (unsigned)bits, hi = (unsigned)(bits >> 32);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            lo |= (unsigned)__shfl_xor((int)lo, o, PBB_WAVE);
            hi |= (unsigned)__shfl_xor((int)hi, o, PBB_WAVE);
        }
        return ((uint64_t)(unsigned)uni((int)hi) << 32) | (unsigned)uni((int)lo);
    }

    // The refinement bounds below work on one child per lane (lane = row
    // cell), both directions, with the child's chains computed per lane
    // (the refinement modes serve instances of up to 30 jobs; LB1 above is
    // the fused hot loop).
    // d = 0: front' = front + job; d = 1: back' = job + back.
    __device__ __forceinline__ void child_chains(int job, int (&chF)[M], int (&chB)[M]) const
    {
        const int *pj = lds.pm + job;  // machine q at pj[q * PM_STRIDE]
        constexpr int ps = PM_STRIDE;
        int t = 0;
#pragma unroll
        for (int q = 0; q < M; q++) {
            int f = fval(q);
            t = (q == 0 ? f : max(f, t)) + pj[q * ps];
            chF[q] = t;
        }
#pragma unroll
        for (int q = M - 1; q >= 0; q--) {
            int b = bval(q);
            t = (q == M - 1 ? b : max(b, t)) + pj[q * ps];
            chB[q] = t;
        }
    }

    // LB1' of this lane's child (bound_fsp_weak::bornes_calculer with the
    // child's front or back)
    __device__ __forceinline__ int lb1p_child(int d, int job, const int (&chF)[M], const int (&chB)[M]) const
    {
        const int *pj = lds.pm + job;  // machine q at pj[q * PM_STRIDE]
        constexpr int ps = PM_STRIDE;
        int t0 = 0, lb = 0;
#pragma unroll
        for (int q = 0; q < M; q++) {
            int f = d == 0 ? chF[q] : fval(q);
            int b = d == 1 ? chB[q] : bval(q);
            int r = rval(q) - pj[q * ps];
            if (q == 0) {
                t0 = f + r;
                lb = t0 + b;
            } else {
                int t1 = max(f + r, t0);
                lb = max(lb, t1 + b);
                t0 = t1;
            }
        }
        return lb;
    }

    // Johnson LB2 of child c (uniform), direction d
    // (bound_fsp_strong::bornes_calculer with the child's partial schedule).
    // The child's front (d = 0) or back (d = 1) completion times come from
    // one lane-parallel chain scan (lane = machine, as a stack push) and go
    // to the wave's LDS scratch, where the pair lanes read them.
    __device__ __forceinline__ int lb2_child(int c, int d, uint64_t rowmask, int k)
    {
        const int cj = flag_abs(rdlane(rowv, c));
        int *fs = lds.scr, *bs = lds.scr + 32;
        if (d == 0) {
            const int f = ftop;
            const int cf = chain_scan(f, pmach(cj));
            if (lane < M) {
                fs[lane] = cf;
                bs[lane] = brow(nb)[lane];
            }
        } else {
            // back chain over machines M-1..0: lane q handles machine M-1-q
            const int base = lane < M ? brow(nb)[M - 1 - lane] : 0;
            const int pr = lane < M ? lds.pmR[lane * PM_STRIDE + cj] : 0;
            const int cb = chain_scan(base, pr);
            if (lane < M) {
                fs[lane] = ftop;
                bs[M - 1 - lane] = cb;
            }
        }
        __builtin_amdgcn_wave_barrier();
        int r;
        if (k <= 2) {
            // l2 - l1 <= 2: makespan of the completed permutation (bound_fsp_strong.cpp:543-546)
            const int other = flag_abs(rdlane(rowv, 1 - c));
            const int fq = lane < M ? fs[lane] : 0;
            const int bq = lane < M ? bs[lane] : 0;
            const int cc = chain_scan(fq, pmach(other));
            r = wave_max(lane < M ? cc + bq : 0);
        } else {
            const uint64_t unsched = rowmask & ~(1ull << cj);
            r = lb2_pairs_any<PCH>(pinfo, lds.jtl, lds.jtabL, lds.jtabG, n, fs, bs, unsched);
        }
        __builtin_amdgcn_wave_barrier();
        return r;
    }

    // mixed mode (computeStrongBounds, ivm_bound.cpp:141-195): refine the
    // children not pruned by LB1 in the chosen direction; 0 for the others
    __device__ __forceinline__ int refine_dir(int d, bool valid, int job, int cw)
    {
        const int k = n - L;
        bool todo = valid && !gt_best(cw);
        uint64_t tmask = ballot(todo);
        jb += (unsigned)__popcll(tmask);
        jhist += lane == L ? __popcll(tmask) : 0;
        if constexpr (REF == 1) {
            int chF[M], chB[M];
            child_chains(job, chF, chB);
            int v = lb1p_child(d, job, chF, chB);
            return todo ? v : 0;
        }
        uint64_t rowmask = row_jobs(valid, job);
        int cs = 0;
        uint64_t m = tmask;
        while (m) {
            int c = __ffsll((unsigned long long)m) - 1;
            m &= m - 1;
            int v = lb2_child(c, d, rowmask, k);
            if (lane == c) cs = v;
        }
        return cs;
    }

    // strong-only mode (strongBoundPrune, ivm_bound.cpp:668-676)
    __device__ __forceinline__ void refine_both(bool valid, int job, int &csF, int &csB)
    {
        const int k = n - L;
        jb += 2u * (unsigned)k;
        jhist += lane == L ? 2 * k : 0;
        if constexpr (REF == 1) {
            int chF[M], chB[M];
            child_chains(job, chF, chB);
            csF = valid ? lb1p_child(0, job, chF, chB) : 0;
            csB = valid ? lb1p_child(1, job, chF, chB) : 0;
            return;
        }
        uint64_t rowmask = row_jobs(valid, job);
        csF = csB = 0;
        for (int c = 0; c < k; c++) {
            int v0 = lb2_child(c, 0, rowmask, k);
            int v1 = lb2_child(c, 1, rowmask, k);
            if (lane == c) {
                csF = v0;
                csB = v1;
            }
        }
    }

    // -------------------------------------------------------------- leaf
    // ivm_bound::boundLeaf (ivm_bound.cpp:238-271): makespan of front + the
    // last row job + back; the incumbent only moves on strict improvement.
    __device__ __forceinline__ int makespan_with(int r) const
    {
        int c = chain_scan(ftop, pmach(r));
        int b = brow(nb)[lane < M ? lane : M - 1];
        b = lane < M ? b : 0;
        return wave_max(lane < M ? c + b : 0);
    }

    __device__ __forceinline__ void leaf()
    {
        leaves++;
        const int r = rdlane(rowv, 0);
        const int cost = makespan_with(r);
        if (!gt_best(cost)) {
            if (cost < local_best) {
                write_perm(r);
                w_min64<SC>(ir.best, ((unsigned long long)(unsigned)cost << 32) | (unsigned)e);
            }
            local_best = cost;
        }
        rowv = wrlane(rowv, flag_neg(r), 0);
    }

    // decode the leaf schedule (ivm_bound::prepareSchedule) into best_perm[e]
    __device__ __forceinline__ void write_perm(int mid)
    {
        int jl = 0;
        if (lane < L) jl = flag_abs((int)lds.mat[lds_mat_row(lane, n) + posv]);
        uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        uint64_t fr = ballot(lane < L && dirv == 0);
        uint64_t bk = ballot(lane < L && dirv != 0);
        int *dst = a.best_perm + (size_t)e * 64;
        if (lane < L) {
            int p = dirv == 0 ? __popcll(fr & below) : n - 1 - __popcll(bk & below);
            dst[p] = jl;
        }
        if (lane == 0) dst[__popcll(fr)] = mid;
    }

    // ------------------------------------------------------- navigation
    __device__ __forceinline__ void store_row(int line) const
    {
        *(lane < n - line ? lds.mat + lds_mat_row(line, n) + lane : (int8_t *)(lds.sink + lane)) = (int8_t)rowv;
    }

    __device__ __forceinline__ void go_up()
    {
        if (L > 0) {
            posv = wrlane(posv, 0, L);
            L--;
            const int v = (int)lds.mat[lds_mat_row(L, n) + min(lane, n - 1 - L)];
            rowv = lane < n - L ? v : 0;
            const int p = rdlane(posv, L);
            pop_job(flag_abs(rdlane(rowv, p)), rdlane(dirv, L));
            if (lane == p) rowv = flag_neg(rowv);
        } else {
            posv = wrlane(posv, rdlane(posv, 0) + 1, 0);
        }
    }

    // branch on cell c of line L: generate line L+1 (ivm::goDown/generateLine)
    __device__ __forceinline__ void go_down(int c, bool explore)
    {
        store_row(L);
        push_job(flag_abs(rdlane(rowv, c)), rdlane(dirv, L));
        // drop cell c: lanes >= c take their right neighbour (DPP wave_shl:1,
        // lane i reads lane i+1; lane 63 reads 0)
        const int nxt = __builtin_amdgcn_update_dpp(0, rowv, 0x130, 0xF, 0xF, false);
        int v = lane >= c ? nxt : rowv;
        L++;
        rowv = lane < n - L ? flag_abs(v) : 0;
        if (explore) {
            posv = wrlane(posv, 0, L);
            dirv = wrlane(dirv, 0, L);
        }
    }

    // sequentialbb::unfold (sequentialbb.cpp:206-264): bounded, uncounted
    __device__ __forceinline__ void unfold()
    {
        while (L < n - 2) {
            int p = rdlane(posv, L);
            if (rdlane(rowv, p) < 0) {
                if (lane > L) posv = 0;
                break;
            }
            go_down(p, false);
            bound_row(false);
        }
    }

    // One call of sequentialbb::next (sequentialbb.cpp:128-163): returns 1
    // when an interior node was decomposed and bounded, 0 when the interval
    // is exhausted.  The reference tests beforeEnd at every iteration; pos
    // only grows in goRight and at the root line's increment (goUp resets a
    // deeper position to 0, goDown opens a line at 0), and end only shrinks
    // when a donation trims it, so the test runs exactly there and gives the
    // same answers.  goRight then goes down or up in the same iteration.
    __device__ __forceinline__ int step()
    {
        if (end_dirty) {
            end_dirty = false;
            if (!before_end()) return 0;
        }
        while (true) {
            const int p = rdlane(posv, L);
            int c = p;
            if (p < n - L) {
                // first live cell at or after p (ivm::goRight over pruned / visited cells)
                const uint64_t live = lanes_below(n - L) & ballot(rowv >= 0);
                const uint64_t ahead = live & (~0ull << p);
                c = ahead ? __ffsll((unsigned long long)ahead) - 1 : n - L;
                if (c != p) {
                    posv = wrlane(posv, c, L);
                    if (!before_end()) return 0;
                }
            }
            if (c >= n - L) {
                const bool root_line = L == 0;
                go_up();
                if (root_line && !before_end()) return 0;
                continue;
            }
            dec++;
            since += since < a.ramp ? 1u : 0u;  // saturates at the ramp
            go_down(c, true);
            hist += (lane == L);
            if (L == n - 1) {
                leaf();
                continue;
            }
            bound_row(false);
            return 1;
        }
    }

    // ------------------------------------------------------ state I/O
    __device__ __forceinline__ void rebuild_stacks()
    {
        reset_stacks();
        for (int l = 0; l < L; l++) {
            int p = rdlane(posv, l);
            int j = uni(flag_abs((int)lds.mat[lds_mat_row(l, n) + p]));
            push_job(j, rdlane(dirv, l));
        }
    }

    __device__ __forceinline__ void load_state()
    {
        L = uni(a.line[e]);
        const int8_t *gm = a.mat + (size_t)e * rows * 64;
        for (int r = 0; r < L; r++)
            if (lane < n - r) lds.mat[lds_mat_row(r, n) + lane] = gm[r * 64 + lane];
        rowv = lane < n - L ? (int)gm[L * 64 + lane] : 0;
        posv = lane < n ? (int)a.pos[(size_t)e * 64 + lane] : 0;
        endv = lane < n ? (int)a.end[(size_t)e * 64 + lane] : 0;
        dirv = lane < n ? (int)a.dir[(size_t)e * 64 + lane] : 0;
        end_dirty = true;
        since = a.ramp;  // a resumed or donated interval is shared at once
        __builtin_amdgcn_wave_barrier();
        rebuild_stacks();
    }

    __device__ __forceinline__ void save_state(int st)
    {
        int8_t *gm = a.mat + (size_t)e * rows * 64;
        for (int r = 0; r < L; r++)
            if (lane < n - r) gm[r * 64 + lane] = lds.mat[lds_mat_row(r, n) + lane];
        if (lane < n) gm[L * 64 + lane] = (int8_t)rowv;
        if (lane < n) {
            a.pos[(size_t)e * 64 + lane] = (int8_t)posv;
            a.end[(size_t)e * 64 + lane] = (int8_t)endv;
            a.dir[(size_t)e * 64 + lane] = (int8_t)dirv;
        }
        if (lane == 0) a.line[e] = L;
        if (lane == 0) st_rlx<SC>(a.state + e, st);
    }

    // fresh interval: row 0 = bounded root row, pos/end from HBM
    __device__ __forceinline__ bool start_unfold()
    {
        L = 0;
        rowv = lane < n ? ir.root_row[lane] : 0;
        posv = lane < n ? (int)a.pos[(size_t)e * 64 + lane] : 0;
        endv = lane < n ? (int)a.end[(size_t)e * 64 + lane] : 0;
        dirv = lane == 0 ? ir.root_dir[0] : 0;
        end_dirty = true;
        since = 0;
        reset_stacks();
        if (!before_end()) return false;
        unfold();
        return true;
    }

    // the whole search space [0...0, (n-1)...0] from the bounded root row
    // already in rowv / dirv lane 0 (sequentialbb::initFullInterval +
    // initAtInterval, sequentialbb.cpp:66-108)
    __device__ __forceinline__ bool start_root()
    {
        L = 0;
        posv = 0;
        endv = lane < n ? n - 1 - lane : 0;
        dirv = lane == 0 ? dirv : 0;
        end_dirty = true;
        since = 0;
        reset_stacks();
        if (!before_end()) return false;
        unfold();
        return true;
    }

    // --------------------------------------------------- work sharing
    // idle-wait pause: 2^k x 512 cycles, k growing to 4 while nothing
    // happens, so thousands of idle wavefronts do not flood the memory
    // system with polls
    __device__ __forceinline__ static void idle_pause(int &nap)
    {
        for (int i = 0; i < nap; i++) __builtin_amdgcn_s_sleep(8);
        if (nap < 16) nap <<= 1;
    }

    __device__ __forceinline__ void fail(int code)
    {
        if (lane == 0) {
            st_rlx<SC>(&run->err, code);
            st_rlx<SC>(&run->stop, 1);
        }
    }

    // bbthread::shareWork + ivm::cuttingPosition (bbthread.cpp:39-72,
    // ivm.cpp:188-237): split at the first line l whose pos differs from end
    // (l <= L, l < n - margin, the size_d - margin rule of
    // gpu/include/loadbalance.cuh:90); the thief gets the explorable cells after the
    // victim's kept half, rows 0..l and the victim's end; no re-bounding.
    // A line whose cells after pos are all pruned or visited holds no work:
    // its end is trimmed to pos (deeper ends become the maximum, so the
    // victim's interval covers exactly the same live nodes) and the split
    // moves down to the next line.  Node counts are unchanged by either step.
    // An explorer that starts a launch on a fresh interval (the root
    // interval, imported intervals) decomposes `ramp` nodes before it hands
    // work on: a small tree then runs depth-first in one explorer, finding
    // the incumbents the reference's order finds, and a large one fans out
    // after that many nodes (woken and resumed explorers share at once).
    // h, t: idle-ring head/tail loaded before the node's work (prefetch).
    // Returns whether idle explorers were waiting (the ring was not empty).
    __device__ __forceinline__ bool try_donate(unsigned h, unsigned t)
    {
        h = (unsigned)uni((int)h);
        t = (unsigned)uni((int)t);
        if (h == t) return false;
        const int lim = min(L, n - 1 - a.steal_margin);  // last line a split may use
        if (lim < 0) return true;                         // margin >= n: never split
        uint64_t ne = lanes_below(n) & ballot(posv != endv);
        uint64_t limmask = lim >= 64 ? ~0ull : ((1ull << lim) - 1ull);
        uint64_t cand = ne & limmask;
        const int first = cand ? __ffsll((unsigned long long)cand) - 1 : lim;
        int l = first, pl = 0, el = 0, rowl = -1;
        uint64_t expl = 0ull;
        for (;; l++) {
            if (l > lim) return true;
            pl = rdlane(posv, l);
            el = (l == first) ? rdlane(endv, l) : n - 1 - l;
            if (pl < el) {
                rowl = (l == L) ? rowv : (lane < n - l ? (int)lds.mat[lds_mat_row(l, n) + lane] : -1);
                expl = (lanes_below(el + 1) & ~lanes_below(pl + 1)) & ballot(rowl >= 0);
                if (expl) break;
            }
        }
        // the victim's end once lines first..l-1 are trimmed
        int endE = endv;
        if (l > first) {
            if (lane >= first && lane < l) endE = posv;
            else if (lane >= l && lane < n) endE = n - 1 - lane;
        }
        int nexp = __popcll(expl);
        // Every busy explorer that saw the ring non-empty races for ticket h.
        // The claim orders nothing (the donor has published nothing yet, and
        // the acquire load of the cell's sequence below orders the read of
        // the thief's id), so it is relaxed: with thousands of contenders an
        // acq_rel attempt, i.e. an L2 writeback + invalidate per loser,
        // dominated the ramp-up of every solve.
        if (!w_cas_rlx<SC>(&ir.ctl->idle_head, h, h + 1)) return true;
        // claim the idle explorer registered at ticket h (bounded MPMC ring
        // cell: the ticket and the explorer id share one word, so reading
        // the id needs no acquire)
        const unsigned slot = h & (unsigned)(ir.ring_cap - 1);
        bool ok = false;
        unsigned long long cell = 0ull;
        for (int spin = 0; spin < PBB_SPIN_LIMIT; spin++) {
            cell = uni64(ld_rlx64<SC>(ir.ring + slot));
            if ((unsigned)(cell >> 32) == h + 1u) {
                ok = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (!ok) {
            fail(PBB_DEV_ERR_RING);
            return true;
        }
        const int thief = (int)(unsigned)cell - 1;
        // the cell waits for the ticket of the next lap
        if (lane == 0) st_rlx64<SC>(ir.ring + slot, (unsigned long long)(h + (unsigned)ir.ring_cap) << 32);
        if (thief < 0 || thief >= a.E) {
            fail(PBB_DEV_ERR_RING);
            return true;
        }
        int keep = nexp / 2;
        int cut;
        if (keep == 0) {
            cut = pl + 1;
        } else {
            uint64_t mm = expl;
            for (int i = 1; i < keep; i++) mm &= mm - 1;
            cut = __ffsll((unsigned long long)mm);  // index of the keep-th cell + 1
        }
        int8_t *gm = a.mat + (size_t)thief * rows * 64;
        for (int r = 0; r < l; r++)
            if (lane < n - r) gm[r * 64 + lane] = lds.mat[lds_mat_row(r, n) + lane];
        if (lane < n) gm[l * 64 + lane] = (int8_t)rowl;
        if (lane < n) {
            a.pos[(size_t)thief * 64 + lane] = (int8_t)(lane < l ? posv : (lane == l ? cut : 0));
            a.end[(size_t)thief * 64 + lane] = (int8_t)endE;
            a.dir[(size_t)thief * 64 + lane] = (int8_t)(lane <= l ? dirv : 0);
        }
        if (lane == 0) a.line[thief] = l;
        // the victim keeps [pos, cut-1] at line l and everything below
        endv = endE;
        if (lane == l) endv = cut - 1;
        if (lane > l && lane < n) endv = n - 1 - lane;
        end_dirty = true;
        // The thief's count is added before the release below publishes its
        // interval, so its own decrement (after its acquire) comes later in
        // the counter's order: relaxed suffices.
        w_add_rlx<SC>(&ir.ctl->work_count, 1);
        if (lane == 0) st_rel<SC>(a.state + thief, ST_ACTIVE);
        return true;
    }

    // enter the idle ring (cell sequence protocol of a bounded MPMC queue:
    // ticket t owns cell t mod cap once the cell shows t; storing t + 1 with
    // this explorer's id hands it to the consumer of ticket t).  All relaxed:
    // the cell word carries everything the consumer reads.
    __device__ __forceinline__ void register_idle()
    {
        if (lane == 0) st_rlx<SC>(a.state + e, ST_IDLE);
        unsigned t = w_addu_rlx<SC>(&ir.ctl->idle_tail, 1u);
        const unsigned slot = t & (unsigned)(ir.ring_cap - 1);
        bool ok = false;
        for (int spin = 0; spin < PBB_SPIN_LIMIT; spin++) {
            if ((unsigned)(uni64(ld_rlx64<SC>(ir.ring + slot)) >> 32) == t) {
                ok = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (!ok) {
            fail(PBB_DEV_ERR_RING);
        } else if (lane == 0) {
            st_rlx64<SC>(ir.ring + slot, ((unsigned long long)(t + 1u) << 32) | (unsigned)(e + 1));
        }
    }

    // give up an exhausted interval; returns true for the explorer whose
    // release emptied the instance (all explorers idle: it is solved)
    __device__ __forceinline__ bool go_idle()
    {
        register_idle();
        int left = w_add_rlx<SC>(&ir.ctl->work_count, -1);
        return left == 1;
    }

    __device__ __forceinline__ void flush_counters()
    {
        if (lane == 0) {
            if (dec) atomicAdd(&ir.ctl->dec, (unsigned long long)dec);
            if (leaves) atomicAdd(&ir.ctl->leaves, (unsigned long long)leaves);
            if (sb) atomicAdd(&ir.ctl->sb, (unsigned long long)sb);
            if (jb) atomicAdd(&ir.ctl->jb, (unsigned long long)jb);
        }
        dec = leaves = sb = jb = 0;
        if (hist && lane <= n) atomicAdd(&ir.ctl->histo[lane], (unsigned long long)hist);
        hist = 0;
        if constexpr (REF != 0) {
            if (jhist && lane <= n) atomicAdd(&ir.ctl->jhisto[lane], (unsigned long long)jhist);
            jhist = 0;
        }
    }

    // incumbent word loaded by the caller (cost << 32 | explorer)
    __device__ __forceinline__ void take_best(unsigned long long b)
    {
        const int bc = uni((int)(unsigned)(b >> 32));
        if (bc < local_best) local_best = bc;
    }

    // periodic: node budget, deadline, host interrupt, incumbent refresh
    __device__ __forceinline__ bool check_stop()
    {
        if (a.node_budget) {
            unsigned long long before = w_add64<SC>(&ir.ctl->dec, (unsigned long long)dec);
            if (before + dec >= a.node_budget && lane == 0) st_rlx<SC>(&run->stop, 1);
            dec = 0;
        }
        if (dec >= (1u << 28) || sb >= (1u << 28) || jb >= (1u << 28)) flush_counters();  // 32-bit deltas
        if (t_end && wall_clock64() >= t_end && lane == 0) st_rlx<SC>(&run->stop, 1);
        poll_host_and_best();
        return uni(ld_rlx<SC>(&run->stop)) != 0;
    }

    // The pool's periodic check: the deadline (the launch's or the request's,
    // whichever is first) raises the instance's own stop word.  Returns 0 to
    // go on, 1 when the launch stops (interrupt, error: the explorer leaves),
    // 2 when only this instance stops.
    __device__ __forceinline__ int check_stop_slot(int *slot_stop)
    {
        if (dec >= (1u << 28) || sb >= (1u << 28) || jb >= (1u << 28)) flush_counters();  // 32-bit deltas
        if (t_end && wall_clock64() >= t_end && lane == 0) st_rlx<SC>(slot_stop, 1);
        poll_host_and_best();
        if (uni(ld_rlx<SC>(&run->stop)) != 0) return 1;
        return uni(ld_rlx<SC>(slot_stop)) != 0 ? 2 : 0;
    }

    __device__ __forceinline__ void poll_host_and_best()
    {
        const bool host_turn = (nchk++ & (PBB_HOST_POLL_CHECKS - 1u)) == 0u;
        if (a.host_flag && host_turn && lane == 0) {
            int hf = __hip_atomic_load(a.host_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            if (hf) st_rlx<SC>(&run->stop, 1);
        }
        unsigned long long b = ld_rlx<SC>(ir.best);
        int bc = uni((int)(unsigned)(b >> 32));
        if (bc < local_best) local_best = bc;
    }
};

}  // namespace pbb
