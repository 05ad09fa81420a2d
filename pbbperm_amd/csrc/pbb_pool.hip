// pbb_pool.hip -- many flow-shop instances in one resident launch.
//
//   k_pool<M,REF>  persistent kernel over a table of instance slots.  A
//                  workgroup of W explorers (one per wavefront) serves one
//                  slot at a time:
//                    * owner: pops a slot id from the request queue, builds
//                      the instance's tables on the device (the same tables
//                      as bound_fsp_weak::init / fillMinTimeArrDep,
//                      bound_fsp_weak.cpp:17-44, :248-286, and
//                      bound_fsp_strong::remplirMachine / remplirLag /
//                      remplirTabJohnson, bound_fsp_strong.cpp:264-436),
//                      bounds the root (ivm_bound::boundRoot,
//                      ivm_bound.cpp:273-330) and explores the full interval
//                      from explorer 0 of the group;
//                    * helper: joins the running slot with the most busy
//                      explorers net of idle ones (up to the per-instance
//                      cap) and receives work
//                      through the slot's idle ring, as explorers of the
//                      one-instance engine do.
//                  The explorer whose release empties a slot records the
//                  optimum; the last member to leave writes the result
//                  record (host-visible when serving).  The launch ends when
//                  the queue is closed and drained and no slot runs, on the
//                  deadline, or on the host interrupt flag.
//   k_pool_init    resets the slot table and the device side of the queue
//                  (a kernel rather than memset, so the pair can be captured
//                  into a HIP graph).
//
// Translation units: compiled once per pool machine count M (-DPBB_M=<m>).
#include <algorithm>
#include <cstddef>

#include "pbb_ig.h"

namespace pbb {

// ---------------------------------------------------- system-scope access
// Request, result and queue records may live in host-mapped memory.
__device__ __forceinline__ int sys_ld_rlx(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void sys_st(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void sys_st64(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void sys_st_rel(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ------------------------------------------------------------ slot access
__device__ __forceinline__ char *slot_base(const PoolArgs &p, int s) { return p.arena + (size_t)s * p.lay.stride; }
__device__ __forceinline__ PoolSlot *slot_hdr(const PoolArgs &p, int s)
{
    return (PoolSlot *)(slot_base(p, s) + p.lay.hdr);
}

template <int M>
__device__ __forceinline__ DevInst slot_inst(const PoolArgs &p, int s, int n, int npairs)
{
    char *b = slot_base(p, s);
    DevInst I;
    I.n = n;
    I.m = M;
    I.mp = (M + 3) & ~3;
    I.ptm = (const int *)(b + p.lay.ptm);
    I.ptmT = (const int *)(b + p.lay.ptmT);
    I.min_arr = (const int *)(b + p.lay.minA);
    I.min_dep = (const int *)(b + p.lay.minD);
    I.npairs = npairs;
    I.pair_ab = (const int *)(b + p.lay.pair);
    I.jtab = (const int4 *)(b + p.lay.jtab);
    return I;
}

__device__ __forceinline__ InstRef slot_ref(const PoolArgs &p, int s, PoolSlot *S, int m, int npairs, int ub)
{
    char *b = slot_base(p, s);
    InstRef r;
    r.ctl = &S->c;
    r.ring = (unsigned long long *)(b + p.lay.ring);
    r.ring_cap = p.ring_cap;
    r.ub0 = ub;
    r.best = &S->best;
    r.root_row = S->root_row;
    r.root_dir = &S->root_dir;
    r.m = m;
    r.npairs = npairs;
    r.pairs_mode = 0;  // the pair list is built for the configured JohnsonMode
    r.pair_ab = (const int *)(b + p.lay.pair);
    return r;
}

// ---------------------------------------------------------- admission
// The whole workgroup builds slot s's tables from its request.  Machines
// m..M-1 get zero times.  Returns false (uniformly) for a malformed request.
template <int M>
static __device__ __forceinline__ bool pool_admit(const PoolArgs &p, int s, bool lb2)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    char *b = slot_base(p, s);
    PoolSlot *S = (PoolSlot *)(b + p.lay.hdr);
    const PoolReq *R = p.req + s;
    const int n = sys_ld_rlx(&R->n), m = sys_ld_rlx(&R->m), ub = sys_ld_rlx(&R->ub), tag = sys_ld_rlx(&R->tag);
    const bool ok = n >= 3 && n <= p.nmax && m >= 2 && m <= M;
    const int mp = (M + 3) & ~3;
    int *ptm = (int *)(b + p.lay.ptm);
    int *ptmT = (int *)(b + p.lay.ptmT);
    int *minA = (int *)(b + p.lay.minA);
    int *minD = (int *)(b + p.lay.minD);
    int *pair = (int *)(b + p.lay.pair);
    int *key = (int *)(b + p.lay.key);
    int4 *jtab = (int4 *)(b + p.lay.jtab);
    int npairs = 0;
    if (lb2 && ok) npairs = p.pairs_mode == 0 ? m * (m - 1) / 2 : m - 1;
    if (ok) {
        for (int i = tid; i < M * n; i += nt) {
            int q = i / n, j = i - q * n;
            ptm[i] = q < m ? sys_ld_rlx(&R->ptm[q * n + j]) : 0;
        }
    }
    __threadfence_block();
    __syncthreads();
    if (ok) {
        for (int i = tid; i < n * mp; i += nt) {
            int j = i / mp, q = i - j * mp;
            ptmT[i] = q < M ? ptm[q * n + j] : 0;
        }
        // fillMinTimeArrDep: min over jobs of the head before / tail after machine q
        for (int q = tid; q < mp; q += nt) {
            int ha = 0, td = 0;
            if (q < M) {
                ha = td = 0x7fffffff;
                for (int j = 0; j < n; j++) {
                    int h = 0, t = 0;
                    for (int k = 0; k < q; k++) h += ptm[k * n + j];
                    for (int k = q + 1; k < M; k++) t += ptm[k * n + j];
                    ha = min(ha, h);
                    td = min(td, t);
                }
            }
            minA[q] = ha;
            minD[q] = td;
        }
        // machine pairs of the configured JohnsonMode over the real machines,
        // lexicographic (bound_fsp_strong::remplirMachine)
        for (int t = tid; t < M * M; t += nt) {
            int a = t / M, c = t - a * M;
            int idx = -1;
            if (npairs && c < m && a < c) {
                if (p.pairs_mode == 0) idx = a * (2 * m - a - 1) / 2 + (c - a - 1);
                else if (p.pairs_mode == 1 && c == m - 1) idx = a;
                else if (p.pairs_mode == 2 && c == a + 1) idx = a;
            }
            if (idx >= 0) pair[idx] = a | (c << 8);
        }
    }
    __threadfence_block();
    __syncthreads();
    // Johnson/Mitten order per pair: set 1 (p_a < p_b) ascending p_a + lag,
    // then set 2 descending p_b + lag, ties by job index.  The sort key packs
    // set and direction so that the order is ascending key, then job.
    for (int i = tid; i < npairs * n; i += nt) {
        int pi = i / n, j = i - pi * n;
        int ab = pair[pi], a = ab & 0xff, c = ab >> 8;
        int pa = ptm[a * n + j], pc = ptm[c * n + j], lag = 0;
        for (int k = a + 1; k < c; k++) lag += ptm[k * n + j];
        bool first = pa < pc;
        int kv = (first ? pa : pc) + lag;
        key[i] = first ? kv : (1 << 30) + ((1 << 30) - 1 - kv);
    }
    __threadfence_block();
    __syncthreads();
    for (int i = tid; i < npairs * n; i += nt) {
        int pi = i / n, j = i - pi * n;
        int kj = key[i], rank = 0;
        const int *kr = key + pi * n;
        for (int jj = 0; jj < n; jj++) {
            int ki = kr[jj];
            rank += (ki < kj || (ki == kj && jj < j)) ? 1 : 0;
        }
        int ab = pair[pi], a = ab & 0xff, c = ab >> 8;
        int lag = 0;
        for (int k = a + 1; k < c; k++) lag += ptm[k * n + j];
        jtab[pi * n + rank] = make_int4(j, ptm[a * n + j], ptm[c * n + j], lag);
    }
    // control block, incumbent, idle ring
    int *ci = (int *)&S->c;
    const int wc = (int)(offsetof(Ctl, work_count) / 4);
    // explorer 0 holds the full interval; an IG helper wave holds one more unit
    const int held = p.ig_iters > 0 ? 2 : 1;
    for (int i = tid; i < (int)(sizeof(Ctl) / 4); i += nt) ci[i] = (i == wc) ? held : 0;
    unsigned long long *rs = (unsigned long long *)(b + p.lay.ring);
    for (int i = tid; i < p.ring_cap; i += nt) rs[i] = (unsigned long long)(unsigned)i << 32;
    if (tid == 0) {
        S->best = ((unsigned long long)(unsigned)ub << 32) | 0xffffffffull;
        S->t_admit = wall_clock64();
        S->n = ok ? n : 0;
        S->m = m;
        S->ub = ub;
        S->tag = tag;
        S->npairs = npairs;
        S->best_cost = ub;
        S->improved = 0;
        S->status = ok ? RES_PENDING : RES_INVALID;
    }
    __threadfence();
    __syncthreads();
    return ok;
}

// -------------------------------------------------------- slot selection
// Wave 0 of a free workgroup: a new instance from the queue first, else the
// running slot with the largest surplus of busy over idle explorers (below
// the per-instance cap), else wait.  Returns the role; *slot receives the slot id.
static __device__ __forceinline__ int pool_pick(const PoolArgs &p, int W, unsigned long long t_end, const int *host_flag, int *slot)
{
    const int lane = lane_id();
    // A waiting group passes here every microsecond or so, and hundreds may
    // wait at once (the tail of a batch, an idle server): the host word is
    // read at every PBB_HOST_POLL_CHECKS-th pass only, and the queue tail
    // relaxed, with the acquire (an L2 invalidate) taken only by the group
    // that claims a request (512-instance batch: 0.92-1.0 s before, 0.48-0.50
    // s with both and the relaxed idle polls).
    for (unsigned it = 0;; it++) {
        if (uni(ld_rlx(&p.run->stop))) return ROLE_EXIT;
        if (t_end && uni64(wall_clock64()) >= t_end) return ROLE_EXIT;
        if (host_flag && (it & (PBB_HOST_POLL_CHECKS - 1u)) == 0u &&
            uni(__hip_atomic_load(host_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)))
            return ROLE_EXIT;
        const int tail = uni(sys_ld_rlx(&p.q->tail));
        const unsigned head = (unsigned)uni(ld_rlx((const int *)p.q_head));
        if ((int)head < tail) {
            if (w_cas(p.q_head, head, head + 1u)) {
                // pairs with the host's release of the tail read above
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
                *slot = uni(sys_ld_rlx(&p.q->ids[head & (unsigned)(p.qcap - 1)]));
                // the slot counts as occupied from here on: a group that finds
                // the queue closed and drained waits while a request is
                // still being admitted, instead of leaving the launch
                if (lane == 0)
                    __hip_atomic_store(&slot_hdr(p, *slot)->occ, W, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return ROLE_OWNER;
            }
            continue;
        }
        int best_score = -1, best_slot = -1, best_occ = 0;
        bool running = false;
        for (int base = 0; base < p.nslots; base += PBB_WAVE) {
            const int s = base + lane;
            int o = 0, wc = 0;
            if (s < p.nslots) o = ld_rlx(&slot_hdr(p, s)->occ);
            const bool live = (o & POOL_RUNNING) && !(o & POOL_DONE);
            if (live) wc = ld_rlx(&slot_hdr(p, s)->c.work_count);
            const int mem = o & POOL_MEMBERS;
            // a slot whose request ran out of time is draining: do not join it
            const bool stopping = live && ld_rlx(&slot_hdr(p, s)->c.stop) != 0;
            // more busy than idle members: work is there to be split
            const bool cand = live && !stopping && mem + W <= p.epi && 2 * wc > mem;
            running = running || ballot(o != 0) != 0ull;
            const int sc = cand ? 2 * wc - mem : -1;
            const int mx = wave_max(sc);
            if (mx > best_score) {
                const int l = __ffsll((unsigned long long)ballot(sc == mx)) - 1;
                best_score = mx;
                best_slot = base + l;
                best_occ = rdlane(o, l);
            }
        }
        if (best_slot >= 0) {
            if (w_cas((unsigned *)&slot_hdr(p, best_slot)->occ, (unsigned)best_occ, (unsigned)(best_occ + W))) {
                *slot = best_slot;
                return ROLE_HELPER;
            }
            continue;
        }
        const int closed = uni(sys_ld_rlx(&p.q->closed));
        if (closed && (int)head >= tail && !running) return ROLE_EXIT;
        __builtin_amdgcn_s_sleep(32);
    }
}

// ------------------------------------------------------------ retirement
// the explorer whose release emptied the slot: record the incumbent and its
// schedule while every member is still inside, then mark the slot done --
// solved, or aborted when its stop word was raised (the request's time
// budget ran out and its explorers gave their intervals up)
template <int M, int REF>
__device__ __forceinline__ void pool_finish(Explorer<M, REF> &X, const KArgs &a, PoolSlot *S)
{
    const unsigned long long b =
        __hip_atomic_load(&S->best, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const int cost = uni((int)(unsigned)(b >> 32));
    const unsigned who = (unsigned)uni((int)(unsigned)(b & 0xffffffffull));
    const bool imp = who < (unsigned)a.E;
    if (imp && X.lane < X.n) S->perm[X.lane] = a.best_perm[(size_t)who * 64 + X.lane];
    const bool stopped = uni(ld_rlx(&S->c.stop)) != 0;
    if (X.lane == 0) {
        S->best_cost = cost;
        if (imp) S->improved = 1;  // else the schedule given at admission (NEH), if any, stands
        S->status = stopped ? RES_ABORTED : RES_SOLVED;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (X.lane == 0) {
        st_rel(&S->c.done, 1);
        __hip_atomic_fetch_or(&S->occ, (int)POOL_DONE, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// wave 0 of a leaving workgroup; the last member writes the result record
static __device__ __forceinline__ void pool_leave(const PoolArgs &p, int s, int W)
{
    const int lane = lane_id();
    PoolSlot *S = slot_hdr(p, s);
    const int old = w_add(&S->occ, -W);
    if ((old & POOL_MEMBERS) != W) return;
    PoolRes *R = p.res + s;
    const int status = (old & POOL_DONE) ? uni(ld_rlx(&S->status)) : (int)RES_ABORTED;
    const int improved = uni(ld_rlx(&S->improved));
    if (lane < PBB_POOL_ROW) sys_st(&R->perm[lane], improved ? ld_rlx(&S->perm[lane]) : -1);
    if (lane == 0) {
        sys_st(&R->status, status);
        sys_st(&R->best, ld_rlx(&S->best_cost));
        sys_st(&R->improved, improved);
        sys_st64(&R->dec, ld_rlx(&S->c.dec));
        sys_st64(&R->leaves, ld_rlx(&S->c.leaves));
        sys_st64(&R->sb, ld_rlx(&S->c.sb));
        sys_st64(&R->jb, ld_rlx(&S->c.jb));
        sys_st64(&R->t_admit, ld_rlx(&S->t_admit));
        sys_st64(&R->t_done, wall_clock64());
    }
    const int tag = uni(ld_rlx(&S->tag));
    if (lane == 0) __hip_atomic_exchange(&S->occ, 0, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    if (lane == 0) sys_st_rel(&R->tag, tag);
}

// ------------------------------------------------------------ IG helper
// The last wave of an instance's admitting workgroup runs one chain of
// IG.cpp runIG (pbb_ig.h) beside the search and publishes every improvement
// as the instance's incumbent: its schedule into this explorer's best_perm
// row, then the incumbent word (the explorers reload it every
// PBB_BEST_EVERY nodes and prune with it).  The wave holds one unit of the
// slot's work counter from admission on, so the instance cannot retire while
// the chain may still write its schedule; the chain stops once it is the
// only holder left (the search is over), on the slot's or the launch's stop
// word, or at the deadline, and the wave then gives the unit up like an
// explorer giving up its interval (the last one out records the result).
// Acceptance temperature as the reference's (IG.cpp:319-338): 0.2 x the
// mean processing time / 10, thresholds floor(2^32 exp(-delta / T)).
// It runs before the wave builds its explorer, so the chain's registers and
// the explorer's are never live together.
template <int M>
__device__ __forceinline__ void pool_ig_helper(const Lds &lds, int n, int e, const KArgs &a, const PoolArgs &p,
                                               PoolSlot *S, int slot, char *smem, unsigned long long t_end)
{
    const int lane = lane_id();
    unsigned long long *thr =
        (unsigned long long *)(smem + pool_lds_bytes(p.nmax, M, p.pmax, p.jt_lds != 0, p.W));
    int *spt = (int *)(thr + IG_THR);
    int *hd = spt + 64, *tl = hd + M * 64, *scr = tl + M * 64;
    int tot = 0;
    if (lane < n)
        for (int q = 0; q < M; q++) tot += lds.pm[q * PM_STRIDE + lane];
    spt[lane] = tot;
    const int m = uni(ld_rlx(&S->m));
    const double T = 0.2 * ((double)wave_sum(tot) / ((double)n * m)) / 10.0;
    for (int d = lane; d < IG_THR; d += PBB_WAVE) {
        const double pr = T > 0.0 ? exp(-(double)d / T) : (d == 0 ? 1.0 : 0.0);
        const double v = floor(pr * 4294967296.0);
        thr[d] = v >= 4294967296.0 ? (1ull << 32) : (unsigned long long)v;
    }
    __builtin_amdgcn_wave_barrier();
    IgWave<M> G{lds.pm, spt, hd, tl, scr, n, lane};
    uint64_t rng = 0x243F6A8885A308D3ull ^ (((uint64_t)(uint32_t)ld_rlx(&S->tag) << 32) | (uint32_t)slot);
    rng = ig_next(rng);
    const auto stop = [&] {
        return uni(ld_rlx(&S->c.work_count)) <= 1 || uni(ld_rlx(&S->c.stop)) != 0 ||
               uni(ld_rlx(&p.run->stop)) != 0 || (t_end && uni64(wall_clock64()) >= t_end);
    };
    const auto publish = [&](int cost, int seq) {
        const int cur = uni((int)(unsigned)(ld_rlx(&S->best) >> 32));
        if (cost >= cur) return;
        if (lane < n) a.best_perm[(size_t)e * 64 + lane] = seq;
        w_min64(&S->best, ((unsigned long long)(unsigned)cost << 32) | (unsigned)e);
    };
    int best = 0;
    ig_chain<M>(G, rng, thr, p.ig_iters, &best, stop, publish);
}

// ----------------------------------------------------------------- kernel
// Minimum resident waves per SIMD the compiler must allow (register
// budget): 3 up to 10 machines (168 VGPRs: the 10-machine LB2 pool would
// otherwise take 169-176 and drop to 2 waves per SIMD), free above.
#ifndef PBB_POOL_MIN_WAVES
#define PBB_POOL_MIN_WAVES(M) ((M) <= 10 ? 3 : 1)
#endif

template <int M, int REF>
__global__ __launch_bounds__(256, PBB_POOL_MIN_WAVES(M)) void k_pool(KArgs a, PoolArgs p)
{
    extern __shared__ __align__(16) char smem[];
    __shared__ int s_pick[2];
    const int W = p.W;
    const int wave = threadIdx.x >> 6;
    const int e = blockIdx.x * W + wave;
    const int lane = lane_id();
    const bool lb2 = REF == 2;
    unsigned long long t_end = 0ull;
    if (threadIdx.x == 0) sys_st(&p.q->started, 1);
    if (a.deadline) {
        unsigned long long t0 = wall_clock64();
        if (threadIdx.x == 0) s_pick[0] = (int)(unsigned)t0, s_pick[1] = (int)(unsigned)(t0 >> 32);
        __syncthreads();
        t_end = (((unsigned long long)(unsigned)s_pick[1] << 32) | (unsigned)s_pick[0]) + a.deadline;
        __syncthreads();
    }
    for (;;) {
        if (wave == 0) {
            int slot = -1;
            const int role = pool_pick(p, W, t_end, a.host_flag, &slot);
            if (lane == 0) {
                s_pick[0] = slot;
                s_pick[1] = role;
            }
        }
        __syncthreads();
        const int slot = s_pick[0], role = s_pick[1];
        if (role == ROLE_EXIT) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        PoolSlot *S = slot_hdr(p, slot);
        bool ok = true;
        if (role == ROLE_OWNER) ok = pool_admit<M>(p, slot, lb2);
        else ok = ld_rlx(&S->status) != RES_INVALID;
        if (!ok) {
            // a malformed request retires at once with RES_INVALID
            if (threadIdx.x == 0)
                __hip_atomic_exchange(&S->occ, (int)(POOL_RUNNING | POOL_DONE) + W, __ATOMIC_RELEASE,
                                      __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const int n = uni(ld_rlx(&S->n)), m = uni(ld_rlx(&S->m)), npairs = uni(ld_rlx(&S->npairs));
            const DevInst I = slot_inst<M>(p, slot, n, npairs);
            const Lds lds = lds_setup(I, smem, p.jt_lds != 0, REF == 2);
            if (role == ROLE_OWNER) {
                // a request without an incumbent gets NEH's (ub < 0)
                if (wave == 0 && uni(ld_rlx(&S->ub)) < 0) {
                    int *seq = (int *)lds.mat;  // the IVM rows are not in use yet
                    const int c = neh_wave<M>(lds.pm, PM_STRIDE, n, seq);
                    if (lane < n) S->perm[lane] = seq[lane];
                    if (lane == 0) {
                        S->ub = c;
                        S->best_cost = c;
                        S->improved = 1;
                        S->best = ((unsigned long long)(unsigned)c << 32) | 0xffffffffull;
                    }
                    __threadfence();
                }
                __syncthreads();
            }
            const int ub = uni(ld_rlx(&S->ub));
            // The request's own deadline (admission + its budget) and the
            // launch's, whichever comes first, raise the slot's stop word: its
            // explorers then give their intervals up as if exhausted, so the
            // slot drains through the ordinary release path and every member
            // leaves together once it is done (none leaves while a donor might
            // still hand it work).
            unsigned long long slot_end = t_end;
            if (p.req_ticks) {
                const unsigned long long dl = uni64(ld_rlx(&S->t_admit)) + p.req_ticks;
                slot_end = t_end ? (t_end < dl ? t_end : dl) : dl;
            }
            // the IG helper wave meets the owner's barrier below, then runs its
            // chain before it builds its explorer
            const bool igw = role == ROLE_OWNER && p.ig_iters > 0 && wave == W - 1;
            if (igw) {
                __syncthreads();
                pool_ig_helper<M>(lds, n, e, a, p, S, slot, smem, slot_end);
            }
            Explorer<M, REF> X(a, lds, e, slot_ref(p, slot, S, m, npairs, ub), p.run, n, I.mp);
            X.t_end = slot_end;
            X.local_best = uni((int)(unsigned)(ld_rlx(&S->best) >> 32));
            int st = ST_IDLE;
            if (igw) {
                if (X.go_idle()) pool_finish(X, a, S);
            } else if (role == ROLE_OWNER) {
                if (wave == 0) {
                    // boundRoot: the identity order bounded with LB1 and pruned
                    // against the initial UB (not counted, as in the engine)
                    X.rowv = lane < n ? lane : 0;
                    X.reset_stacks();
                    X.bound_row(true);
                    X.sb = 0;
                }
                __syncthreads();
                if (threadIdx.x == 0)
                    __hip_atomic_exchange(&S->occ, (int)POOL_RUNNING + W, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                if (wave == 0) {
                    if (X.start_root()) {
                        st = ST_ACTIVE;
                    } else if (X.go_idle()) {
                        pool_finish(X, a, S);
                    }
                } else {
                    X.register_idle();
                }
            } else {
                X.register_idle();
            }
            unsigned iter = 0;
            for (;;) {
                if (st == ST_IDLE) {
                    bool leave = false;
                    int nap = 1;
                    while (true) {
                        const int sv = uni(ld_rlx(a.state + e));
                        if (sv == ST_ACTIVE) {
                            // pairs with the donor's release of the state word
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                            st = ST_ACTIVE;
                            X.load_state();
                            X.take_best(ld_rlx(&S->best));
                            break;
                        }
                        if (uni(ld_rlx(&S->c.done)) || uni(ld_rlx(&p.run->stop)) ||
                            (t_end && uni64(wall_clock64()) >= t_end)) {
                            leave = true;
                            break;
                        }
                        X.idle_pause(nap);
                    }
                    if (leave) break;
                }
                const bool poll = (iter & (PBB_POLL_EVERY - 1)) == 0u;
                unsigned rh = 0, rt = 0;
                if (poll) {
                    rh = ld_rlx(&S->c.idle_head);
                    rt = ld_rlx(&S->c.idle_tail);
                }
                const bool refresh = (iter & (PBB_BEST_EVERY - 1)) == 0u;
                unsigned long long bw = ~0ull;
                if (refresh) bw = ld_rlx(&S->best);
                if (!X.step()) {
                    if (X.go_idle()) pool_finish(X, a, S);
                    st = ST_IDLE;
                    continue;
                }
                if (refresh) X.take_best(bw);
                if ((++iter & 63u) == 0u) {
                    const int sc = X.check_stop_slot(&S->c.stop);
                    if (sc == 1) break;  // the launch stops
                    if (sc == 2) {       // this request stops: release the interval
                        if (X.go_idle()) pool_finish(X, a, S);
                        st = ST_IDLE;
                        continue;
                    }
                }
                if (poll) X.try_donate(rh, rt);
            }
            X.flush_counters();
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (wave == 0) pool_leave(p, slot, W);
        __syncthreads();
    }
}

// reset the slot table, the queue's device side, the launch control block;
// a batch launch (count > 0) also fills the queue with slots 0..count-1
template <int M>
__global__ void k_pool_init(PoolArgs p, int count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = gridDim.x * blockDim.x;
    for (int s = t; s < p.nslots; s += nt) {
        slot_hdr(p, s)->occ = 0;
        if (count > 0) {
            p.res[s].tag = -1;
            p.res[s].status = RES_PENDING;
        }
    }
    if (count > 0)
        for (int i = t; i < count; i += nt) p.q->ids[i] = i;
    if (t == 0) {
        *p.q_head = 0u;
        p.run->stop = 0;
        p.run->err = 0;
        if (count > 0) {
            p.q->tail = count;
            p.q->closed = 1;
        }
    }
}

// NEH for many instances, one wavefront each (times [m][n] at ptms + off[i],
// padded to M machines with zero times, which changes no partial makespan).
// A shape outside 3 <= n <= 63, 2 <= m <= M gets cost -1 and order -1 (the
// host checks shapes before the allocating entry point launches this; the
// allocation-free one leaves the check to the device).
template <int M>
__global__ __launch_bounds__(64) void k_neh(const int *ptms, const long long *off, const int *ns, const int *ms,
                                            int count, int nmax, int *perms, int *costs)
{
    extern __shared__ __align__(16) char smem[];
    const int s = blockIdx.x;
    if (s >= count) return;
    const int n = ns[s], m = ms[s], lane = lane_id();
    if (n < 3 || n > PBB_POOL_ROW - 1 || m < 2 || m > M) {
        perms[(size_t)s * 64 + lane] = -1;
        if (lane == 0) costs[s] = -1;
        return;
    }
    const int ps = lds_pm_stride(nmax);
    int *pm = (int *)smem;
    int *seq = pm + M * ps;
    const int *src = ptms + off[s];
    for (int i = lane; i < M * n; i += PBB_WAVE) {
        const int q = i / n, j = i - q * n;
        pm[q * ps + j] = q < m ? src[q * n + j] : 0;
    }
    __builtin_amdgcn_wave_barrier();
    const int c = neh_wave<M>(pm, ps, n, seq);
    if (lane < n) perms[(size_t)s * 64 + lane] = seq[lane];
    if (lane == 0) costs[s] = c;
}

// ---------------------------------------------------------- host launchers
template <int M>
static const void *pool_kernel(const KArgs &a)
{
    if (a.bound_mode == 0) return (const void *)k_pool<M, 0>;
    return a.refine == 1 ? (const void *)k_pool<M, 2> : (const void *)k_pool<M, 1>;
}

template <int M>
hipError_t launch_pool_m(const KArgs &a, const PoolArgs &p, int groups, int init_count, hipStream_t s)
{
    const size_t lds = pool_lds_total(p, M);
    hipLaunchKernelGGL(k_pool_init<M>, dim3(std::max(1, std::min(1024, (std::max(p.nslots, init_count) + 255) / 256))),
                       dim3(256), 0, s, p, init_count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (const hipError_t err =
            hipFuncSetAttribute(pool_kernel<M>(a), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds))
        return err;
    if (a.bound_mode == 0)
        hipLaunchKernelGGL((k_pool<M, 0>), dim3(groups), dim3(64 * p.W), lds, s, a, p);
    else if (a.refine == 1)
        hipLaunchKernelGGL((k_pool<M, 2>), dim3(groups), dim3(64 * p.W), lds, s, a, p);
    else
        hipLaunchKernelGGL((k_pool<M, 1>), dim3(groups), dim3(64 * p.W), lds, s, a, p);
    return hipGetLastError();
}

template <int M>
hipError_t pool_occupancy_m(const KArgs &a, const PoolArgs &p, int *groups_per_cu)
{
    const size_t lds = pool_lds_total(p, M);
    if (lds > 160 * 1024) {
        *groups_per_cu = 0;
        return hipSuccess;
    }
    const size_t q = residency_lds_bytes(lds);
    if (const hipError_t err =
            hipFuncSetAttribute(pool_kernel<M>(a), hipFuncAttributeMaxDynamicSharedMemorySize, (int)q))
        return err;
    if (a.bound_mode == 0)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(groups_per_cu, k_pool<M, 0>, 64 * p.W, q);
    if (a.refine == 1)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(groups_per_cu, k_pool<M, 2>, 64 * p.W, q);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(groups_per_cu, k_pool<M, 1>, 64 * p.W, q);
}

template <int M>
hipError_t launch_neh_m(const int *ptms, const long long *off, const int *ns, const int *ms, int count, int nmax,
                        int *perms, int *costs, hipStream_t s)
{
    const size_t lds = (size_t)M * lds_pm_stride(nmax) * 4 + 64 * 4;
    hipLaunchKernelGGL(k_neh<M>, dim3(count), dim3(64), lds, s, ptms, off, ns, ms, count, nmax, perms, costs);
    return hipGetLastError();
}

#ifdef PBB_M
template hipError_t launch_neh_m<PBB_M>(const int *, const long long *, const int *, const int *, int, int, int *,
                                        int *, hipStream_t);
template hipError_t launch_pool_m<PBB_M>(const KArgs &, const PoolArgs &, int, int, hipStream_t);
template hipError_t pool_occupancy_m<PBB_M>(const KArgs &, const PoolArgs &, int *);
#endif

}  // namespace pbb
