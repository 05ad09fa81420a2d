// pbb_kernels.h -- kernel argument block and launchers shared by the HIP
// kernels (pbb_kernels.hip) and the host engine (pbb_host.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pbb_device.h"

namespace pbb {

enum { ST_IDLE = 0, ST_ACTIVE = 1, ST_UNFOLD = 2 };

enum { PBB_DEV_ERR_RING = 1 };

#define PBB_SPIN_LIMIT (1 << 24)

// A busy explorer of a device-wide launch reads the idle ring every
// PBB_POLL_EVERY decomposed nodes (power of two), whether or not explorers
// wait in it.  Polling at every node while some waited (the earlier policy)
// had every busy explorer of a small tree race for the same ring ticket at
// every node: on a 5.0 M-node 19x10 LB2 tree with a frozen incumbent the
// plain engine took 94 / 275 / 354 ms at 256 / 1024 / 3072 explorers; polling
// every 4 / 8 / 16 nodes instead, 89 / 81 / 95, 88 / 52 / 61 and 89 / 47 / 53
// ms; the 512-instance batch at 1024 explorers per instance 599-929 ms
// before, 178-189 ms at 16; the 50x20 explorer 1984 M nodes/s before,
// 2036 M at 16.
#ifndef PBB_POLL_EVERY
#define PBB_POLL_EVERY 16
#endif

// A busy explorer reloads the shared incumbent every PBB_BEST_EVERY decomposed
// nodes (power of two; the load is issued before the node's work and applied
// after it), and an idle one when a donation wakes it, so an improvement found
// by one explorer prunes the others' subtrees within that many nodes.  With a
// frozen incumbent nothing changes.  Measured on the 20x5 NEH-incumbent solve
// (4096 explorers): refreshing only at the periodic check (every 64 nodes,
// which most explorers of a small tree never reach) decomposed 126-136 k nodes
// in 5-6 ms; on wake-up and every 64 / 16 / 8 nodes, 12-13 k / 6-10 k / 6 k
// nodes in 1.1-1.4 ms; the 50x20 explorer is within noise at 16.
#ifndef PBB_BEST_EVERY
#define PBB_BEST_EVERY 16
#endif

// The host's interrupt word lives in host memory, so reading it is a PCIe
// round trip.  A busy explorer reads it at every PBB_HOST_POLL_CHECKS-th
// periodic check (one check per 64 decomposed nodes; explorers start at
// staggered phases), and the first one to see it raises the device-side stop
// word that every explorer reads at each check.  Measured at 50x20 with 4096
// explorers: a host read every 64 nodes held the explorer at 720-1100 M
// nodes/s, varying with the host's load; one every 1024 nodes, 1530-1545 M.
#ifndef PBB_HOST_POLL_CHECKS
#define PBB_HOST_POLL_CHECKS 16
#endif

// Device-resident control block of one engine (or pool slot).  The ring
// tickets and the work counter take atomics from every donor; the flags are
// polled by idle explorers; each group sits on its own 128-byte line so the
// polls do not queue behind the atomics.
struct Ctl {
    unsigned idle_head;        // ring tickets (dequeue side)
    unsigned idle_tail;        // ring tickets (enqueue side)
    int pad0[30];
    int work_count;            // explorers holding an interval
    int pad1[31];
    int done;                  // set when work_count reaches 0
    int stop;                  // budget / deadline / interrupt / error
    int err;                   // device error code
    int pad2[29];
    unsigned long long dec;    // decomposed nodes
    unsigned long long leaves;
    unsigned long long sb;     // simple (LB1) bounds
    unsigned long long jb;     // refinement bounds
    unsigned long long histo[65];   // decomposed nodes per depth (fixed jobs)
    unsigned long long jhisto[65];  // refinement bounds per depth of the bounded node
};

// The instance one explorer works on: its control block, idle ring,
// incumbent and bounded root row (one per engine, one per pool slot).
struct InstRef {
    Ctl *ctl;
    unsigned long long *ring;   // idle-ring cells: ticket << 32 | explorer + 1
    int ring_cap;
    int ub0;                    // initial UB (branching thresholds when not singleNode)
    unsigned long long *best;   // cost << 32 | explorer
    const int *root_row;
    const int *root_dir;
    int m;                      // machines the pair table refers to
    int npairs;
    int pairs_mode;
    const int *pair_ab;
};

struct KArgs {
    DevInst I;
    int E;              // explorers (wavefronts)
    int rows;           // rows per explorer in mat (the largest n served)
    // explorer state in HBM (row stride 64 bytes)
    int8_t *mat;        // [E][n][64]
    int8_t *pos;        // [E][64]
    int8_t *end;        // [E][64]
    int8_t *dir;        // [E][64]
    int *line;          // [E]
    int *state;         // [E]
    // idle ring (bounded MPMC cells; each one word, ticket << 32 | explorer + 1,
    // so a cell moves with single relaxed atomics and no fence)
    unsigned long long *ring;   // [ring_cap]
    int ring_cap;               // power of two >= E
    Ctl *ctl;
    unsigned long long *best;   // cost << 32 | explorer
    int *best_perm;             // [E][64]
    const int *root_row;        // bounded root row [64]
    const int *root_dir;        // [1]
    int *root_row_out;          // written by k_bound_root
    int *root_dir_out;
    // run control
    unsigned long long node_budget;   // 0 = none
    unsigned long long deadline;      // wall_clock64 ticks, 0 = none
    const int *host_flag;             // host-mapped interrupt flag (may be NULL)
    // configuration
    int bound_mode, branching, refine, pairs_mode, find_all, single_node, initial_ub;
    int steal_margin;
    unsigned ramp;      // nodes an explorer decomposes in a fresh interval before it donates
};

// The root order of k_bound_root, passed by value (no upload before the launch).
struct RootPerm {
    int p[64];
};

// LDS footprint.  Workgroup tables: the machine-major processing times
// pm[m][PM_STRIDE] and the same rows in reversed machine order pmR, with a
// fixed odd row stride (lanes reading one job on different machines, or
// different jobs on one machine, hit distinct banks, and every per-machine
// offset is a compile-time immediate), min_arr, min_dep, per-machine totals,
// and the Johnson table npairs*n int4 when LB2 is used and it fits.  Per
// wavefront: the IVM matrix (row l holds its n-l cells, rows packed), the
// front/back stack ((n+1) x m ints), the LB1 operand block (2 x m int2) and,
// for the LB2 variant, 64 ints of scratch, then a 64 x int2 sink for the
// lanes past a row in stores that would otherwise need an exec mask.
#define PM_STRIDE 65
__device__ __host__ inline int lds_pm_stride(int) { return PM_STRIDE; }
__device__ __host__ inline size_t lds_table_bytes(int n, int m, int npairs, bool lb2)
{
    const int mp = (m + 3) & ~3;
    size_t off = ((size_t)2 * m * PM_STRIDE + 3 * (size_t)mp) * 4;
    off = (off + 15) & ~(size_t)15;
    if (lb2) off += (size_t)npairs * n * 16;
    return off;
}
// offset of IVM row l (rows 0..l-1 hold n, n-1, ..., n-l+1 cells)
__device__ __host__ inline int lds_mat_row(int l, int n) { return l * n - (l * (l - 1)) / 2; }
__device__ __host__ inline size_t lds_mat_bytes(int n)
{
    size_t t = (size_t)n * (n + 1) / 2;
    if (t < (size_t)4 * n) t = (size_t)4 * n;  // also holds n ints (NEH order at admission)
    return (t + 15) & ~(size_t)15;
}
__device__ __host__ inline size_t lds_stack_bytes(int n, int m) { return ((size_t)(n + 1) * m * 4 + 15) & ~(size_t)15; }
__device__ __host__ inline size_t lds_wave_bytes(int n, int m, bool scratch)
{
    return lds_mat_bytes(n) + lds_stack_bytes(n, m) + (size_t)m * 16 + (scratch ? 64 * 4 : 0) + 64 * 8;
}
// The Johnson table is staged in LDS when it is at most 64 KB (m <= 10 at
// n <= 50, m = 20 at n <= 21); larger tables are read from global memory.
__device__ __host__ inline bool jtab_in_lds(int n, int npairs) { return (size_t)npairs * n * 16 <= 64 * 1024; }
// Explorers (wavefronts) per workgroup of the one-instance engine; they share
// the workgroup's tables.
#ifndef PBB_EXPLORE_WAVES
#define PBB_EXPLORE_WAVES 4
#endif
inline size_t explore_lds_bytes(const DevInst &I, bool lb2)
{
    return lds_table_bytes(I.n, I.m, I.npairs, lb2 && jtab_in_lds(I.n, I.npairs)) +
           PBB_EXPLORE_WAVES * lds_wave_bytes(I.n, I.m, lb2);
}
// The team launch (one workgroup) holds at most 16 explorers with LB1 only
// (4 per SIMD, the register budget of the LB1 explorer) and 8 with a
// refinement bound (whose explorer needs more than 128 registers).
__device__ __host__ constexpr int team_waves_max(int ref) { return ref == 0 ? 16 : 8; }
inline size_t team_lds_bytes(const DevInst &I, bool lb2, int waves)
{
    return lds_table_bytes(I.n, I.m, I.npairs, lb2 && jtab_in_lds(I.n, I.npairs)) +
           (size_t)waves * lds_wave_bytes(I.n, I.m, lb2);
}
// LDS size handed to the occupancy query of the persistent kernels.  Their
// explorers wait on one another (the idle ring, work_count), so every
// workgroup of the grid must be resident at once.  The query's count can be
// one workgroup per CU too high when the workgroups fill the CU's LDS to
// within a granule (12x20 with the Johnson table: 3 x 53904 B of 163840 B
// queried as 3, held 2); rounding each workgroup up to 2 KB keeps the grid
// inside what the hardware actually co-schedules.
inline size_t residency_lds_bytes(size_t lds) { return (lds + 2047) & ~(size_t)2047; }
inline size_t root_lds_bytes(const DevInst &I)
{
    return lds_table_bytes(I.n, I.m, I.npairs, false) + lds_wave_bytes(I.n, I.m, false);
}

// ------------------------------------------------------------------ pool
// Many instances in one resident launch.  Workgroups of W explorers take one
// instance slot at a time: the first group to pop a slot from the request
// queue admits it (tables, idle ring, root), later groups join running slots
// whose explorers are all busy, up to a per-instance explorer cap; work moves
// between the members of a slot through its idle ring exactly as in the
// one-instance engine.  The last member to leave a solved slot writes its
// result record.  Instances with fewer machines than the pool's M are padded
// with zero-time machines, which changes none of the bounds or makespans.
enum { POOL_RUNNING = 1 << 30, POOL_DONE = 1 << 29, POOL_MEMBERS = (1 << 29) - 1 };
enum { RES_PENDING = 0, RES_SOLVED = 1, RES_ABORTED = 2, RES_INVALID = 3 };
enum { ROLE_EXIT = 0, ROLE_OWNER = 1, ROLE_HELPER = 2 };

#define PBB_POOL_MAXM 20
#define PBB_POOL_ROW 64

struct PoolReq {                  // one request (host-visible when serving)
    int n, m, ub, tag;
    int ptm[PBB_POOL_MAXM * PBB_POOL_ROW];   // [m][n] machine-major, row stride n
};

struct PoolRes {                  // written once, when the instance retires
    int tag;                      // the request's tag, stored last (completion flag)
    int status;                   // RES_*
    int best;                     // optimum (or the UB when no better schedule exists)
    int improved;                 // 1 when perm holds a schedule of cost best
    unsigned long long dec, leaves, sb, jb;
    unsigned long long t_admit, t_done;      // device wall clock
    int perm[PBB_POOL_ROW];
};

struct PoolQueue {                // request queue: slot ids published by tail
    int tail;
    int closed;
    int started;                  // set by every workgroup of k_pool that starts
    int pad[61];
    int ids[1];                   // [qcap]
};

struct PoolSlot {                 // device header of one instance slot
    Ctl c;
    unsigned long long best;      // cost << 32 | explorer
    unsigned long long t_admit;
    int occ;                      // POOL_RUNNING | POOL_DONE | member explorers
    int n, m, ub, tag, npairs, root_dir, best_cost, improved, status, pad[2];
    int root_row[PBB_POOL_ROW];
    int perm[PBB_POOL_ROW];
};

// byte offsets inside one slot's arena (all 256-byte aligned)
struct PoolLayout {
    size_t hdr, ptm, ptmT, minA, minD, pair, key, jtab, ring, stride;
};
__device__ __host__ inline size_t pool_al(size_t x) { return (x + 255) & ~(size_t)255; }
__device__ __host__ inline PoolLayout pool_layout(int nmax, int M, int pmax, int ring_cap)
{
    const int mp = (M + 3) & ~3;
    PoolLayout L;
    size_t o = 0;
    L.hdr = o;      o += pool_al(sizeof(PoolSlot));
    L.ptm = o;      o += pool_al((size_t)M * nmax * 4);
    L.ptmT = o;     o += pool_al((size_t)nmax * mp * 4);
    L.minA = o;     o += pool_al((size_t)mp * 4);
    L.minD = o;     o += pool_al((size_t)mp * 4);
    L.pair = o;     o += pool_al((size_t)(pmax > 0 ? pmax : 1) * 4);
    L.key = o;      o += pool_al((size_t)(pmax > 0 ? pmax : 1) * nmax * 4);
    L.jtab = o;     o += pool_al((size_t)(pmax > 0 ? pmax : 1) * nmax * 16);
    L.ring = o;     o += pool_al((size_t)ring_cap * 8);
    L.stride = o;
    return L;
}

struct PoolArgs {
    char *arena;                  // nslots x layout.stride
    PoolLayout lay;
    int nslots;
    int W;                        // explorers (wavefronts) per workgroup
    int epi;                      // explorer cap per instance
    int nmax;                     // largest n served
    int pmax;                     // pair-table capacity
    int ring_cap;                 // idle-ring cells per slot (power of two >= epi)
    int jt_lds;                   // Johnson tables staged in LDS
    int pairs_mode;               // JohnsonMode: pair list built at admission
    int qcap;                     // queue capacity (power of two)
    int sys;                      // req/res/q live in host-mapped memory
    const PoolReq *req;           // [nslots]
    PoolRes *res;                 // [nslots]
    PoolQueue *q;
    unsigned *q_head;             // consumer ticket (device memory)
    Ctl *run;                     // stop / error of the launch
    unsigned long long req_ticks; // per-request budget after admission (wall_clock64 ticks), 0 = none
    int ig_iters;                 // IG helper rounds per instance (0 = no helper wave)
};

__device__ __host__ inline size_t pool_lds_bytes(int nmax, int M, int pmax, bool jt_lds, int W)
{
    return lds_table_bytes(nmax, M, pmax, jt_lds) + (size_t)W * lds_wave_bytes(nmax, M, pmax > 0);
}

// per-M kernels (pbb_kernels.hip, one translation unit per M)
template <int M> hipError_t launch_explore_m(const KArgs &a, int blocks, hipStream_t s);
template <int M> hipError_t launch_team_m(const KArgs &a, hipStream_t s);
template <int M> hipError_t occupancy_m(const KArgs &a, int *blocks_per_cu);
template <int M> hipError_t launch_root_m(const KArgs &a, const RootPerm &root_perm, hipStream_t s);
template <int M>
hipError_t launch_bounds_m(int which, const KArgs &a, int count, const int *perms, const int *l1,
                           const int *l2, int arg, int *o1, int *o2, hipStream_t s);

// pool kernels (pbb_pool.hip, one translation unit per M)
template <int M> hipError_t launch_pool_m(const KArgs &a, const PoolArgs &p, int groups, int init_count,
                                          hipStream_t s);
template <int M> hipError_t pool_occupancy_m(const KArgs &a, const PoolArgs &p, int *groups_per_cu);
template <int M>
hipError_t launch_neh_m(const int *ptms, const long long *off, const int *ns, const int *ms, int count, int nmax,
                        int *perms, int *costs, hipStream_t s);

// iterated greedy (pbb_ig.hip, one translation unit per M): acceptance
// thresholds per instance for cost deltas 0..IG_THR-1
#define IG_THR 256
// LDS of the pool's IG helper wave, after the pool's own LDS: acceptance
// thresholds (IG_THR x u64), job totals (64), heads and tails (2 x M x 64)
// and scratch (64 ints)
__device__ __host__ inline size_t pool_ig_bytes(int M) { return (size_t)IG_THR * 8 + ((size_t)2 * M * 64 + 128) * 4; }
__device__ __host__ inline size_t pool_lds_total(const PoolArgs &p, int M)
{
    return pool_lds_bytes(p.nmax, M, p.pmax, p.jt_lds != 0, p.W) + (p.ig_iters > 0 ? pool_ig_bytes(M) : 0);
}
template <int M>
hipError_t launch_ig_m(const int *ptms, const long long *off, const int *ns, const int *ms, int count, int chains,
                       int iters, unsigned long long seed, const unsigned long long *thr, int *perms, int *costs,
                       hipStream_t s);

// runtime-M dispatch (pbb_dispatch.cpp)
hipError_t launch_explore(const KArgs &a, int blocks, hipStream_t s);
// one workgroup of a.E (<= PBB_TEAM_WAVES) explorers
hipError_t launch_team(const KArgs &a, hipStream_t s);
hipError_t explore_occupancy(const KArgs &a, int *blocks_per_cu);
hipError_t launch_root(const KArgs &a, const RootPerm &root_perm, hipStream_t s);
// engine set-up kernels (pbb_setup.hip, one translation unit): the whole
// search space to explorer 0 and every other explorer idle, with the
// control block cleared; the idle ring emptied and the stop word cleared
// before a launch
hipError_t launch_setup_full(const KArgs &a, hipStream_t s);
hipError_t launch_prelaunch(const KArgs &a, hipStream_t s);
// which: 0 lb1 children, 1 lb1 node, 2 lb2 node, 3 makespan
hipError_t launch_bounds(int which, const KArgs &a, int count, const int *perms, const int *l1,
                         const int *l2, int arg, int *o1, int *o2, hipStream_t s);
// init_count > 0: batch launch, the queue holds slots 0..init_count-1
hipError_t launch_pool(int M, const KArgs &a, const PoolArgs &p, int groups, int init_count, hipStream_t s);
hipError_t pool_occupancy(int M, const KArgs &a, const PoolArgs &p, int *groups_per_cu);
hipError_t launch_neh(int M, const int *ptms, const long long *off, const int *ns, const int *ms, int count,
                      int nmax, int *perms, int *costs, hipStream_t s);
hipError_t launch_ig(int M, const int *ptms, const long long *off, const int *ns, const int *ms, int count,
                     int chains, int iters, unsigned long long seed, const unsigned long long *thr, int *perms,
                     int *costs, hipStream_t s);

}  // namespace pbb
