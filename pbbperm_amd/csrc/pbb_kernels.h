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

// Device-resident control block of one engine.
struct Ctl {
    unsigned idle_head;        // ring tickets (dequeue side)
    unsigned idle_tail;        // ring tickets (enqueue side)
    int work_count;            // explorers holding an interval
    int done;                  // set when work_count reaches 0
    int stop;                  // budget / deadline / interrupt / error
    int err;                   // device error code
    int pad[2];
    unsigned long long dec;    // decomposed nodes
    unsigned long long leaves;
    unsigned long long sb;     // simple (LB1) bounds
    unsigned long long jb;     // refinement bounds
    unsigned long long histo[65];
};

struct KArgs {
    DevInst I;
    int E;              // explorers (wavefronts)
    // explorer state in HBM (row stride 64 bytes)
    int8_t *mat;        // [E][n][64]
    int8_t *pos;        // [E][64]
    int8_t *end;        // [E][64]
    int8_t *dir;        // [E][64]
    int *line;          // [E]
    int *state;         // [E]
    // idle ring (bounded MPMC cells)
    int *ring_id;       // [ring_cap]
    unsigned *ring_seq; // [ring_cap]
    int ring_cap;       // power of two >= E
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
};

// LDS footprint.  Workgroup tables: job-major processing times ptmT[n][mp],
// the machine-major copy pm[m][n] (stored in an n*mp slot), min_arr, min_dep, per-machine
// totals, and the Johnson table npairs*n int4 when LB2 is used and it fits.
// Per wavefront: the IVM matrix (n rows of 64 bytes), the front/back stack
// ((n+1) x m ints), the LB1 operand block (2 x m int2) and 64 ints of scratch.
__device__ __host__ inline size_t lds_table_bytes(int n, int mp, int npairs, bool lb2)
{
    size_t off = (size_t)(2 * n + 3) * mp * 4;
    off = (off + 15) & ~(size_t)15;
    if (lb2) off += (size_t)npairs * n * 16;
    return off;
}
__device__ __host__ inline size_t lds_stack_bytes(int n, int m) { return ((size_t)(n + 1) * m * 4 + 15) & ~(size_t)15; }
__device__ __host__ inline size_t lds_wave_bytes(int n, int m)
{
    return (size_t)n * 64 + lds_stack_bytes(n, m) + (size_t)m * 16 + 64 * 4;
}
// The Johnson table is staged in LDS when it is at most 64 KB (m <= 10 at
// n <= 50, m = 20 at n <= 21); larger tables are read from global memory.
__device__ __host__ inline bool jtab_in_lds(int n, int npairs) { return (size_t)npairs * n * 16 <= 64 * 1024; }
inline size_t explore_lds_bytes(const DevInst &I, bool lb2)
{
    return lds_table_bytes(I.n, I.mp, I.npairs, lb2 && jtab_in_lds(I.n, I.npairs)) + 4 * lds_wave_bytes(I.n, I.m);
}
inline size_t root_lds_bytes(const DevInst &I)
{
    return lds_table_bytes(I.n, I.mp, I.npairs, false) + lds_wave_bytes(I.n, I.m);
}

// per-M kernels (pbb_kernels.hip, one translation unit per M)
template <int M> hipError_t launch_explore_m(const KArgs &a, int blocks, hipStream_t s);
template <int M> hipError_t occupancy_m(const KArgs &a, int *blocks_per_cu);
template <int M> hipError_t launch_root_m(const KArgs &a, const int *root_perm, hipStream_t s);
template <int M>
hipError_t launch_bounds_m(int which, const KArgs &a, int count, const int *perms, const int *l1,
                           const int *l2, int arg, int *o1, int *o2, hipStream_t s);

// runtime-M dispatch (pbb_dispatch.cpp)
hipError_t launch_explore(const KArgs &a, int blocks, hipStream_t s);
hipError_t explore_occupancy(const KArgs &a, int *blocks_per_cu);
hipError_t launch_root(const KArgs &a, const int *root_perm, hipStream_t s);
// which: 0 lb1 children, 1 lb1 node, 2 lb2 node, 3 makespan
hipError_t launch_bounds(int which, const KArgs &a, int count, const int *perms, const int *l1,
                         const int *l2, int arg, int *o1, int *o2, hipStream_t s);

}  // namespace pbb
