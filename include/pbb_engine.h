/*
 * pbb_engine.h -- C ABI of the MI355X IVM branch-and-bound engine for the
 * permutation flow-shop problem (makespan objective).
 *
 * Drop-in boundary for three reference surfaces (jangmys/PBBPerm paths):
 *   (i)   the lower-bound library    bounds/include/bound_abstract.h:6-34
 *   (ii)  the GPU worker engine      gpu/include/gpubb.h:24-173 as driven by
 *                                    distributed/src/worker_gpu.cpp:26-108
 *   (iii) the factoradic work unit   distributed/include/fact_work.h:13-41
 * and the gpuconfig.ini option surface (common/src/arguments.cpp:68-133).
 *
 * Conventions: every entry point returns 0 on success or a negative PBB_E*
 * code (never exits the process); pbb_last_error() gives a message for the
 * calling thread.  Processing-time matrices are m x n, machine-major int32
 * (the stream layout of bounds/src/instance_flowshop.cpp:166-170).  One
 * engine handle belongs to one host thread at a time; pbb_interrupt may be
 * called from any thread.
 */
#ifndef PBB_ENGINE_H
#define PBB_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PBB_OK 0
#define PBB_E_ARG -1         /* invalid argument / unsupported option value */
#define PBB_E_HIP -2         /* HIP runtime error */
#define PBB_E_STATE -3       /* call not valid in the engine's current state */
#define PBB_E_CAPACITY -4    /* buffer / device capacity exceeded */
#define PBB_E_NODEVICE -5    /* no usable gfx950 device */
#define PBB_E_BUSY -6        /* serving: no free slot or HBM budget for this request now */

#define PBB_MAX_JOBS 63      /* n <= 63: IVM row cells and stack entries 0..n fit one 64-lane wavefront */
#define PBB_MAX_MACHINES 20  /* m <= 20 (the reference GPU cap, gpu_fsp_bound.cuh:12) */

/* ------------------------------------------------------------------ config */

/* Mirrors the [bb]/[gpu]/[initial] keys read by arguments::readIniFile. */
typedef struct pbb_config {
    int bound_mode;         /* [bb] boundingMode: 0 LB1, 1 strong only, 2 LB1 + refinement */
    int branching;          /* [bb] adaptiveBranchingMode, CPU numbering:
                               -1 front, 1 MaxSum, 2 MinBranch, 3 MinMin
                               (ivm_bound.cpp:413-507) */
    int refine;             /* refinement bound of modes 1/2: 0 = LB1' as shipped
                               (ivm_bound.cpp:170), 1 = Johnson LB2 (ivm_bound.cpp:171) */
    int johnson_pairs;      /* [bb] JohnsonMode: 0 all pairs, 1 (i,m), 2 (i,i+1) */
    int early_stop_johnson; /* [bb] earlyStopJohnson: accepted where it cannot change
                               the search (boundingMode 0, 2, or 1 with refine 0);
                               PBB_E_ARG with boundingMode 1 and refine 1 */
    int find_all;           /* [bb] findAll: prune with > instead of >= */
    int single_node;        /* [bb] singleNode: MinBranch/MinMin use the live UB (1)
                               or the UB at engine creation (0) */
    int sorted_dfs;         /* [bb] sortedDFS: must be 0 */
    int nb_ivms;            /* [gpu] nbIVMs: requested explorers (clamped to the
                               device's resident-wavefront capacity) */
    int steal_margin;       /* split intervals only at lines < n - steal_margin */
    int init_mode;          /* [initial] ub: 0 from file, 1 heuristic (NEH), 3 infinity */
    int init_solution;      /* [initial] sol: 0 identity root order, 1 NEH order */
    int checkpoint_s;       /* [time] checkpoint */
    int balance_s;          /* [time] balance: pbb_run returns after this period
                               in worker mode */
    int ramp_nodes;         /* an explorer donates to idle explorers only after
                               decomposing this many nodes of its current
                               interval (0: at once); engine-specific, no
                               reference counterpart */
    int team_nodes;         /* a search from pbb_init_full first runs on one
                               workgroup of explorers (16 with LB1 only, 8 with
                               a refinement bound) for up to this many nodes,
                               then on the whole device (0: whole device at
                               once); engine-specific, no reference counterpart */
    int team_waves;         /* explorers of that workgroup (0: the largest the
                               kernel allows, 16 or 8; larger values are
                               clamped to it); engine-specific, no reference
                               counterpart */
} pbb_config;

/* Defaults of arguments.cpp:135-175 with sortedDFS forced to 0. */
int pbb_config_default(pbb_config *cfg);

/* Parse a gpuconfig.ini/bbconfig.ini file (same sections and keys as the
 * reference).  problem/instance receive [problem] problem/instance (buffers of
 * at least 64 bytes, may be NULL).  Unsupported values (sortedDFS != 0,
 * JohnsonMode 3, earlyStopJohnson with boundingMode 1 and the Johnson
 * refinement, a negative steal margin) return PBB_E_ARG. */
int pbb_config_load_ini(const char *path, pbb_config *cfg, char *problem, char *instance);

/* ------------------------------------------------------------- instances */

/* Taillard instance ta<id> (bounds/src/instance_flowshop.cpp:146-181).
 * ptm must hold m*n ints (at most 20*500). */
int pbb_instance_taillard(int id, int *n, int *m, int32_t *ptm);

/* Seeded synthetic instance: SplitMix64 stream, machine-major draws,
 * p = 1 + (u mod 99) with rejection of the biased tail. */
int pbb_instance_synth(uint64_t seed, int n, int m, int32_t *ptm);

/* Shape draw for instance streams: n ~ U[nlo,nhi], m from mchoices. */
int pbb_instance_synth_shape(uint64_t seed, int nlo, int nhi, const int *mchoices, int nchoices,
                             int *n, int *m);

/* VRF text instance (bounds/src/instance_vrf.cpp:31-69): "n m" then per job
 * (machine, time) pairs.  Writes the machine-major PTM. */
int pbb_instance_read_vrf(const char *path, int *n, int *m, int32_t *ptm, int cap);

/* NEH heuristic (heuristic/src/fastNEH.cpp:151-243): perm in = start order,
 * out = NEH order; returns the makespan in *cost. */
int pbb_neh(const int32_t *ptm, int n, int m, int *perm, int *cost);

/* Makespan of one permutation on the host (bound_fsp_weak::evalSolution). */
int pbb_makespan_host(const int32_t *ptm, int n, int m, const int *perm, int *cost);

/* ---------------------------------------------------------- bound library */

/* Batched bound_abstract on device buffers (all pointers are device
 * pointers; `stream` is a hipStream_t or NULL for the default stream).
 * A partial schedule k is perms[k*n .. k*n+n-1] with front perms[0..l1[k]]
 * and back perms[l2[k]..n-1], -1 <= l1 < l2 <= n. */
typedef struct pbb_bounds pbb_bounds;

/* bound_fsp_weak::init + bound_fsp_strong::init for one instance. */
int pbb_bounds_create(const int32_t *ptm, int n, int m, pbb_bounds **out);
int pbb_bounds_destroy(pbb_bounds *b);

/* bound_fsp_weak::boundChildren (bounds/src/bound_fsp_weak.cpp:193-221):
 * job-indexed cb/ce (count*n ints, 0 for non-children). both=0 -> front only
 * (ce untouched). */
int pbb_lb1_children_batch(pbb_bounds *b, int count, const int *perms, const int *l1,
                           const int *l2, int both, int *cb, int *ce, void *stream);

/* bound_fsp_weak::bornes_calculer (bounds/src/bound_fsp_weak.cpp:288-311). */
int pbb_lb1_node_batch(pbb_bounds *b, int count, const int *perms, const int *l1, const int *l2,
                       int *lb, void *stream);

/* bound_fsp_strong::bornes_calculer with earlyExit=0, best=-1
 * (bounds/src/bound_fsp_strong.cpp:540-560); pairs = JohnsonMode 0/1/2. */
int pbb_lb2_node_batch(pbb_bounds *b, int count, const int *perms, const int *l1, const int *l2,
                       int pairs, int *lb, void *stream);

/* bound_fsp_weak::evalSolution over count full permutations. */
int pbb_makespan_batch(pbb_bounds *b, int count, const int *perms, int *cmax, void *stream);

/* ----------------------------------------------------------- the engine */

typedef struct pbb_engine pbb_engine;

typedef struct pbb_stats {
    uint64_t decomposed;     /* pbab stats.totDecomposed (sequentialbb.cpp:129) */
    uint64_t leaves;         /* stats.leaves (ivm_bound.cpp:243) */
    uint64_t simple_bounds;  /* stats.simpleBounds (ivm_bound.cpp:225-229) */
    uint64_t johnson_bounds; /* stats.johnsonBounds (ivm_bound.cpp:178) */
    int explorers;           /* explorers actually running (resident wavefronts) */
    int active;              /* explorers holding an interval after the last run */
    double kernel_ms;        /* device time of the last pbb_run */
    uint64_t depth_histo[PBB_MAX_JOBS + 2]; /* decomposed nodes per depth */
    uint64_t refine_histo[PBB_MAX_JOBS + 2]; /* refinement bounds per depth of the bounded node */
} pbb_stats;

/* gpubb(pbab*) + initialize + initializeBoundFSP + copyH2D
 * (gpu/src/gpubb.cu:24-91, :974-1038).  ub is the incumbent cost (INT_MAX
 * for none; negative is PBB_E_ARG), ub_perm its permutation (may be NULL;
 * PBB_E_ARG when not a permutation of 0..n-1).  The engine sizes its grid
 * for the whole device; explorers only receive work once their wavefront
 * runs, so another kernel on the device delays the search but cannot stall
 * it. */
int pbb_create(const int32_t *ptm, int n, int m, const pbb_config *cfg, int ub,
               const int *ub_perm, pbb_engine **out);
int pbb_destroy(pbb_engine *e);

/* Root order (root_sltn->perm, ivm_bound.cpp:288-293); identity by default. */
int pbb_set_root(pbb_engine *e, const int *root_perm);

/* gpubb::initFullInterval (gpu/src/gpubb.cu:95-154): bound the root with LB1
 * and give the whole interval [0, n!-1] to explorer 0. Resets statistics. */
int pbb_init_full(pbb_engine *e);

/* gpubb::initFromFac (gpu/src/gpubb.cu:1244-1310): nb closed factoradic
 * intervals pos[k*n..], end[k*n..] (digit i in [0, n-1-i]) for explorers
 * ids[k] < explorers.  Each is unfolded (bounded, uncounted) on device. */
int pbb_init_intervals(pbb_engine *e, int nb, const int *ids, const int *pos, const int *end);

/* gpubb::next (gpu/src/gpubb.cu:272-306): explore until every explorer is
 * idle (*all_done = 1), or until node_budget decomposed nodes (0 = none) /
 * time_budget_s seconds (0 = none) / pbb_interrupt. */
int pbb_run(pbb_engine *e, uint64_t node_budget, double time_budget_s, int *all_done);

/* gpubb::getIntervals (gpu/src/gpubb.cu:1189-1241): export the interval of
 * every active explorer; max must be >= explorers.  Node/leaf counters are
 * returned in *decomposed / *leaves (may be NULL) and reset, as the
 * reference moves them into the worker's work unit. */
int pbb_get_intervals(pbb_engine *e, int *pos, int *end, int *ids, int *nb, int max,
                      uint64_t *decomposed, uint64_t *leaves);

/* Incumbent in/out (pbb->sltn, common/src/solution.cpp:26-88). set only
 * lowers the cost (strict <), as solution::update. perm may be NULL. */
int pbb_get_incumbent(pbb_engine *e, int *cost, int *perm);
int pbb_set_incumbent(pbb_engine *e, int cost, const int *perm);

int pbb_get_stats(pbb_engine *e, pbb_stats *st);

/* gpubb::getDeepSubproblem (gpu/src/gpubb.cu:1313-1371): the current
 * schedule of every active explorer (front jobs, the current row in order,
 * back jobs: ivm_bound::prepareSchedule, multicore/src/ivm_bound.cpp:111-138),
 * lowest makespan first (ties by explorer), at most max of them, n ints each
 * in perms; *count receives how many.  Used by the worker to seed its
 * heuristic (distributed/src/worker_gpu.cpp:93-110). */
int pbb_get_deep_subproblems(pbb_engine *e, int *perms, int max, int *count);

/* gpubb::interruptExploration (gpu/src/gpubb.cu:156-162); honoured by the
 * running kernel within a few thousand nodes. */
int pbb_interrupt(pbb_engine *e);

/* ------------------------------------------------ host co-worker
 * IVM explorers on host threads (the reference's multicore explorers,
 * multicore/src/matrix_controller.cpp and bbthread.cpp, as a worker of
 * distributed/src/worker_mc.cpp:26-111) on the same work units as the
 * engine: pbb_cpu_init_intervals / pbb_cpu_get_intervals take and return
 * the factoradic intervals that pbb_get_intervals / pbb_init_intervals
 * exchange, so idle host cores explore intervals moved off the GPU.  Every
 * node is processed with the engine's semantics (bounds, branching, pruning,
 * counters): with a frozen incumbent an interval's node counts are the same
 * on either side.  Threads share work among themselves as bbthread::shareWork
 * does. */
typedef struct pbb_cpu pbb_cpu;

/* threads = 0: one per hardware thread.  ub / ub_perm as pbb_create. */
int pbb_cpu_create(const int32_t *ptm, int n, int m, const pbb_config *cfg, int threads, int ub,
                   const int *ub_perm, pbb_cpu **out);
int pbb_cpu_destroy(pbb_cpu *c);
int pbb_cpu_set_root(pbb_cpu *c, const int *root_perm);
/* matrix_controller::initFromFac (multicore/src/matrix_controller.cpp:112-140):
 * interval k to thread ids[k] < threads, unfolded (bounded, uncounted). */
int pbb_cpu_init_intervals(pbb_cpu *c, int nb, const int *ids, const int *pos, const int *end);
int pbb_cpu_init_full(pbb_cpu *c);
/* explore until every thread is dry (*all_done = 1), node_budget decomposed
 * nodes, time_budget_s seconds (0 = none) or pbb_cpu_interrupt */
int pbb_cpu_run(pbb_cpu *c, uint64_t node_budget, double time_budget_s, int *all_done);
/* matrix_controller::getIntervals (:142-170); counters as pbb_get_intervals */
int pbb_cpu_get_intervals(pbb_cpu *c, int *pos, int *end, int *ids, int *nb, int max, uint64_t *decomposed,
                          uint64_t *leaves);
int pbb_cpu_get_incumbent(pbb_cpu *c, int *cost, int *perm);
int pbb_cpu_set_incumbent(pbb_cpu *c, int cost, const int *perm);
int pbb_cpu_get_stats(pbb_cpu *c, pbb_stats *st);
int pbb_cpu_interrupt(pbb_cpu *c);

/* The engine and the host co-worker on one search (no counterpart in the
 * reference, which runs them as separate workers of its master): periods of
 * period_s seconds run both at once; whenever the host threads are dry, a
 * quarter of the engine's intervals (at most one per thread) moves to them;
 * incumbents are exchanged after every period.  Ends when both are done or
 * after time_budget_s (0 = none).  The decomposed nodes of each side since
 * the call go to *gpu_decomposed and *cpu_decomposed (may be NULL). */
int pbb_hybrid_run(pbb_engine *e, pbb_cpu *c, double period_s, double time_budget_s, uint64_t *gpu_decomposed,
                   uint64_t *cpu_decomposed, int *all_done);

/* ------------------------------------------------ work-unit codec
 * Factoradic digit vectors (n digits, digit i in [0, n-1-i], the fact_work
 * pos/end layout) <-> decimal interval ranks sum_i digit_i * (n-1-i)!
 * (fact_work::VectToBigint / BigintToVect, distributed/src/fact_work.cpp:47-86,
 * weights in distributed/src/weights.cpp:10-26; fixed 384-bit arithmetic in
 * place of GMP, ranks up to 63! - 1). */
int pbb_fact_to_dec(int n, const int *digits, char *dec, int cap);
int pbb_dec_to_fact(int n, const char *dec, int *digits);

/* The checkpoint text of works::save (distributed/src/works.cpp:172-191):
 * root solution, simpleBounds, incumbent (solution operator<<,
 * common/src/solution.cpp:132-142), then the work units (works operator<<,
 * distributed/src/works.cpp:558-569; work::writeHeader/writeIntervals,
 * distributed/src/work.cpp:544-583: "id end_updated max_intervals
 * exploredNodes nb" then "id begin end" per interval, then the total size). */
typedef struct pbb_checkpoint {
    int n;
    int root_cost;
    int root_perm[PBB_MAX_JOBS];
    uint64_t simple_bounds;
    int best_cost;
    int best_perm[PBB_MAX_JOBS];
    int work_id;
    int max_intervals;
    uint64_t explored;      /* exploredNodes (sum over units when parsing) */
    int nb;                 /* intervals */
} pbb_checkpoint;

/* Text of one work unit holding intervals k < c->nb (ids, pos/end digit
 * vectors); *len = text length; buf may be NULL to query it. */
int pbb_checkpoint_format(const pbb_checkpoint *c, const int *ids, const int *pos, const int *end,
                          char *buf, int cap, int *len);
/* Parse a checkpoint text for n-job instances: every unit's non-empty
 * intervals (begin <= end) into ids/pos/end (at most max). */
int pbb_checkpoint_parse(const char *text, int n, pbb_checkpoint *c, int *ids, int *pos, int *end, int max);
/* works::save of an engine: its active intervals, incumbent, root order,
 * decomposed nodes and simple bounds, written to path. */
int pbb_checkpoint_save(pbb_engine *e, const char *path);
/* works::init_infile for an engine: root order, incumbent and intervals
 * from path (interval i goes to explorer i); info (may be NULL) receives
 * the file's counters. */
int pbb_checkpoint_load(pbb_engine *e, const char *path, pbb_checkpoint *info);

/* Explorer capacity the engine would use for (n, m, cfg) on this device. */
int pbb_query_explorers(int n, int m, const pbb_config *cfg, int *explorers);

/* --------------------------------------------- multi-instance pool
 * Many instances in one resident launch: batched solves (all instances in
 * one HIP-graph launch) and a serving loop (requests admitted while others
 * run).  There is no reference counterpart: the reference solves one
 * instance per process (its strong-bound tables are process-global,
 * bounds/src/bound_fsp_strong.cpp:5-10); each instance is solved with the
 * same semantics as pbb_create + pbb_init_full + pbb_run with the root order
 * the identity (initial.sol=0).  Instances with m < machines are padded
 * with zero-time machines, which leaves every bound and makespan unchanged. */
typedef struct pbb_pool pbb_pool;

#define PBB_RES_PENDING 0
#define PBB_RES_SOLVED 1     /* explored to the end: best is the optimum (or no schedule beats ub) */
#define PBB_RES_ABORTED 2    /* time budget or interrupt before the end */
#define PBB_RES_INVALID 3    /* malformed request */

typedef struct pbb_pool_params {
    int max_jobs;               /* largest n accepted (3..63) */
    int machines;               /* M: instances with 2 <= m <= M are accepted */
    int slots;                  /* instance slots (instances in flight) */
    int waves_per_group;        /* explorers per workgroup: 1, 2 or 4 */
    int explorers_per_instance; /* cap on the explorers working one instance */
    int steal_margin;           /* split only at lines < n - margin; 0 = the config's */
    int groups;                 /* workgroups launched; 0 = every resident one */
    double hbm_budget_gb;       /* serving admission budget; 0 = device memory free at creation */
    double request_budget_s;    /* per-instance time after admission (0 = none): an instance
                                   still running then is stopped and retires PBB_RES_ABORTED,
                                   with the best schedule found so far */
    int ig_iterations;          /* 0 = none; else the last wave of each instance's admitting
                                   workgroup runs one iterated-greedy chain of this many
                                   rounds beside the search, publishing every improvement as
                                   the instance's incumbent (needs waves_per_group >= 2);
                                   engine-specific, no reference counterpart */
} pbb_pool_params;

typedef struct pbb_result {
    int tag;                    /* caller's tag */
    int status;                 /* PBB_RES_* */
    int best;                   /* optimum, or ub when nothing below ub exists */
    int improved;               /* 1: perm holds a schedule of cost best */
    uint64_t decomposed, leaves, simple_bounds, johnson_bounds;
    double admit_ms, done_ms;   /* device clock, ms after the launch started */
    int perm[PBB_MAX_JOBS];
} pbb_result;

int pbb_pool_params_default(pbb_pool_params *prm);
int pbb_pool_create(const pbb_config *cfg, const pbb_pool_params *prm, pbb_pool **out);
int pbb_pool_destroy(pbb_pool *p);
/* Explorers and workgroups the pool launches. */
int pbb_pool_info(pbb_pool *p, int *explorers, int *groups, int *lds_bytes);

/* Change the per-instance explorer cap and the steal margin of an idle pool
 * (the cap stays within the idle-ring capacity chosen at creation). */
int pbb_pool_retune(pbb_pool *p, int explorers_per_instance, int steal_margin);

/* Depth histograms of the last pbb_pool_solve batch, instance i at
 * [64 i .. 64 i + 63]: decomposed[.. + d] = nodes decomposed with d jobs
 * fixed (leaves included, at d = n - 1), refined[.. + d] = refinement bounds
 * (LB1' or Johnson) computed for the children of such nodes.  For roofline
 * accounting: the work of a node and of a refinement depends on d.  count
 * must not exceed the last batch's count; PBB_E_STATE while serving. */
int pbb_pool_batch_histograms(pbb_pool *p, int count, uint64_t *decomposed, uint64_t *refined);

/* A request's ub may be PBB_UB_NEH: the admitting workgroup then computes
 * the incumbent with NEH on the device (pbb_neh_device semantics) and the
 * result's schedule is the NEH order unless the search improves on it. */
#define PBB_UB_NEH -1

/* NEH for count instances on the device, one wavefront each (the
 * pbb_neh order with identity start; instance i's times at ptms + off_i as
 * in pbb_pool_solve).  perms: count x 64 ints, instance i's order at
 * perms[64 i .. 64 i + n[i] - 1]; costs: count makespans. */
int pbb_neh_device(int count, const int32_t *ptms, const int *n, const int *m, int *perms, int *costs);

/* pbb_neh_device on buffers already in device memory, allocating nothing.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  d_ptms,
 * d_n, d_m, d_perms (count x 64 ints) and d_costs are device pointers laid
 * out as in pbb_neh_device; max_m (2..20) bounds every m[i] and picks the
 * kernel.  Asynchronous on `stream` (a hipStream_t, NULL = default stream).
 * The instances' offsets go to a fixed device array of PBB_NEH_NOALLOC_MAX
 * entries (256 KB per device, held until the process ends) shared by every
 * call on the device: count must not exceed PBB_NEH_NOALLOC_MAX
 * (PBB_E_CAPACITY), and calls that overlap on one device race.  An instance
 * with n outside [3, 63] or m outside [2, max_m] gets cost -1 and order -1;
 * its n*m times still count in the offsets of the later instances when
 * 1 <= n <= 63 and 1 <= m <= 20 (other sizes count as 0). */
#define PBB_NEH_NOALLOC_MAX 32768
int pbb_neh_device_noalloc(int count, const int32_t *d_ptms, const int *d_n, const int *d_m, int max_m,
                           int *d_perms, int *d_costs, void *stream);

/* Iterated greedy on the device (heuristic/src/IG.cpp IG::runIG with the
 * k-insertion local search of heuristic/src/fspnhoods.cpp:46-100, run as the
 * reference's initial heuristic runs it: independent chains from NEH, the
 * best kept, common/src/pbab.cpp:124-188).  For each of count instances
 * (times laid out as in pbb_neh_device, 3 <= n <= 63, 2 <= m <= 20) run
 * `chains` chains of `iterations` rounds each, one wavefront per chain.
 * Chain c of instance i draws from a SplitMix64 stream seeded with
 * seed ^ (i << 32 | c) (the reference seeds from std::random_device), and a
 * worse solution is accepted with probability exp(-delta / T), T = 0.2 x
 * the mean processing time / 10, through integer thresholds computed here
 * (deltas >= 256 are refused).  perms: count x 64 ints (instance i's best
 * order at perms[64 i ..]); costs: count best makespans, never above NEH's;
 * chain_costs / chain_perms (may be NULL): count x chains costs and
 * count x chains x 64 orders of every chain. */
int pbb_ig_device(int count, const int32_t *ptms, const int *n, const int *m, int chains, int iterations,
                  uint64_t seed, int *perms, int *costs, int *chain_costs, int *chain_perms);

/* Batch: count <= slots instances, instance i's m[i] x n[i] machine-major
 * times at ptms[off_i], off_i = sum of n[k]*m[k] for k < i.  Upload, reset,
 * the resident launch and the result download are one graph launch.
 * tags may be NULL (tag = index). kernel_ms (may be NULL) = graph time. */
int pbb_pool_solve(pbb_pool *p, int count, const int32_t *ptms, const int *n, const int *m,
                   const int *ub, const int *tags, double time_budget_s, pbb_result *out,
                   double *kernel_ms);

/* Serving: start the resident launch (ends after time_budget_s at most). */
int pbb_serve_start(pbb_pool *p, double time_budget_s);
/* Admit one request if a slot and the HBM budget allow, else PBB_E_BUSY. */
int pbb_serve_submit(pbb_pool *p, const int32_t *ptm, int n, int m, int ub, int tag);
/* Completed requests since the last poll (at most max). */
int pbb_serve_poll(pbb_pool *p, pbb_result *out, int max, int *count);
/* Close the queue, let admitted requests finish, wait for the launch. */
int pbb_serve_stop(pbb_pool *p);
/* Admission footprint of an n-job request: explorers_per_instance x
 * (n*n + 2n) int32 (the IVM of one explorer) plus the slot's tables. */
int pbb_serve_footprint(pbb_pool *p, int n, uint64_t *bytes);

/* Replay an arrival stream through the serving loop: request i (as in
 * pbb_pool_solve) arrives at arrival_s[i] seconds; latency_ms[i] = completion
 * seen by the host - arrival, or -1 for a request not completed within
 * time_budget_s (out[i].status then says whether it was solved while the
 * pool stopped, aborted, or never admitted).  *wall_s = first arrival to last
 * completion within the budget (0 when none completed). */
int pbb_serve_replay(pbb_pool *p, int count, const int32_t *ptms, const int *n, const int *m,
                     const int *ub, const double *arrival_s, double time_budget_s, pbb_result *out,
                     double *latency_ms, double *wall_s);
/* The same stream through several pools resident on the device together
 * (no counterpart in the reference): request i goes to the pool with the
 * fewest machines whose shape holds it (n <= max_jobs, m <= machines; ties go
 * to the smaller max_jobs), so a 5- or 10-machine request is not padded to
 * the largest machine count.  1 to 8 distinct pools; a request that no pool
 * holds fails the call with PBB_E_ARG before anything is submitted.  The
 * serving kernels start in decreasing order of machines, and each stays
 * resident until the replay ends, so a later pool only gets the registers
 * and LDS the earlier ones leave: give every pool an explicit `groups` small
 * enough that each later pool keeps room for its groups (a group that is
 * not resident only delays, but a pool with none serves nothing until the
 * time budget ends the replay).  Measured on one MI355X: 192 groups of a
 * 20-machine pool started before 512 groups of a 10-machine pool. */
int pbb_serve_replay_routed(pbb_pool *const *pools, int npools, int count, const int32_t *ptms, const int *n,
                            const int *m, const int *ub, const double *arrival_s, double time_budget_s,
                            pbb_result *out, double *latency_ms, double *wall_s);
/* The same stream on the one-instance engine, one launch per instance in
 * arrival order (first come, first served).  ub[i] may be PBB_UB_NEH (the
 * host NEH's incumbent, computed after the request arrives).
 * request_budget_s (0 = none) caps each instance's launch as the pool's
 * request_budget_s does: an instance still running then retires
 * PBB_RES_ABORTED with the best schedule found so far. */
int pbb_plain_replay(const pbb_config *cfg, int count, const int32_t *ptms, const int *n,
                     const int *m, const int *ub, const double *arrival_s, double time_budget_s,
                     double request_budget_s, pbb_result *out, double *latency_ms, double *wall_s);

/* Autotuner: time pbb_pool_solve on the given instances for every
 * (waves_per_group in {1,2,4}, explorers_per_instance in {64,256,1024},
 * steal_margin in {2,4,8}) candidate (each trial capped at twice the best
 * time so far plus 0.5 s; capped trials count as failed) and
 * write the fastest into *best (other fields taken from *base).  trials
 * (may be NULL) receives up to max_trials rows {W, epi, margin, ms}. */
int pbb_pool_autotune(const pbb_config *cfg, const pbb_pool_params *base, int count,
                      const int32_t *ptms, const int *n, const int *m, const int *ub,
                      pbb_pool_params *best, double *trials, int max_trials, int *ntrials);

const char *pbb_last_error(void);
const char *pbb_version(void);

#ifdef __cplusplus
}
#endif
#endif
