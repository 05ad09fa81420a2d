/*
 * pbb_oracle.h -- CPU restatement of the reference IVM branch-and-bound path
 * for the permutation flow-shop (makespan).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP engine is compared
 * against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product library never links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the jangmys/PBBPerm tree).  Parity status of each
 * function is recorded beside it:
 *   pinned:<how>   checked against reference vectors / closed forms / brute force
 */
#ifndef PBB_ORACLE_H
#define PBB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXN 128
#define ORC_MAXM 64

/* ---------------- instance generators ---------------- */

/* Taillard's Park-Miller/Schrage generator, bounds/src/instance_flowshop.cpp:146-171.
 * Writes the m x n machine-major PTM of ta<id>; returns n (m in *m_out).
 * pinned: ta001-ta020 PTMs from the compiled reference + optima in parameters/instances.data */
int orc_taillard(int id, int *m_out, int32_t *ptm);

/* Project generator: SplitMix64 stream, machine-major draws, p = 1 + (u mod 99)
 * with rejection of the biased tail.  Not in the reference (its only generator
 * maps seed 0 to all-ones); defined here so instances are reproducible. */
void orc_synth_ptm(uint64_t seed, int n, int m, int32_t *ptm);
/* Batched-stream shape draw: n ~ U[nlo,nhi], m from mchoices, then the PTM,
 * all from the same SplitMix64 stream of `seed`. */
void orc_synth_shape(uint64_t seed, int nlo, int nhi, const int *mchoices, int nchoices,
                     int *n_out, int *m_out);
uint64_t orc_splitmix_next(uint64_t *state);
uint64_t orc_uniform(uint64_t *state, uint64_t k);

/* ---------------- instance tables ---------------- */

typedef struct {
    int n, m;
    int32_t ptm[ORC_MAXM * ORC_MAXN];      /* [m][n] machine-major              */
    int32_t min_arr[ORC_MAXM];             /* bound_fsp_weak.cpp:270-283        */
    int32_t min_dep[ORC_MAXM];             /* bound_fsp_weak.cpp:253-267        */
    int npairs;                            /* m(m-1)/2, bound_fsp_strong.cpp:241-246 */
    int32_t pair_a[ORC_MAXM * (ORC_MAXM - 1) / 2];
    int32_t pair_b[ORC_MAXM * (ORC_MAXM - 1) / 2];
    int32_t *lag;                          /* [npairs][n] bound_fsp_strong.cpp:264-280 */
    int32_t *johnson;                      /* [npairs][n] job order, :342-436   */
} orc_inst;

orc_inst *orc_inst_new(int n, int m, const int32_t *ptm);
void orc_inst_free(orc_inst *I);

/* ---------------- bounds ---------------- */

/* Makespan of a full permutation, bound_fsp_weak.cpp:223-244.  pinned: reference vectors */
int orc_makespan(const orc_inst *I, const int *perm);

/* LB1 of every child, both directions (branching > 0) or front only,
 * bound_fsp_weak.cpp:193-221 (+ :61-191).  Job-indexed outputs, zero for
 * non-children.  prio arrays (sum of |idle|) are also produced.
 * pinned: reference vectors + brute-force validity (LB <= best completion). */
void orc_lb1_children(const orc_inst *I, const int *sched, int l1, int l2, int both,
                      int *cb, int *ce, int *pb, int *pe);

/* Node bound LB1' (machine-chained), bound_fsp_weak.cpp:288-311.  pinned: reference vectors */
int orc_lb1_node(const orc_inst *I, const int *perm, int l1, int l2);

/* Johnson two-machine bound LB2, bound_fsp_strong.cpp:540-560 with
 * earlyExit=0 / best=-1 (machine pairs mode 0,1,2).  *idle receives couts[1].
 * pinned: reference vectors + brute force of each pair's two-machine optimum. */
int orc_lb2_node(const orc_inst *I, const int *perm, int l1, int l2, int pairs_mode, int *idle);

/* NEH: fastNEH::initialSort + runNEH, heuristic/src/fastNEH.cpp:151-243.
 * perm in: starting order (identity in the reference usage); out: NEH order. */
int orc_neh(const orc_inst *I, int *perm);

/* ---------------- IVM single-explorer B&B ---------------- */

typedef struct {
    int bound_mode;     /* 0 LB1, 1 strong-only, 2 LB1 + refinement (arguments.cpp:230) */
    int branching;      /* CPU numbering: -3,-2,-1,1 MaxSum,2 MinBranch,3 MinMin */
    int refine;         /* 0: LB1' as shipped (ivm_bound.cpp:170); 1: Johnson LB2 (:171) */
    int johnson_pairs;  /* JohnsonMode 0,1,2 */
    int find_all;
    int single_node;    /* 1: MinBranch/MinMin use the live UB, 0: the initial UB */
} orc_cfg;

typedef struct {
    uint64_t decomposed, leaves, simple_bounds, johnson_bounds;
} orc_stats;

typedef struct orc_bb orc_bb;

orc_bb *orc_bb_new(const orc_inst *I, const orc_cfg *cfg, int ub, const int *ub_perm);
void orc_bb_free(orc_bb *B);
/* sequentialbb::setRoot + ivm_bound::boundRoot, sequentialbb.cpp:27-45, ivm_bound.cpp:273-330 */
void orc_bb_set_root(orc_bb *B, const int *root_perm);
/* sequentialbb::initAtInterval + unfold, sequentialbb.cpp:83-108, :206-264 */
int orc_bb_init_interval(orc_bb *B, const int *pos, const int *end);
int orc_bb_init_full(orc_bb *B);
/* sequentialbb::next, sequentialbb.cpp:110-191: returns 1 if a node was decomposed */
int orc_bb_next(orc_bb *B);
/* Run next() until done or `max_nodes` decomposed nodes (0 = unbounded). Returns 1 if finished. */
int orc_bb_run(orc_bb *B, uint64_t max_nodes);
void orc_bb_stats(const orc_bb *B, orc_stats *st);
int orc_bb_best(const orc_bb *B, int *perm);
/* Raw IVM access (interval export, stealing tests) */
int orc_bb_line(const orc_bb *B);
void orc_bb_interval(const orc_bb *B, int *pos, int *end);
/* bbthread::shareWork (multicore/src/bbthread.cpp:39-72) from B (victim) to T (thief). */
int orc_bb_share(orc_bb *victim, orc_bb *thief, int min_depth_margin);
/* Per-depth histogram of decomposed nodes (depth = line after goDown). */
void orc_bb_depth_histo(const orc_bb *B, uint64_t *histo);

/* ---------------- iterated greedy ---------------- */

/* fastInsertRemove::insertMakespans / removeMakespans
 * (heuristic/src/fastinsertremove.cpp:196-279): makespans of seq[0..len-1]
 * with job inserted before each position 0..len, and without each position.
 * pinned: the reference's own functions (oracle/_ref), tests/golden/ig_moves.json */
void orc_insert_makespans(const orc_inst *I, const int *seq, int len, int job, int *out);
void orc_remove_makespans(const orc_inst *I, const int *seq, int len, int *out);
/* IG::IG visitOrder, heuristic/src/IG.cpp:38-52.  pinned: the reference's
 * order, printed by oracle/_ref */
void orc_ig_visit_order(int n, int *vo);
/* The IG's generator: uniform in [0, k) from the SplitMix64 stream, and the
 * stream state of chain `chain` of instance `inst`. */
uint32_t orc_ig_below(uint64_t *state, uint32_t k);
uint64_t orc_ig_chain_seed(uint64_t seed, int inst, int chain);
/* Acceptance thresholds of IG::acceptance (IG.cpp:319-338) for cost deltas
 * 0..ORC_IG_THR-1 (see pbb_oracle.c). */
#define ORC_IG_THR 256
void orc_ig_thresholds(const orc_inst *I, uint64_t *thr);
/* IG::runIG (heuristic/src/IG.cpp:341-404) from the NEH order, `iters`
 * rounds, driven by the SplitMix64 stream `state` (the reference's
 * generators are seeded from std::random_device).  Returns the best
 * makespan; perm receives its order.  pinned: the moves above against the
 * reference + properties (<= NEH, the order realises the cost, optimum on
 * small instances by brute force); the seeded steps themselves have no
 * reference run to match (parity unpinned beyond that). */
int orc_ig(const orc_inst *I, uint64_t state, int iters, const uint64_t *thr, int *perm);

/* Batched form of the bound functions above (NULL outputs are skipped). */
void orc_bounds_batch(const orc_inst *I, int count, const int *perms, const int *l1, const int *l2,
                      int pairs_mode, int *cb, int *ce, int *cbf, int *lb1n, int *lb2, int *mk);

#ifdef __cplusplus
}
#endif
#endif
