/*
 * pbb_oracle.c -- CPU restatement of the reference IVM branch-and-bound path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py.  The product library never links it.
 *
 * Reference files restated (paths relative to the jangmys/PBBPerm tree):
 *   bounds/src/bound_fsp_weak.cpp     LB1 children, LB1', makespan
 *   bounds/src/bound_fsp_strong.cpp   Johnson LB2
 *   bounds/src/instance_flowshop.cpp  Taillard generator
 *   heuristic/src/fastNEH.cpp         NEH
 *   multicore/src/ivm.cpp             IVM navigation
 *   multicore/src/ivm_bound.cpp       decode / bound / branch / prune / leaf
 *   multicore/src/sequentialbb.cpp    driver loop, interval init, unfold
 *   multicore/src/bbthread.cpp        interval split (work sharing)
 */
#include "pbb_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int imax(int a, int b) { return a > b ? a : b; }

/* ======================================================================
 * Generators
 * ====================================================================== */

/* bounds/src/instance_flowshop.cpp:8-32 (seeds), :104-144 (shapes) */
static const long ta_seeds[120] = {
    873654221, 379008056, 1866992158, 216771124, 495070989, 402959317, 1369363414, 2021925980,
    573109518, 88325120, 587595453, 1401007982, 873136276, 268827376, 1634173168, 691823909,
    73807235, 1273398721, 2065119309, 1672900551, 479340445, 268827376, 1958948863, 918272953,
    555010963, 2010851491, 1519833303, 1748670931, 1923497586, 1829909967, 1328042058, 200382020,
    496319842, 1203030903, 1730708564, 450926852, 1303135678, 1273398721, 587288402, 248421594,
    1958948863, 575633267, 655816003, 1977864101, 93805469, 1803345551, 49612559, 1899802599,
    2013025619, 578962478, 1539989115, 691823909, 655816003, 1315102446, 1949668355, 1923497586,
    1805594913, 1861070898, 715643788, 464843328, 896678084, 1179439976, 1122278347, 416756875,
    267829958, 1835213917, 1328833962, 1418570761, 161033112, 304212574, 1539989115, 655816003,
    960914243, 1915696806, 2013025619, 1168140026, 1923497586, 167698528, 1528387973, 993794175,
    450926852, 1462772409, 1021685265, 83696007, 508154254, 1861070898, 26482542, 444956424,
    2115448041, 118254244, 471503978, 1215892992, 135346136, 1602504050, 160037322, 551454346,
    519485142, 383947510, 1968171878, 540872513, 2013025619, 475051709, 914834335, 810642687,
    1019331795, 2056065863, 1342855162, 1325809384, 1988803007, 765656702, 1368624604, 450181436,
    1927888393, 1759567256, 606425239, 19268348, 1298201670, 2041736264, 379756761, 28837162};

static int ta_jobs(int id)
{
    if (id > 110) return 500;
    if (id > 90) return 200;
    if (id > 60) return 100;
    if (id > 30) return 50;
    return 20;
}

static int ta_machines(int id)
{
    static const int lim[] = {110, 100, 90, 80, 70, 60, 50, 40, 30, 20, 10};
    static const int val[] = {20, 20, 10, 20, 10, 5, 20, 10, 5, 20, 10};
    for (int i = 0; i < 11; i++)
        if (id > lim[i]) return val[i];
    return 5;
}

/* Park-Miller minimal standard step via Schrage's decomposition, and the
 * single-precision scaling of instance_flowshop.cpp:146-157. */
static long ta_unif(long *seed, long low, long high)
{
    const long M = 2147483647, A = 16807, Q = 127773, R = 2836;
    long k = *seed / Q;
    *seed = A * (*seed % Q) - k * R;
    if (*seed < 0) *seed += M;
    float ratio = (float)*seed / (float)M;
    double v = ratio;
    return low + (long)(v * (double)(high - low + 1));
}

int orc_taillard(int id, int *m_out, int32_t *ptm)
{
    int n = ta_jobs(id), m = ta_machines(id);
    long seed = ta_seeds[id - 1];
    for (int k = 0; k < m; k++)
        for (int j = 0; j < n; j++)
            ptm[k * n + j] = (int32_t)ta_unif(&seed, 1, 99);
    *m_out = m;
    return n;
}

uint64_t orc_splitmix_next(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t orc_uniform(uint64_t *state, uint64_t k)
{
    /* reject the top (2^64 mod k) values so every residue is equally likely */
    uint64_t limit = UINT64_MAX - (UINT64_MAX % k + 1) % k;
    uint64_t x;
    do {
        x = orc_splitmix_next(state);
    } while (x > limit);
    return x % k;
}

void orc_synth_ptm(uint64_t seed, int n, int m, int32_t *ptm)
{
    uint64_t s = seed;
    for (int k = 0; k < m; k++)
        for (int j = 0; j < n; j++)
            ptm[k * n + j] = 1 + (int32_t)orc_uniform(&s, 99);
}

void orc_synth_shape(uint64_t seed, int nlo, int nhi, const int *mchoices, int nchoices,
                     int *n_out, int *m_out)
{
    uint64_t s = seed ^ 0x5348415045ull; /* shape stream, distinct from the PTM stream */
    *n_out = nlo + (int)orc_uniform(&s, (uint64_t)(nhi - nlo + 1));
    *m_out = mchoices[orc_uniform(&s, (uint64_t)nchoices)];
}

/* ======================================================================
 * Instance tables
 * ====================================================================== */

/* Johnson/Mitten order key: set 1 (p_a < p_b) first, ascending p_a + lag;
 * set 2 descending p_b + lag (bound_fsp_strong.cpp:358-436).  Tie order
 * inside a key does not change the two-machine optimum, so a stable
 * insertion sort is used instead of the reference's Hoare quicksort. */
static int johnson_before(int set_i, int key_i, int set_j, int key_j)
{
    if (set_i != set_j) return set_i < set_j;
    if (set_i == 1) return key_i < key_j;
    return key_i > key_j;
}

orc_inst *orc_inst_new(int n, int m, const int32_t *ptm)
{
    if (n < 2 || n > ORC_MAXN || m < 2 || m > ORC_MAXM) return NULL;
    orc_inst *I = (orc_inst *)calloc(1, sizeof(orc_inst));
    I->n = n;
    I->m = m;
    memcpy(I->ptm, ptm, sizeof(int32_t) * n * m);

    /* bound_fsp_weak.cpp:248-286 */
    for (int k = 0; k < m; k++) I->min_dep[k] = INT_MAX;
    I->min_dep[m - 1] = 0;
    for (int k = 0; k < m; k++) I->min_arr[k] = INT_MAX;
    I->min_arr[0] = 0;
    for (int j = 0; j < n; j++) {
        int tail = 0; /* sum of p on machines k+1..m-1 */
        for (int k = m - 2; k >= 0; k--) {
            tail += ptm[(k + 1) * n + j];
            if (tail < I->min_dep[k]) I->min_dep[k] = tail;
        }
        int head = 0; /* sum of p on machines 0..k-1 */
        for (int k = 1; k < m; k++) {
            head += ptm[(k - 1) * n + j];
            if (head < I->min_arr[k]) I->min_arr[k] = head;
        }
    }

    /* machine pairs in (i<j) lexicographic order, bound_fsp_strong.cpp:282-297 */
    int P = m * (m - 1) / 2;
    I->npairs = P;
    int c = 0;
    for (int a = 0; a < m - 1; a++)
        for (int b = a + 1; b < m; b++) {
            I->pair_a[c] = a;
            I->pair_b[c] = b;
            c++;
        }
    I->lag = (int32_t *)calloc((size_t)P * n, sizeof(int32_t));
    I->johnson = (int32_t *)calloc((size_t)P * n, sizeof(int32_t));
    for (int p = 0; p < P; p++) {
        int a = I->pair_a[p], b = I->pair_b[p];
        int set[ORC_MAXN], key[ORC_MAXN];
        for (int j = 0; j < n; j++) {
            int lag = 0;
            for (int k = a + 1; k < b; k++) lag += ptm[k * n + j];
            I->lag[p * n + j] = lag;
            if (ptm[a * n + j] < ptm[b * n + j]) {
                set[j] = 1;
                key[j] = ptm[a * n + j] + lag;
            } else {
                set[j] = 2;
                key[j] = ptm[b * n + j] + lag;
            }
        }
        int32_t *ord = I->johnson + (size_t)p * n;
        for (int j = 0; j < n; j++) {
            int i = j;
            while (i > 0 && johnson_before(set[j], key[j], set[ord[i - 1]], key[ord[i - 1]])) {
                ord[i] = ord[i - 1];
                i--;
            }
            ord[i] = j;
        }
    }
    return I;
}

void orc_inst_free(orc_inst *I)
{
    if (!I) return;
    free(I->lag);
    free(I->johnson);
    free(I);
}

/* ======================================================================
 * Bounds
 * ====================================================================== */

#define P_(I, k, j) ((I)->ptm[(k) * (I)->n + (j)])

int orc_makespan(const orc_inst *I, const int *perm)
{
    int c[ORC_MAXM] = {0};
    for (int i = 0; i < I->n; i++) {
        int j = perm[i];
        c[0] += P_(I, 0, j);
        for (int k = 1; k < I->m; k++) c[k] = imax(c[k - 1], c[k]) + P_(I, k, j);
    }
    return c[I->m - 1];
}

/* front / back / remain of a partial schedule, bound_fsp_weak.cpp:61-130 */
static void lb1_partial(const orc_inst *I, const int *s, int l1, int l2, int *front, int *back,
                        int *remain)
{
    int n = I->n, m = I->m;
    if (l1 == -1) {
        for (int k = 0; k < m; k++) front[k] = I->min_arr[k];
    } else {
        for (int k = 0; k < m; k++) front[k] = 0;
        for (int i = 0; i <= l1; i++) {
            int j = s[i];
            front[0] += P_(I, 0, j);
            for (int k = 1; k < m; k++) front[k] = imax(front[k - 1], front[k]) + P_(I, k, j);
        }
    }
    if (l2 == n) {
        for (int k = 0; k < m; k++) back[k] = I->min_dep[k];
    } else {
        for (int k = 0; k < m; k++) back[k] = 0;
        for (int i = n - 1; i >= l2; i--) {
            int j = s[i];
            back[m - 1] += P_(I, m - 1, j);
            for (int k = m - 2; k >= 0; k--) back[k] = imax(back[k], back[k + 1]) + P_(I, k, j);
        }
    }
    for (int k = 0; k < m; k++) remain[k] = 0;
    for (int i = l1 + 1; i < l2; i++)
        for (int k = 0; k < m; k++) remain[k] += P_(I, k, s[i]);
}

void orc_lb1_children(const orc_inst *I, const int *sched, int l1, int l2, int both, int *cb,
                      int *ce, int *pb, int *pe)
{
    int n = I->n, m = I->m;
    int front[ORC_MAXM], back[ORC_MAXM], remain[ORC_MAXM];
    memset(cb, 0, sizeof(int) * n);
    if (both) memset(ce, 0, sizeof(int) * n);
    lb1_partial(I, sched, l1, l2, front, back, remain);
    for (int i = l1 + 1; i < l2; i++) {
        int j = sched[i];
        /* addFrontAndBound, bound_fsp_weak.cpp:133-161 */
        int t0 = front[0] + P_(I, 0, j);
        int lb = front[0] + remain[0] + back[0];
        int prio = 0;
        for (int k = 1; k < m; k++) {
            int d = t0 - front[k];
            prio += d < 0 ? -d : d;
            int t1 = imax(front[k], t0);
            int c = t1 + remain[k] + back[k];
            t0 = t1 + P_(I, k, j);
            if (c > lb) lb = c;
        }
        cb[j] = lb;
        if (pb) pb[j] = prio;
        if (!both) continue;
        /* addBackAndBound, bound_fsp_weak.cpp:163-191 */
        t0 = back[m - 1] + P_(I, m - 1, j);
        lb = front[m - 1] + remain[m - 1] + back[m - 1];
        prio = 0;
        for (int k = m - 2; k >= 0; k--) {
            int d = t0 - back[k];
            prio += d < 0 ? -d : d;
            int t1 = imax(t0, back[k]);
            int c = front[k] + remain[k] + t1;
            t0 = t1 + P_(I, k, j);
            if (c > lb) lb = c;
        }
        ce[j] = lb;
        if (pe) pe[j] = prio;
    }
}

int orc_lb1_node(const orc_inst *I, const int *perm, int l1, int l2)
{
    int front[ORC_MAXM], back[ORC_MAXM], remain[ORC_MAXM];
    lb1_partial(I, perm, l1, l2, front, back, remain);
    int t0 = front[0] + remain[0];
    int lb = t0 + back[0];
    for (int k = 1; k < I->m; k++) {
        int t1 = imax(front[k] + remain[k], t0);
        lb = imax(lb, t1 + back[k]);
        t0 = t1;
    }
    return lb;
}

int orc_lb2_node(const orc_inst *I, const int *perm, int l1, int l2, int pairs_mode, int *idle)
{
    int n = I->n, m = I->m;
    /* bound_fsp_strong.cpp:543-546: at most one unscheduled job -> exact makespan */
    if (l2 - l1 <= 2) return orc_makespan(I, perm);

    int flag[ORC_MAXN];
    memset(flag, 0, sizeof(int) * n);
    for (int i = 0; i <= l1; i++) flag[perm[i]] = -1;
    for (int i = l2; i < n; i++) flag[perm[i]] = perm[i] + 1;

    int front[ORC_MAXM], back[ORC_MAXM];
    int id = 0;
    /* scheduleFront with idle accumulation, :456-478 */
    if (l1 == -1) {
        for (int k = 0; k < m; k++) front[k] = I->min_arr[k];
        id = 0;
    } else {
        for (int k = 0; k < m; k++) front[k] = 0;
        for (int i = 0; i <= l1; i++) {
            int j = perm[i];
            front[0] += P_(I, 0, j);
            for (int k = 1; k < m; k++) {
                id += imax(0, front[k - 1] - front[k]);
                front[k] = imax(front[k], front[k - 1]) + P_(I, k, j);
            }
        }
    }
    /* scheduleBack, :489-509 (an empty back resets the idle sum) */
    if (l2 == n) {
        for (int k = 0; k < m; k++) back[k] = I->min_dep[k];
        id = 0;
    } else {
        for (int k = 0; k < m; k++) back[k] = 0;
        for (int i = n - 1; i >= l2; i--) {
            int j = perm[i];
            back[m - 1] += P_(I, m - 1, j);
            for (int k = m - 2; k >= 0; k--) {
                id += imax(0, back[k + 1] - back[k]);
                back[k] = imax(back[k], back[k + 1]) + P_(I, k, j);
            }
        }
    }
    if (idle) *idle = id;

    /* borneInfMakespan without early exit, :102-162 */
    int best = 0;
    for (int p = 0; p < I->npairs; p++) {
        int a = I->pair_a[p], b = I->pair_b[p];
        if (pairs_mode == 1 && b < m - 1) continue;
        if (pairs_mode == 2 && b != a + 1) continue;
        int t0 = front[a], t1 = front[b];
        const int32_t *ord = I->johnson + (size_t)p * n;
        const int32_t *lag = I->lag + (size_t)p * n;
        for (int i = 0; i < n; i++) {
            int j = ord[i];
            if (flag[j] != 0) continue;
            t0 += P_(I, a, j);
            t1 = imax(t1, t0 + lag[j]) + P_(I, b, j);
        }
        /* cmaxFin, :64-72 */
        int c = imax(t1 + back[b], t0 + back[a]);
        if (c > best) best = c;
    }
    return best;
}

/* fastNEH.cpp:151-243.  Insertion uses the plain makespan of the partial
 * sequence at every position, which equals the head/inser/tail evaluation. */
static int partial_makespan(const orc_inst *I, const int *seq, int len)
{
    int c[ORC_MAXM] = {0};
    for (int i = 0; i < len; i++) {
        int j = seq[i];
        c[0] += P_(I, 0, j);
        for (int k = 1; k < I->m; k++) c[k] = imax(c[k - 1], c[k]) + P_(I, k, j);
    }
    return c[I->m - 1];
}

int orc_neh(const orc_inst *I, int *perm)
{
    int n = I->n, m = I->m;
    int key[ORC_MAXN];
    for (int j = 0; j < n; j++) {
        key[j] = 0;
        for (int k = 0; k < m; k++) key[j] -= P_(I, k, j);
    }
    /* stable gnome sort by key ascending, common/include/misc.h:32-48 */
    for (int i = 1; i < n; i++) {
        int v = perm[i], t = i;
        while (t > 0 && key[perm[t - 1]] > key[v]) {
            perm[t] = perm[t - 1];
            t--;
        }
        perm[t] = v;
    }
    /* first two jobs: keep the sorted order only if strictly better, :233-236 */
    int c1 = partial_makespan(I, perm, 2);
    int sw[2] = {perm[1], perm[0]};
    int c2 = partial_makespan(I, sw, 2);
    if (!(c1 < c2)) {
        perm[0] = sw[0];
        perm[1] = sw[1];
    }
    int tmp[ORC_MAXN];
    for (int k = 2; k < n; k++) {
        int job = perm[k];
        int bestc = INT_MAX, bestpos = 0;
        for (int p = 0; p <= k; p++) {
            int t = 0;
            for (int i = 0; i < p; i++) tmp[t++] = perm[i];
            tmp[t++] = job;
            for (int i = p; i < k; i++) tmp[t++] = perm[i];
            int c = partial_makespan(I, tmp, k + 1);
            if (c < bestc) {
                bestc = c;
                bestpos = p;
            }
        }
        for (int p = k; p > bestpos; p--) perm[p] = perm[p - 1];
        perm[bestpos] = job;
    }
    return orc_makespan(I, perm);
}

/* ======================================================================
 * IVM single-explorer B&B
 * ====================================================================== */

#define FRONT 0
#define BACK 1

struct orc_bb {
    const orc_inst *I;
    orc_cfg cfg;
    int n;
    /* ivm (multicore/include/ivm.h) */
    int *mat, *pos, *end, *dir;
    int line;
    /* decoded node (ivm_bound::node) */
    int *sched;
    int l1, l2;
    /* costs[WEAK|STRONG][BEGIN|END], job indexed */
    int *cw_b, *cw_e, *cs_b, *cs_e;
    int local_best, initial_ub;
    int best_cost;
    int *best_perm;
    orc_stats st;
    uint64_t histo[ORC_MAXN + 1];
};

static inline int negative(int i) { return i < 0 ? i : -i - 1; } /* misc.h:19-23 */
static inline int absolute(int i) { return i >= 0 ? i : -i - 1; } /* misc.h:25-29 */

orc_bb *orc_bb_new(const orc_inst *I, const orc_cfg *cfg, int ub, const int *ub_perm)
{
    int n = I->n;
    orc_bb *B = (orc_bb *)calloc(1, sizeof(orc_bb));
    B->I = I;
    B->cfg = *cfg;
    B->n = n;
    B->mat = (int *)calloc((size_t)n * n, sizeof(int));
    B->pos = (int *)calloc(n, sizeof(int));
    B->end = (int *)calloc(n, sizeof(int));
    B->dir = (int *)calloc(n, sizeof(int));
    B->sched = (int *)calloc(n, sizeof(int));
    B->cw_b = (int *)calloc(n, sizeof(int));
    B->cw_e = (int *)calloc(n, sizeof(int));
    B->cs_b = (int *)calloc(n, sizeof(int));
    B->cs_e = (int *)calloc(n, sizeof(int));
    B->best_perm = (int *)calloc(n, sizeof(int));
    B->pos[0] = n; /* empty interval, ivm.cpp:15 */
    B->best_cost = ub;
    B->local_best = ub;
    B->initial_ub = ub;
    for (int i = 0; i < n; i++) B->best_perm[i] = ub_perm ? ub_perm[i] : i;
    return B;
}

void orc_bb_free(orc_bb *B)
{
    if (!B) return;
    free(B->mat);
    free(B->pos);
    free(B->end);
    free(B->dir);
    free(B->sched);
    free(B->cw_b);
    free(B->cw_e);
    free(B->cs_b);
    free(B->cs_e);
    free(B->best_perm);
    free(B);
}

static inline int gt_best(const orc_bb *B, int v) /* ivm_bound.cpp:678-682 */
{
    return B->cfg.find_all ? v > B->local_best : v >= B->local_best;
}

/* ivm_bound::prepareSchedule, ivm_bound.cpp:111-138 */
static void decode(orc_bb *B)
{
    int n = B->n;
    B->l1 = -1;
    B->l2 = n;
    for (int l = 0; l < B->line; l++) {
        int job = absolute(B->mat[l * n + B->pos[l]]);
        if (B->dir[l] == 0)
            B->sched[++B->l1] = job;
        else
            B->sched[--B->l2] = job;
    }
    for (int l = 0; l < n - B->line; l++) B->sched[B->l1 + 1 + l] = absolute(B->mat[B->line * n + l]);
}

/* ivm_bound::computeWeakBounds, :215-236 */
static void compute_weak(orc_bb *B)
{
    int n = B->n;
    memset(B->cs_b, 0, sizeof(int) * n);
    memset(B->cs_e, 0, sizeof(int) * n);
    int both = B->cfg.branching > 0;
    orc_lb1_children(B->I, B->sched, B->l1, B->l2, both, B->cw_b, B->cw_e, NULL, NULL);
    B->st.simple_bounds += (uint64_t)(both ? 2 : 1) * (B->l2 - B->l1 - 1);
}

/* ivm_bound::chooseChildrenSet, :413-507 (sort priorities unused: sortedDFS=0) */
static int choose_dir(const orc_bb *B, const int *cb, const int *ce)
{
    int n = B->n;
    switch (B->cfg.branching) {
    case -3: return (B->line % 2) ? BACK : FRONT;
    case -2: return BACK;
    case -1: return FRONT;
    case 1: {
        int sum = 0;
        for (int i = 0; i < n; i++) sum += cb[i] - ce[i];
        return sum < 0 ? BACK : FRONT;
    }
    case 2: {
        int ub = B->cfg.single_node ? B->local_best : B->initial_ub;
        int elim = 0, sum = 0;
        for (int i = 0; i < n; i++) {
            if (cb[i] >= ub) elim++;
            else sum += cb[i];
            if (ce[i] >= ub) elim--;
            else sum -= ce[i];
        }
        if (elim > 0) return FRONT;
        if (elim < 0) return BACK;
        return sum < 0 ? BACK : FRONT;
    }
    case 3: {
        int ub = B->cfg.single_node ? B->local_best : B->initial_ub;
        int mn = INT_MAX, minc = 0, elimc = 0;
        for (int i = 0; i < n; i++) {
            if (cb[i] < mn && cb[i] > 0) mn = cb[i];
            if (ce[i] < mn && ce[i] > 0) mn = ce[i];
        }
        for (int i = 0; i < n; i++) {
            if (cb[i] == mn) minc++;
            if (ce[i] == mn) minc--;
            if (cb[i] >= ub) elimc++;
            if (ce[i] >= ub) elimc--;
        }
        if (minc > 0) return BACK;
        if (minc < 0) return FRONT;
        return elimc > 0 ? FRONT : BACK;
    }
    }
    return -1;
}

/* refinement bound of one child schedule: ivm_bound.cpp:170 (LB1') or :171 (LB2) */
static int refine_bound(const orc_bb *B, const int *s, int l1, int l2)
{
    if (B->cfg.refine == 0) return orc_lb1_node(B->I, s, l1, l2);
    return orc_lb2_node(B->I, s, l1, l2, B->cfg.johnson_pairs, NULL);
}

/* ivm_bound::computeStrongBounds, :141-195 */
static void compute_strong(orc_bb *B, int be)
{
    int n = B->n;
    int L1 = B->l1 + (be == FRONT ? 1 : 0);
    int L2 = B->l2 - (be == BACK ? 1 : 0);
    int *first = be == FRONT ? B->cs_b : B->cs_e;
    const int *second = be == FRONT ? B->cw_b : B->cw_e;
    memset(first, 0, sizeof(int) * n);
    int fill = be == FRONT ? L1 : L2;
    for (int i = B->l1 + 1; i < B->l2; i++) {
        int job = B->sched[i];
        if (!gt_best(B, second[job])) {
            int t = B->sched[fill];
            B->sched[fill] = B->sched[i];
            B->sched[i] = t;
            first[job] = refine_bound(B, B->sched, L1, L2);
            t = B->sched[fill];
            B->sched[fill] = B->sched[i];
            B->sched[i] = t;
            B->st.johnson_bounds++;
        }
    }
}

/* ivm_bound::prune + eliminateJobs (sortedDFS=0), :337-410, :610-646 */
static void prune(orc_bb *B, const int *c1b, const int *c1e, const int *c2b, const int *c2e)
{
    int n = B->n, L = B->line;
    int *row = B->mat + L * n;
    int prev = L > 0 ? B->dir[L - 1] : 0;
    if (prev != B->dir[L]) {
        for (int i1 = 0, i2 = n - L - 1; i1 < i2; i1++, i2--) {
            int t = row[i1];
            row[i1] = row[i2];
            row[i2] = t;
        }
    }
    if (prev == 1 && B->dir[L] == 0)
        for (int l = 0; l < n - L; l++) B->sched[B->l1 + 1 + l] = absolute(row[l]);
    const int *c1 = B->dir[L] == 1 ? c1e : (B->dir[L] == 0 ? c1b : NULL);
    const int *c2 = B->dir[L] == 1 ? c2e : (B->dir[L] == 0 ? c2b : NULL);
    if (!c1) return;
    for (int i = 0; i < n - L; i++) {
        int job = row[i];
        if (gt_best(B, c1[job]) || gt_best(B, c2[job])) row[i] = negative(job);
    }
}

static void weak_prune(orc_bb *B)
{
    compute_weak(B);
    B->dir[B->line] = choose_dir(B, B->cw_b, B->cw_e);
    prune(B, B->cw_b, B->cw_e, B->cw_b, B->cw_e);
}

static void mixed_prune(orc_bb *B)
{
    compute_weak(B);
    B->dir[B->line] = choose_dir(B, B->cw_b, B->cw_e);
    compute_strong(B, B->dir[B->line] == 1 ? BACK : FRONT);
    prune(B, B->cw_b, B->cw_e, B->cs_b, B->cs_e);
}

/* strongBoundPrune, :668-676: weak costs are memset with INT_MAX bytes (-1
 * ints), so every child is refined in both directions. */
static void strong_prune(orc_bb *B)
{
    int n = B->n;
    B->dir[B->line] = -1;
    for (int i = 0; i < n; i++) B->cw_b[i] = B->cw_e[i] = -1;
    compute_strong(B, FRONT);
    compute_strong(B, BACK);
    B->dir[B->line] = choose_dir(B, B->cs_b, B->cs_e);
    prune(B, B->cs_b, B->cs_e, B->cs_b, B->cs_e);
}

static void bound_node(orc_bb *B)
{
    switch (B->cfg.bound_mode) {
    case 0: weak_prune(B); break;
    case 1: strong_prune(B); break;
    case 2: mixed_prune(B); break;
    }
}

/* ivm_bound::boundLeaf, :238-271 + solution::update (strict <), solution.cpp:26-41 */
static void bound_leaf(orc_bb *B)
{
    int n = B->n;
    B->st.leaves++;
    int cost = orc_makespan(B->I, B->sched);
    if (!gt_best(B, cost)) {
        B->local_best = cost;
        if (cost < B->best_cost) {
            B->best_cost = cost;
            memcpy(B->best_perm, B->sched, sizeof(int) * n);
        }
    }
    int *cell = B->mat + B->line * n + B->pos[B->line];
    *cell = negative(*cell);
}

void orc_bb_set_root(orc_bb *B, const int *root_perm)
{
    int n = B->n;
    memset(B->mat, 0, sizeof(int) * n * n);
    memset(B->pos, 0, sizeof(int) * n);
    memset(B->end, 0, sizeof(int) * n);
    B->pos[0] = n;
    for (int i = 0; i < n; i++) B->mat[i] = root_perm[i];
    B->line = 0;
    decode(B);
    /* boundRoot: the root row is always bounded with LB1 (weakBoundPrune) */
    B->local_best = B->best_cost;
    B->l1 = -1;
    B->l2 = n;
    for (int i = 0; i < n; i++) B->sched[i] = root_perm[i];
    weak_prune(B);
    memset(B->cs_b, 0, sizeof(int) * n);
    memset(B->cs_e, 0, sizeof(int) * n);
}

/* ivm::beforeEnd, ivm.cpp:152-160 (inclusive) */
static int before_end(const orc_bb *B)
{
    for (int i = 0; i < B->n; i++) {
        if (B->pos[i] < B->end[i]) return 1;
        if (B->pos[i] > B->end[i]) return 0;
    }
    return 1;
}

/* ivm::generateLine, ivm.cpp:115-131 */
static void generate_line(orc_bb *B, int L, int explore)
{
    int n = B->n;
    int col = B->pos[L - 1];
    const int *src = B->mat + (L - 1) * n;
    int *dst = B->mat + L * n;
    for (int i = 0; i < col; i++) dst[i] = absolute(src[i]);
    for (int i = col; i < n - L; i++) dst[i] = absolute(src[i + 1]);
    if (explore) {
        B->pos[L] = 0;
        B->dir[L] = 0;
    }
}

/* sequentialbb::unfold, sequentialbb.cpp:206-264 (bounded, uncounted) */
static void unfold(orc_bb *B)
{
    int n = B->n;
    while (B->line < n - 2) {
        if (B->mat[B->line * n + B->pos[B->line]] < 0) {
            for (int i = B->line + 1; i < n; i++) B->pos[i] = 0;
            break;
        }
        B->line++;
        generate_line(B, B->line, 0);
        decode(B);
        bound_node(B);
    }
}

int orc_bb_init_interval(orc_bb *B, const int *pos, const int *end)
{
    B->line = 0;
    memcpy(B->pos, pos, sizeof(int) * B->n);
    memcpy(B->end, end, sizeof(int) * B->n);
    B->st.simple_bounds = 0;
    for (int i = 0; i < B->n; i++)
        if (pos[i] < 0 || pos[i] >= B->n - i || end[i] < 0 || end[i] >= B->n - i) return -1;
    if (before_end(B)) {
        unfold(B);
        return 1;
    }
    return 0;
}

int orc_bb_init_full(orc_bb *B)
{
    int z[ORC_MAXN], e[ORC_MAXN];
    for (int i = 0; i < B->n; i++) {
        z[i] = 0;
        e[i] = B->n - i - 1;
    }
    return orc_bb_init_interval(B, z, e);
}

/* sequentialbb::next, sequentialbb.cpp:110-191 with ivm.cpp:50-104 */
int orc_bb_next(orc_bb *B)
{
    int n = B->n;
    B->local_best = B->best_cost;
    int state = 0;
    while (before_end(B)) {
        int L = B->line;
        if (B->pos[L] >= n - L) { /* lineEndState -> goUp */
            if (L > 0) {
                B->pos[L] = 0;
                B->line = --L;
                int *cell = B->mat + L * n + B->pos[L];
                *cell = negative(*cell);
            } else {
                B->pos[0]++;
            }
            continue;
        }
        if (B->mat[L * n + B->pos[L]] < 0) { /* pruningCellState -> goRight */
            B->pos[L]++;
            continue;
        }
        state = 1;
        B->st.decomposed++;
        B->line = ++L; /* goDown */
        generate_line(B, L, 1);
        B->histo[L]++;
        decode(B);
        if (L == n - 1) {
            bound_leaf(B);
            state = 0;
            continue;
        }
        break;
    }
    if (state == 1) bound_node(B);
    return state;
}

int orc_bb_run(orc_bb *B, uint64_t max_nodes)
{
    uint64_t start = B->st.decomposed;
    while (orc_bb_next(B)) {
        if (max_nodes && B->st.decomposed - start >= max_nodes) return 0;
    }
    return 1;
}

void orc_bb_stats(const orc_bb *B, orc_stats *st) { *st = B->st; }

int orc_bb_best(const orc_bb *B, int *perm)
{
    if (perm) memcpy(perm, B->best_perm, sizeof(int) * B->n);
    return B->best_cost;
}

int orc_bb_line(const orc_bb *B) { return B->line; }

void orc_bb_interval(const orc_bb *B, int *pos, int *end)
{
    memcpy(pos, B->pos, sizeof(int) * B->n);
    memcpy(end, B->end, sizeof(int) * B->n);
}

void orc_bb_depth_histo(const orc_bb *B, uint64_t *histo)
{
    memcpy(histo, B->histo, sizeof(uint64_t) * (B->n + 1));
}

/* ivm::countExplorableSubtrees + cuttingPosition, ivm.cpp:188-237 */
static int cutting_position(const orc_bb *B, int L, int division)
{
    int n = B->n;
    int expl = 0;
    for (int i = B->pos[L] + 1; i <= B->end[L]; i++)
        if (B->mat[L * n + i] >= 0) expl++;
    int keep = expl / division;
    int p = B->pos[L] + 1, kept = 0;
    while (kept < keep) {
        if (B->mat[L * n + p] >= 0) kept++;
        p++;
    }
    return p;
}

/* bbthread::shareWork, bbthread.cpp:39-72 (split the victim's interval at
 * the first line whose pos differs from end; rows are copied, no re-bounding) */
int orc_bb_share(orc_bb *V, orc_bb *T, int margin)
{
    int n = V->n;
    int l = 0;
    while (V->pos[l] == V->end[l] && l < V->line && l < n - margin) l++;
    if (!(V->pos[l] < V->end[l])) return 0;
    for (int i = 0; i < l; i++) {
        T->pos[i] = V->pos[i];
        memcpy(T->mat + i * n, V->mat + i * n, sizeof(int) * n);
        T->dir[i] = V->dir[i];
    }
    memcpy(T->end, V->end, sizeof(int) * n);
    memcpy(T->mat + l * n, V->mat + l * n, sizeof(int) * n);
    T->dir[l] = V->dir[l];
    T->pos[l] = cutting_position(V, l, 2);
    V->end[l] = T->pos[l] - 1;
    for (int i = l + 1; i < n; i++) T->pos[i] = 0;
    for (int i = l + 1; i < n; i++) V->end[i] = n - i - 1;
    T->line = l;
    return 1;
}

/* ======================================================================
 * Batched checkers (the same functions over arrays of partial schedules)
 * ====================================================================== */

void orc_bounds_batch(const orc_inst *I, int count, const int *perms, const int *l1, const int *l2,
                      int pairs_mode, int *cb, int *ce, int *cbf, int *lb1n, int *lb2, int *mk)
{
    int n = I->n;
    for (int s = 0; s < count; s++) {
        const int *p = perms + (size_t)s * n;
        if (cb) orc_lb1_children(I, p, l1[s], l2[s], 1, cb + (size_t)s * n, ce + (size_t)s * n, NULL, NULL);
        if (cbf) {
            int dummy[ORC_MAXN];
            orc_lb1_children(I, p, l1[s], l2[s], 0, cbf + (size_t)s * n, dummy, NULL, NULL);
        }
        if (lb1n) lb1n[s] = orc_lb1_node(I, p, l1[s], l2[s]);
        if (lb2) lb2[s] = orc_lb2_node(I, p, l1[s], l2[s], pairs_mode, NULL);
        if (mk) mk[s] = orc_makespan(I, p);
    }
}

/* ======================================================================
 * Iterated greedy (heuristic/src/IG.cpp, fspnhoods.cpp, fastinsertremove.cpp)
 *
 * The reference draws from std::random_device-seeded engines and compares a
 * float exp in its acceptance, so no run of it can be repeated.  This is the
 * same sequence of steps driven by a seeded SplitMix64 stream, with the
 * acceptance probabilities given as integer thresholds and the removal
 * weights as integers; the engine's device IG follows these steps exactly.
 * ====================================================================== */

/* fastInsertRemove::insertMakespans (fastinsertremove.cpp:196-214, with
 * computeHeads/computeTails/computeInser :75-134): out[i], i = 0..len, is the
 * makespan of seq[0..len-1] with job inserted before position i.  The tail
 * beyond the sequence is zero (the reference reads tail[k][len], which
 * computeTails never writes).  pinned: the reference's insertMakespans on
 * seeded sequences with that cell zeroed (oracle/_ref, tests/golden). */
void orc_insert_makespans(const orc_inst *I, const int *seq, int len, int job, int *out)
{
    const int m = I->m;
    for (int i = 0; i <= len; i++) {
        int c[ORC_MAXM] = {0};
        for (int t = 0; t < i; t++) {
            int j = seq[t];
            c[0] += P_(I, 0, j);
            for (int k = 1; k < m; k++) c[k] = imax(c[k - 1], c[k]) + P_(I, k, j);
        }
        c[0] += P_(I, 0, job);
        for (int k = 1; k < m; k++) c[k] = imax(c[k - 1], c[k]) + P_(I, k, job);
        for (int t = i; t < len; t++) {
            int j = seq[t];
            c[0] += P_(I, 0, j);
            for (int k = 1; k < m; k++) c[k] = imax(c[k - 1], c[k]) + P_(I, k, j);
        }
        out[i] = c[m - 1];
    }
}

/* fastInsertRemove::removeMakespans (fastinsertremove.cpp:243-279): out[i],
 * i = 0..len-1, is the makespan of seq without position i (len >= 2).
 * pinned: the reference's removeMakespans on seeded sequences. */
void orc_remove_makespans(const orc_inst *I, const int *seq, int len, int *out)
{
    int tmp[ORC_MAXN];
    for (int i = 0; i < len; i++) {
        int t = 0;
        for (int q = 0; q < len; q++)
            if (q != i) tmp[t++] = seq[q];
        out[i] = partial_makespan(I, tmp, len - 1);
    }
}

/* IG::IG visitOrder (IG.cpp:38-52): the middle position, then alternately
 * the positions below and above it, outwards. */
void orc_ig_visit_order(int n, int *vo)
{
    int start = n / 2, ind = 0;
    for (int i = start; i < n; i++, ind += 2) vo[ind] = i;
    ind = 1;
    for (int i = start - 1; i >= 0; i--, ind += 2) vo[ind] = i;
}

/* uniform draw in [0, k), k >= 1, from the high 32 bits of the stream
 * (multiply-shift with rejection of the biased low products) */
uint32_t orc_ig_below(uint64_t *s, uint32_t k)
{
    for (;;) {
        uint32_t x = (uint32_t)(orc_splitmix_next(s) >> 32);
        uint64_t mm = (uint64_t)x * k;
        uint32_t lo = (uint32_t)mm;
        if (lo < k) {
            uint32_t t = (uint32_t)(0u - k) % k;
            if (lo < t) continue;
        }
        return (uint32_t)(mm >> 32);
    }
}

uint64_t orc_ig_chain_seed(uint64_t seed, int inst, int chain)
{
    uint64_t s = seed ^ (((uint64_t)(uint32_t)inst << 32) | (uint32_t)chain);
    return orc_splitmix_next(&s);
}

/* IG::acceptance (IG.cpp:319-338): a worse (or equal) cost delta >= 0 is
 * accepted with probability exp(-delta / T), T = 0.2 * avgPT / 10, avgPT the
 * mean processing time (IG.cpp:24-32); thr[d] = floor(2^32 exp(-d/T)) for
 * d < ORC_IG_THR (thr[0] = 2^32: equal costs are always accepted), larger
 * deltas are refused.  A draw x (the high 32 bits of the stream) accepts
 * when x < thr[delta]. */
void orc_ig_thresholds(const orc_inst *I, uint64_t *thr)
{
    double sum = 0.0;
    for (int i = 0; i < I->n * I->m; i++) sum += I->ptm[i];
    double T = 0.2 * (sum / ((double)I->n * I->m)) / 10.0;
    for (int d = 0; d < ORC_IG_THR; d++) {
        double p = T > 0.0 ? exp(-(double)d / T) : (d == 0 ? 1.0 : 0.0);
        double v = floor(p * 4294967296.0);
        thr[d] = v >= 4294967296.0 ? (1ull << 32) : (uint64_t)v;
    }
}

static void seq_remove(int *seq, int *len, int pos)
{
    for (int p = pos; p < *len - 1; p++) seq[p] = seq[p + 1];
    (*len)--;
}

static void seq_insert(int *seq, int *len, int pos, int job)
{
    for (int p = *len; p > pos; p--) seq[p] = seq[p - 1];
    seq[pos] = job;
    (*len)++;
}

/* fspnhood::kImove (fspnhoods.cpp:46-100) with fastInsertRemove::bestInsert
 * (fastinsertremove.cpp:326-363) and bestRemove2 (:463-496): remove the job
 * at pos, insert it at the best non-tabu position; while that is worse than
 * the starting makespan cmax0 (and fewer than kmax tries), remove a
 * non-tabu job drawn with weight (gain / total time)^2 and insert it best.
 * The weight is w = q^2, q = floor(256 gain / total time) (gain <= total
 * time, so q <= 256); with every weight zero the first non-tabu position
 * is removed.  Returns the makespan of seq on exit. */
static int ig_ki_move(const orc_inst *I, int *seq, int n, int pos, int kmax, int cmax0, const int *sumpt,
                      uint64_t *rng)
{
    int len = n;
    int tabu_jobs[ORC_MAXN], ntj = 0, tabu_pos;
    int ms[ORC_MAXN + 1];
    int rjob = seq[pos];
    seq_remove(seq, &len, pos);
    tabu_jobs[ntj++] = rjob;
    tabu_pos = pos;
    int cmax1 = 0;
    for (int k = 0;;) {
        orc_insert_makespans(I, seq, len, rjob, ms);
        int best = INT_MAX, bp = 0;
        for (int i = 0; i <= len; i++) {
            if (i == tabu_pos) continue;
            if (ms[i] < best) {
                best = ms[i];
                bp = i;
            }
        }
        seq_insert(seq, &len, bp, rjob);
        cmax1 = best;
        if (cmax1 <= cmax0) break;
        if (++k == kmax) break;
        int old = partial_makespan(I, seq, len);
        orc_remove_makespans(I, seq, len, ms);
        uint32_t w[ORC_MAXN], W = 0;
        int first_free = -1;
        for (int i = 0; i < len; i++) {
            int job = seq[i], tabu = 0;
            for (int t = 0; t < ntj; t++) tabu |= tabu_jobs[t] == job;
            w[i] = 0;
            if (!tabu) {
                if (first_free < 0) first_free = i;
                uint64_t g = (uint64_t)(old - ms[i]);
                if (sumpt[job] > 0) {
                    uint64_t q = (g << 8) / (uint64_t)sumpt[job];
                    w[i] = (uint32_t)(q * q);
                }
            }
            W += w[i];
        }
        int pick = first_free < 0 ? 0 : first_free;
        if (W > 0) {
            uint32_t u = orc_ig_below(rng, W), acc = 0;
            for (int i = 0; i < len; i++) {
                acc += w[i];
                if (u < acc) {
                    pick = i;
                    break;
                }
            }
        }
        cmax1 = ms[pick];
        rjob = seq[pick];
        seq_remove(seq, &len, pick);
        tabu_jobs[ntj++] = rjob;
        tabu_pos = pick;
    }
    return cmax1;
}

/* IG::localSearchKI (IG.cpp:566-610): first improvement over the positions
 * in visit order, restarted after every improvement, until none improves. */
static int ig_local_search(const orc_inst *I, int *seq, int n, int cost, int kmax, const int *vo,
                           const int *sumpt, uint64_t *rng)
{
    int tmp[ORC_MAXN];
    for (int it = 0; it < 10000; it++) {
        int found = 0;
        for (int j = 0; j < n; j++) {
            memcpy(tmp, seq, sizeof(int) * n);
            int c = ig_ki_move(I, tmp, n, vo[j], kmax, cost, sumpt, rng);
            if (c < cost) {
                cost = c;
                memcpy(seq, tmp, sizeof(int) * n);
                found = 1;
                break;
            }
        }
        if (!found) break;
    }
    return cost;
}

/* IG::runIG (IG.cpp:341-404): from the NEH order, `iters` rounds of
 * destruction (d jobs at positions drawn from the first r of the visit
 * order, r uniform in [d, n], IG.cpp:94-150), greedy re-insertion at the
 * first best position (IG.cpp:295-317), local search, and acceptance; d
 * restarts at 2 after an acceptance and grows to 5 (then wraps to 2) after
 * refusals.  Returns the best makespan seen; perm receives its order.
 * Restatement with its own seeded generator: parity pinned by the moves
 * above against the reference, and by the properties of its result (no
 * worse than NEH, the order realises the cost, optimal on small instances). */
int orc_ig(const orc_inst *I, uint64_t state, int iters, const uint64_t *thr, int *perm)
{
    const int n = I->n, m = I->m;
    int cur[ORC_MAXN], tmp[ORC_MAXN], best[ORC_MAXN], vo[ORC_MAXN], sumpt[ORC_MAXN];
    int v[ORC_MAXN], removed[ORC_MAXN], ms[ORC_MAXN + 1];
    for (int j = 0; j < n; j++) {
        cur[j] = j;
        sumpt[j] = 0;
        for (int k = 0; k < m; k++) sumpt[j] += P_(I, k, j);
    }
    int curcost = orc_neh(I, cur);
    int bestcost = curcost;
    memcpy(best, cur, sizeof(int) * n);
    orc_ig_visit_order(n, vo);
    int kmax = 0;
    while ((kmax + 1) * (kmax + 1) <= n) kmax++;
    int d = 2;
    for (int it = 0; it < iters && n > 2; it++) {
        memcpy(tmp, cur, sizeof(int) * n);
        int dd = d < n - 1 ? d : n - 1;
        int r = dd + (int)orc_ig_below(&state, (uint32_t)(n - dd + 1));
        for (int i = 0; i < r; i++) v[i] = vo[i];
        for (int i = r - 1; i > 0; i--) {
            int j = (int)orc_ig_below(&state, (uint32_t)(i + 1));
            int t = v[i];
            v[i] = v[j];
            v[j] = t;
        }
        int gone[ORC_MAXN] = {0};
        for (int i = 0; i < dd; i++) {
            removed[i] = tmp[v[i]];
            gone[v[i]] = 1;
        }
        int len = 0;
        for (int i = 0; i < n; i++)
            if (!gone[i]) tmp[len++] = tmp[i];
        for (int i = 0; i < dd; i++) {
            orc_insert_makespans(I, tmp, len, removed[i], ms);
            int bc = INT_MAX, bp = 0;
            for (int p = 0; p <= len; p++)
                if (ms[p] < bc) {
                    bc = ms[p];
                    bp = p;
                }
            seq_insert(tmp, &len, bp, removed[i]);
        }
        int tc = ig_local_search(I, tmp, n, partial_makespan(I, tmp, n), kmax, vo, sumpt, &state);
        int acc;
        if (tc < curcost) {
            acc = 1;
        } else {
            int delta = tc - curcost;
            uint64_t x = orc_splitmix_next(&state) >> 32;
            acc = delta < ORC_IG_THR && x < thr[delta];
        }
        if (acc) {
            memcpy(cur, tmp, sizeof(int) * n);
            curcost = tc;
            d = 2;
        } else if (++d > 5) {
            d = 2;
        }
        if (tc < bestcost) {
            bestcost = tc;
            memcpy(best, tmp, sizeof(int) * n);
        }
    }
    memcpy(perm, best, sizeof(int) * n);
    return bestcost;
}
