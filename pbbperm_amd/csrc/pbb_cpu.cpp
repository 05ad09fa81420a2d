// pbb_cpu.cpp -- host multicore explorers working beside the GPU engine on
// the same search: the co-worker of include/pbb_engine.h ("host co-worker").
//
// The reference runs IVM explorers on CPU threads (multicore/src/
// matrix_controller.cpp, bbthread.cpp, sequentialbb.cpp) and hands intervals
// between a master and its workers as factoradic digit vectors
// (distributed/src/worker_mc.cpp:26-111).  Here the host threads take the
// same work units as the GPU engine -- pbb_cpu_init_intervals /
// pbb_cpu_get_intervals mirror pbb_init_intervals / pbb_get_intervals -- so
// intervals exported by the engine can be explored on idle host cores and
// returned.  pbb_hybrid_run drives the engine and the host explorers on one
// search, moving intervals to the host when its explorers run dry.
//
// Every node is processed with the engine's semantics (which are the
// reference CPU path's, multicore/src/ivm_bound.cpp and sequentialbb.cpp):
// the same LB1 / LB1' / Johnson LB2 values, branching rules, pruning and
// counters, so with a frozen incumbent the node counts of an interval do not
// depend on which side explores it.  The front and back completion times are
// kept as stacks (one push per goDown, one pop per goUp) instead of being
// recomputed per node (bound_fsp_weak::computePartial); the recurrences are
// the same, so the values are identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/pbb_engine.h"
#include "pbb_internal.h"

using namespace pbb;

namespace {

constexpr int MAXN = PBB_MAX_JOBS + 1;
constexpr int MAXM = PBB_MAX_MACHINES;

inline int neg_flag(int j) { return j < 0 ? j : -j - 1; }  // common/include/misc.h negative()
inline int abs_flag(int j) { return j >= 0 ? j : -j - 1; } // common/include/misc.h absolute()

using Clock = std::chrono::steady_clock;

struct CpuShared;

// One IVM explorer (multicore/include/ivm.h + ivm_bound) on one host thread.
struct Cpx {
    CpuShared *S = nullptr;
    int id = 0;
    int n = 0, m = 0;
    int mat[MAXN][MAXN];       // row l: candidate jobs at depth l, pruned/visited cells negative
    int pos[MAXN], end[MAXN], dir[MAXN];
    int line = 0;
    // completion-time stacks: F[s] = front after s front jobs (F[0] = min_arr),
    // B[s] = back after s back jobs (B[0] = min_dep); rem = unscheduled work
    int F[MAXN + 1][MAXM], B[MAXN + 1][MAXM], rem[MAXM];
    int nf = 0, nb = 0;
    int local_best = INT_MAX;
    bool has_work = false;
    uint64_t dec = 0, leaves = 0, sb = 0, jb = 0;
    uint64_t hist[MAXN + 1], jhist[MAXN + 1];
    std::atomic<int> handed{0};  // set by a donor when it wrote an interval here

    void reset_stacks();
    void push(int j, int d);
    void pop(int j, int d);
    bool before_end() const;
    bool gt_best(int v) const;
    int choose(const int *cb, const int *ce, int k) const;
    int refine(int i, int d, int k) const;
    void bound_node(bool root);
    void leaf();
    void go_down(bool explore);
    void go_up();
    bool next();
    void unfold();
    bool start(const int *p, const int *e);
    void rebuild_stacks();
};

struct CpuShared {
    HostTables T;
    pbb_config cfg;
    int n = 0, m = 0;
    std::vector<int> tot;          // per-machine total work
    std::vector<int> pair_rows;    // Johnson table rows (pair indices) the JohnsonMode keeps
    int ub0 = INT_MAX;             // incumbent when the root was bounded (branching thresholds)
    std::vector<int> root_perm, root_row;
    int root_dir = 0;
    bool root_bounded = false;
    // incumbent: cost in the high word; its schedule under mu
    std::atomic<long long> best{INT_MAX};
    std::mutex mu;
    int best_cost = INT_MAX;
    std::vector<int> best_perm;
    // run control
    std::atomic<int> stop{0};
    std::atomic<int> interrupt{0};
    std::atomic<int> holders{0};   // explorers holding an interval
    std::atomic<int> done{0};
    std::atomic<uint64_t> run_dec{0};
    uint64_t node_budget = 0;
    Clock::time_point deadline;
    bool has_deadline = false;
    // idle explorers waiting for a donation
    std::mutex qmu;
    std::vector<int> waiting;
    std::atomic<int> nwaiting{0};
    int steal_margin = 6;
    std::vector<Cpx *> ex;

    void offer(int cost, const int *perm)
    {
        long long cur = best.load(std::memory_order_relaxed);
        while (cost < cur && !best.compare_exchange_weak(cur, cost)) {
        }
        std::lock_guard<std::mutex> g(mu);
        if (cost < best_cost) {
            best_cost = cost;
            best_perm.assign(perm, perm + n);
        }
    }
};

void Cpx::reset_stacks()
{
    nf = nb = 0;
    for (int q = 0; q < m; q++) {
        F[0][q] = S->T.min_arr[q];
        B[0][q] = S->T.min_dep[q];
        rem[q] = S->tot[q];
    }
}

// schedule job j in direction d (the max-plus chains of
// bound_fsp_weak::scheduleFront / scheduleBack, bound_fsp_weak.cpp:61-110)
void Cpx::push(int j, int d)
{
    const int *p = S->T.ptm.data();
    if (d == 0) {
        const int *f = F[nf];
        int *g = F[nf + 1];
        int c = f[0] + p[j];
        g[0] = c;
        for (int q = 1; q < m; q++) {
            c = std::max(f[q], c) + p[q * n + j];
            g[q] = c;
        }
        nf++;
    } else {
        const int *b = B[nb];
        int *g = B[nb + 1];
        int c = b[m - 1] + p[(m - 1) * n + j];
        g[m - 1] = c;
        for (int q = m - 2; q >= 0; q--) {
            c = std::max(b[q], c) + p[q * n + j];
            g[q] = c;
        }
        nb++;
    }
    for (int q = 0; q < m; q++) rem[q] -= p[q * n + j];
}

void Cpx::pop(int j, int d)
{
    const int *p = S->T.ptm.data();
    if (d == 0) nf--;
    else nb--;
    for (int q = 0; q < m; q++) rem[q] += p[q * n + j];
}

// ivm::beforeEnd (multicore/src/ivm.cpp:152-160): lexicographic pos <= end
bool Cpx::before_end() const
{
    for (int i = 0; i < n; i++) {
        if (pos[i] < end[i]) return true;
        if (pos[i] > end[i]) return false;
    }
    return true;
}

// ivm_bound::gtBest (multicore/src/ivm_bound.cpp:678-682)
bool Cpx::gt_best(int v) const { return S->cfg.find_all ? v > local_best : v >= local_best; }

// ivm_bound::chooseChildrenSet (multicore/src/ivm_bound.cpp:413-507), CPU
// numbering; cb / ce indexed by row cell
int Cpx::choose(const int *cb, const int *ce, int k) const
{
    const pbb_config &c = S->cfg;
    switch (c.branching) {
    case -3: return (line % 2) ? 1 : 0;
    case -2: return 1;
    case -1: return 0;
    case 1: {
        int s = 0;
        for (int i = 0; i < k; i++) s += cb[i] - ce[i];
        return s < 0 ? 1 : 0;
    }
    case 2: {
        const int ub = c.single_node ? local_best : S->ub0;
        int elim = 0, s = 0;
        for (int i = 0; i < k; i++) {
            if (cb[i] >= ub) elim++;
            else s += cb[i];
            if (ce[i] >= ub) elim--;
            else s -= ce[i];
        }
        if (elim > 0) return 0;
        if (elim < 0) return 1;
        return s < 0 ? 1 : 0;
    }
    default: {
        const int ub = c.single_node ? local_best : S->ub0;
        int mn = INT_MAX;
        for (int i = 0; i < k; i++) {
            if (cb[i] > 0) mn = std::min(mn, cb[i]);
            if (ce[i] > 0) mn = std::min(mn, ce[i]);
        }
        int minc = 0, elimc = 0;
        for (int i = 0; i < k; i++) {
            minc += (cb[i] == mn) - (ce[i] == mn);
            elimc += (cb[i] >= ub) - (ce[i] >= ub);
        }
        if (minc > 0) return 1;
        if (minc < 0) return 0;
        return elimc > 0 ? 0 : 1;
    }
    }
}

// Refinement bound of the child at row cell i scheduled in direction d:
// bound_fsp_weak::bornes_calculer (LB1', bound_fsp_weak.cpp:288-311) or
// bound_fsp_strong::bornes_calculer (Johnson LB2 over the JohnsonMode's
// pairs, bound_fsp_strong.cpp:50-162, :540-560; the exact makespan when at
// most one job stays unscheduled).
int Cpx::refine(int i, int d, int k) const
{
    const int *p = S->T.ptm.data();
    const int j = abs_flag(mat[line][i]);
    int fc[MAXM], bc[MAXM], rc[MAXM];
    const int *f = F[nf], *b = B[nb];
    if (d == 0) {
        int c = f[0] + p[j];
        fc[0] = c;
        for (int q = 1; q < m; q++) {
            c = std::max(f[q], c) + p[q * n + j];
            fc[q] = c;
        }
        for (int q = 0; q < m; q++) bc[q] = b[q];
    } else {
        int c = b[m - 1] + p[(m - 1) * n + j];
        bc[m - 1] = c;
        for (int q = m - 2; q >= 0; q--) {
            c = std::max(b[q], c) + p[q * n + j];
            bc[q] = c;
        }
        for (int q = 0; q < m; q++) fc[q] = f[q];
    }
    if (S->cfg.refine == 0) {
        for (int q = 0; q < m; q++) rc[q] = rem[q] - p[q * n + j];
        int t0 = fc[0] + rc[0];
        int lb = t0 + bc[0];
        for (int q = 1; q < m; q++) {
            const int t1 = std::max(fc[q] + rc[q], t0);
            lb = std::max(lb, t1 + bc[q]);
            t0 = t1;
        }
        return lb;
    }
    if (k <= 2) {
        const int o = abs_flag(mat[line][1 - i]);
        int c = fc[0] + p[o];
        int mk = c + bc[0];
        for (int q = 1; q < m; q++) {
            c = std::max(fc[q], c) + p[q * n + o];
            mk = std::max(mk, c + bc[q]);
        }
        return mk;
    }
    uint64_t unsched = 0;
    for (int c = 0; c < k; c++)
        if (c != i) unsched |= 1ull << abs_flag(mat[line][c]);
    int best = 0;
    for (int pr : S->pair_rows) {
        const int ab = S->T.pair_ab[pr];
        const int a = ab & 0xff, bm = ab >> 8;
        int t0 = fc[a], t1 = fc[bm];
        const int4 *row = S->T.jtab.data() + (size_t)pr * n;
        for (int r = 0; r < n; r++) {
            const int4 e = row[r];
            if ((unsched >> e.x) & 1ull) {
                t0 += e.y;
                t1 = std::max(t1, t0 + e.w) + e.z;
            }
        }
        best = std::max(best, std::max(t1 + bc[bm], t0 + bc[a]));
    }
    return best;
}

// Bound the children of the node at the current line, choose the direction,
// refine, prune and reverse (ivm_bound.cpp:141-236, :337-410, :610-676).
// The root is always bounded with LB1 (boundRoot, ivm_bound.cpp:273-330).
void Cpx::bound_node(bool root)
{
    const pbb_config &c = S->cfg;
    const int *p = S->T.ptm.data();
    const int L = line, k = n - L;
    int *row = mat[L];
    const int mode = root ? 0 : c.bound_mode;
    const bool both = c.branching > 0;
    int cb[MAXN], ce[MAXN], c1[MAXN], c2[MAXN];
    bool refined = false;
    int d = 0;
    if (mode != 1) {
        const int *f = F[nf], *b = B[nb];
        // bound_fsp_weak::addFrontAndBound / addBackAndBound (bound_fsp_weak.cpp:133-191)
        for (int i = 0; i < k; i++) {
            const int j = row[i];
            int t0 = f[0] + p[j];
            int lb = f[0] + rem[0] + b[0];
            for (int q = 1; q < m; q++) {
                const int t1 = std::max(f[q], t0);
                lb = std::max(lb, t1 + rem[q] + b[q]);
                t0 = t1 + p[q * n + j];
            }
            cb[i] = lb;
            if (both) {
                t0 = b[m - 1] + p[(m - 1) * n + j];
                lb = f[m - 1] + rem[m - 1] + b[m - 1];
                for (int q = m - 2; q >= 0; q--) {
                    const int t1 = std::max(t0, b[q]);
                    lb = std::max(lb, f[q] + rem[q] + t1);
                    t0 = t1 + p[q * n + j];
                }
                ce[i] = lb;
            } else {
                ce[i] = 0;  // costsEnd[WEAK] is never written with one direction
            }
        }
        if (!root) {
            sb += (uint64_t)(both ? 2 * k : k);
        }
        d = choose(cb, ce, k);
        for (int i = 0; i < k; i++) c1[i] = d ? ce[i] : cb[i];
        if (mode == 2) {
            // computeStrongBounds in the chosen direction: children LB1 keeps
            int nref = 0;
            for (int i = 0; i < k; i++) {
                if (!gt_best(c1[i])) {
                    c2[i] = refine(i, d, k);
                    nref++;
                } else {
                    c2[i] = 0;
                }
            }
            jb += (uint64_t)nref;
            jhist[L] += (uint64_t)nref;
            refined = true;
        }
    } else {
        // strongBoundPrune: every child refined in both directions
        int csF[MAXN], csB[MAXN];
        for (int i = 0; i < k; i++) {
            csF[i] = refine(i, 0, k);
            csB[i] = refine(i, 1, k);
        }
        jb += (uint64_t)(2 * k);
        jhist[L] += (uint64_t)(2 * k);
        d = choose(csF, csB, k);
        for (int i = 0; i < k; i++) c1[i] = c2[i] = d ? csB[i] : csF[i];
        refined = true;
    }
    for (int i = 0; i < k; i++)
        if (gt_best(c1[i]) || (refined && gt_best(c2[i]))) row[i] = neg_flag(row[i]);
    const int prev = L > 0 ? dir[L - 1] : 0;
    if (prev != d) std::reverse(row, row + k);
    dir[L] = d;
}

// ivm_bound::boundLeaf (ivm_bound.cpp:238-271) + solution::update (strict <)
void Cpx::leaf()
{
    const int *p = S->T.ptm.data();
    leaves++;
    const int r = mat[line][0];
    const int *f = F[nf], *b = B[nb];
    int c = f[0] + p[r];
    int cost = c + b[0];
    for (int q = 1; q < m; q++) {
        c = std::max(f[q], c) + p[q * n + r];
        cost = std::max(cost, c + b[q]);
    }
    if (!gt_best(cost)) {
        if (cost < local_best) {
            // the schedule: front jobs in order, the leaf job, back jobs
            int perm[MAXN];
            int l1 = 0, l2 = n - 1;
            for (int l = 0; l < line; l++) {
                const int j = abs_flag(mat[l][pos[l]]);
                if (dir[l] == 0) perm[l1++] = j;
                else perm[l2--] = j;
            }
            perm[l1] = r;
            S->offer(cost, perm);
        }
        local_best = cost;
    }
    mat[line][0] = neg_flag(r);
}

// ivm::goDown + generateLine (ivm.cpp:73-131): the new row is the current
// row without the cell at pos, flags cleared
void Cpx::go_down(bool explore)
{
    const int L = line, c = pos[L];
    const int *src = mat[L];
    int *dst = mat[L + 1];
    push(abs_flag(src[c]), dir[L]);
    for (int i = 0; i < c; i++) dst[i] = abs_flag(src[i]);
    for (int i = c; i < n - L - 1; i++) dst[i] = abs_flag(src[i + 1]);
    line = L + 1;
    if (explore) {
        pos[line] = 0;
        dir[line] = 0;
    }
}

// ivm::goUp (ivm.cpp:56-70): back to the parent, whose cell is marked visited
void Cpx::go_up()
{
    if (line > 0) {
        pos[line] = 0;
        line--;
        int &cell = mat[line][pos[line]];
        pop(abs_flag(cell), dir[line]);
        cell = neg_flag(cell);
    } else {
        pos[0]++;
    }
}

// sequentialbb::next (sequentialbb.cpp:110-191): true when an interior node
// was decomposed and bounded, false when the interval is exhausted
bool Cpx::next()
{
    local_best = (int)std::min<long long>(local_best, S->best.load(std::memory_order_relaxed));
    while (before_end()) {
        const int L = line;
        if (pos[L] >= n - L) {
            go_up();
            continue;
        }
        if (mat[L][pos[L]] < 0) {
            pos[L]++;
            continue;
        }
        dec++;
        go_down(true);
        hist[line]++;
        if (line == n - 1) {
            leaf();
            continue;
        }
        bound_node(false);
        return true;
    }
    return false;
}

// sequentialbb::unfold (sequentialbb.cpp:206-264): bounded, uncounted
void Cpx::unfold()
{
    while (line < n - 2) {
        if (mat[line][pos[line]] < 0) {
            for (int i = line + 1; i < n; i++) pos[i] = 0;
            break;
        }
        go_down(false);
        bound_node(false);
    }
}

// sequentialbb::initAtInterval (sequentialbb.cpp:83-108) from the bounded root row
bool Cpx::start(const int *p, const int *e)
{
    line = 0;
    for (int i = 0; i < n; i++) {
        pos[i] = p[i];
        end[i] = e[i];
        mat[0][i] = S->root_row[i];
    }
    dir[0] = S->root_dir;
    reset_stacks();
    local_best = (int)std::min<long long>(INT_MAX, S->best.load());
    if (!before_end()) return false;
    unfold();
    return true;
}

void Cpx::rebuild_stacks()
{
    reset_stacks();
    for (int l = 0; l < line; l++) push(abs_flag(mat[l][pos[l]]), dir[l]);
}

// bbthread::shareWork + ivm::cuttingPosition (bbthread.cpp:39-72,
// ivm.cpp:188-237): the thief T takes the upper half of the explorable
// cells after pos on the first line l (l <= line, l < n - margin) whose pos
// differs from end; rows 0..l, the directions and end are copied, no
// re-bounding, so node counts are unchanged.  A line whose cells after pos
// are all pruned or visited holds no work: its end is trimmed to pos (deeper
// ends become the maximum, which covers the same live nodes) and the split
// moves down, as the engine's donations do.
bool share(Cpx &V, Cpx &T, int margin)
{
    const int n = V.n;
    int l = 0, expl = 0;
    for (; l <= V.line && l <= n - 1 - margin; l++) {
        if (V.pos[l] == V.end[l]) continue;
        expl = 0;
        for (int i = V.pos[l] + 1; i <= V.end[l]; i++) expl += V.mat[l][i] >= 0;
        if (expl > 0) break;
        V.end[l] = V.pos[l];
        for (int i = l + 1; i < n; i++) V.end[i] = n - 1 - i;
    }
    if (l > V.line || l > n - 1 - margin || expl == 0) return false;
    const int keep = expl / 2;
    int cut = V.pos[l] + 1, kept = 0;
    while (kept < keep) {
        kept += V.mat[l][cut] >= 0;
        cut++;
    }
    for (int i = 0; i <= l; i++) {
        std::memcpy(T.mat[i], V.mat[i], sizeof(int) * (size_t)n);
        T.dir[i] = V.dir[i];
    }
    for (int i = 0; i < l; i++) T.pos[i] = V.pos[i];
    for (int i = 0; i < n; i++) T.end[i] = V.end[i];
    T.pos[l] = cut;
    for (int i = l + 1; i < n; i++) T.pos[i] = 0;
    T.line = l;
    T.rebuild_stacks();
    V.end[l] = cut - 1;
    for (int i = l + 1; i < n; i++) V.end[i] = n - 1 - i;
    return true;
}

void flush(CpuShared *S, Cpx &X, uint64_t &reported)
{
    if (X.dec > reported) {
        S->run_dec.fetch_add(X.dec - reported, std::memory_order_relaxed);
        reported = X.dec;
    }
}

// one host thread: explore, donate to waiting explorers, wait when dry
void worker(CpuShared *S, Cpx *X)
{
    uint64_t reported = X->dec;
    unsigned iter = 0;
    while (!S->stop.load(std::memory_order_relaxed)) {
        if (X->has_work) {
            if (!X->next()) {
                X->has_work = false;
                flush(S, *X, reported);
                // the last explorer to run dry ends the search
                if (S->holders.fetch_sub(1) == 1) S->done.store(1);
                continue;
            }
            if ((++iter & 15u) == 0u && S->nwaiting.load(std::memory_order_relaxed) > 0) {
                int thief = -1;
                {
                    std::lock_guard<std::mutex> g(S->qmu);
                    if (!S->waiting.empty()) {
                        thief = S->waiting.back();
                        S->waiting.pop_back();
                        S->nwaiting.fetch_sub(1);
                    }
                }
                if (thief >= 0) {
                    Cpx &T = *S->ex[thief];
                    if (share(*X, T, S->steal_margin)) {
                        S->holders.fetch_add(1);
                        T.handed.store(1, std::memory_order_release);
                    } else {
                        std::lock_guard<std::mutex> g(S->qmu);
                        S->waiting.push_back(thief);
                        S->nwaiting.fetch_add(1);
                    }
                }
            }
            if ((iter & 63u) == 0u) {
                flush(S, *X, reported);
                if (S->node_budget && S->run_dec.load(std::memory_order_relaxed) >= S->node_budget) S->stop.store(1);
                if (S->has_deadline && Clock::now() >= S->deadline) S->stop.store(1);
                if (S->interrupt.load(std::memory_order_relaxed)) S->stop.store(1);
            }
            continue;
        }
        // dry: wait for a donation, the end of the search or a stop
        if (S->done.load()) break;
        X->handed.store(0);
        {
            std::lock_guard<std::mutex> g(S->qmu);
            S->waiting.push_back(X->id);
            S->nwaiting.fetch_add(1);
        }
        while (!X->handed.load(std::memory_order_acquire)) {
            if (S->done.load() || S->stop.load(std::memory_order_relaxed)) break;
            if (S->has_deadline && Clock::now() >= S->deadline) S->stop.store(1);
            if (S->interrupt.load(std::memory_order_relaxed)) S->stop.store(1);
            std::this_thread::yield();
        }
        if (X->handed.load(std::memory_order_acquire)) {
            X->has_work = true;
            X->local_best = (int)std::min<long long>(X->local_best, S->best.load());
            continue;
        }
        // Leave the queue.  A donor that already took this explorer off it
        // either hands it an interval or puts it back, without stopping in
        // between, so wait for one of the two.
        while (true) {
            {
                std::lock_guard<std::mutex> g(S->qmu);
                auto it = std::find(S->waiting.begin(), S->waiting.end(), X->id);
                if (it != S->waiting.end()) {
                    S->waiting.erase(it);
                    S->nwaiting.fetch_sub(1);
                    break;
                }
            }
            if (X->handed.load(std::memory_order_acquire)) {
                X->has_work = true;
                break;
            }
            std::this_thread::yield();
        }
        break;
    }
    flush(S, *X, reported);
}

}  // namespace

struct pbb_cpu {
    CpuShared S;
    std::vector<Cpx *> ex;
    int threads = 0;
    bool initialized = false;
    unsigned long long exported_dec = 0, exported_leaves = 0;
};

static void cpu_free(pbb_cpu *c)
{
    for (Cpx *x : c->ex) delete x;
    delete c;
}

extern "C" int pbb_cpu_create(const int32_t *ptm, int n, int m, const pbb_config *cfg, int threads, int ub,
                              const int *ub_perm, pbb_cpu **out)
{
    if (!out || !cfg) return fail(PBB_E_ARG, "null argument");
    *out = nullptr;
    int rc = check_shape(ptm, n, m);
    if (rc) return rc;
    if ((rc = validate(cfg))) return rc;
    if (ub < 0) return fail(PBB_E_ARG, "incumbent cost %d is negative (INT_MAX = none)", ub);
    if (ub_perm && !is_perm(ub_perm, n)) return fail(PBB_E_ARG, "incumbent order is not a permutation of 0..%d", n - 1);
    if (threads < 0 || threads > 1024) return fail(PBB_E_ARG, "threads %d outside [0, 1024]", threads);
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    pbb_cpu *c = new pbb_cpu();
    CpuShared &S = c->S;
    build_tables(ptm, n, m, S.T);
    S.cfg = *cfg;
    S.n = n;
    S.m = m;
    S.steal_margin = cfg->steal_margin;
    S.tot.assign(m, 0);
    for (int q = 0; q < m; q++)
        for (int j = 0; j < n; j++) S.tot[q] += ptm[q * n + j];
    // the JohnsonMode's machine pairs (bound_fsp_strong.cpp:133-135, :282-297)
    for (int p = 0; p < S.T.npairs; p++) {
        const int a = S.T.pair_ab[p] & 0xff, b = S.T.pair_ab[p] >> 8;
        if (cfg->johnson_pairs == 1 && b < m - 1) continue;
        if (cfg->johnson_pairs == 2 && b != a + 1) continue;
        S.pair_rows.push_back(p);
    }
    S.best.store(ub);
    S.best_cost = ub;
    S.ub0 = ub;
    S.best_perm.resize(n);
    for (int i = 0; i < n; i++) S.best_perm[i] = ub_perm ? ub_perm[i] : i;
    S.root_perm.resize(n);
    for (int i = 0; i < n; i++) S.root_perm[i] = i;
    c->threads = threads;
    for (int t = 0; t < threads; t++) {
        Cpx *x = new Cpx();
        x->S = &S;
        x->id = t;
        x->n = n;
        x->m = m;
        std::memset(x->hist, 0, sizeof x->hist);
        std::memset(x->jhist, 0, sizeof x->jhist);
        c->ex.push_back(x);
    }
    S.ex = c->ex;
    *out = c;
    return PBB_OK;
}

extern "C" int pbb_cpu_destroy(pbb_cpu *c)
{
    if (c) cpu_free(c);
    return PBB_OK;
}

extern "C" int pbb_cpu_set_root(pbb_cpu *c, const int *root_perm)
{
    if (!c || !root_perm) return fail(PBB_E_ARG, "null argument");
    if (!is_perm(root_perm, c->S.n)) return fail(PBB_E_ARG, "root order is not a permutation");
    c->S.root_perm.assign(root_perm, root_perm + c->S.n);
    c->S.root_bounded = false;
    return PBB_OK;
}

// ivm_bound::boundRoot (ivm_bound.cpp:273-330): the root row bounded with
// LB1 against the incumbent, once, shared by every explorer
static void cpu_bound_root(pbb_cpu *c)
{
    CpuShared &S = c->S;
    if (S.root_bounded) return;
    Cpx &X = *c->ex[0];
    S.ub0 = (int)S.best.load();
    X.line = 0;
    for (int i = 0; i < S.n; i++) X.mat[0][i] = S.root_perm[i];
    X.reset_stacks();
    X.local_best = S.ub0;
    X.bound_node(true);
    S.root_row.assign(X.mat[0], X.mat[0] + S.n);
    S.root_dir = X.dir[0];
    S.root_bounded = true;
}

static void cpu_reset_counters(pbb_cpu *c)
{
    for (Cpx *x : c->ex) {
        x->dec = x->leaves = x->sb = x->jb = 0;
        std::memset(x->hist, 0, sizeof x->hist);
        std::memset(x->jhist, 0, sizeof x->jhist);
    }
    c->exported_dec = c->exported_leaves = 0;
}

extern "C" int pbb_cpu_init_intervals(pbb_cpu *c, int nb, const int *ids, const int *pos, const int *end)
{
    if (!c || nb < 0 || (nb > 0 && (!ids || !pos || !end))) return fail(PBB_E_ARG, "bad arguments");
    const int n = c->S.n;
    if (nb > c->threads) return fail(PBB_E_CAPACITY, "cannot handle more than %d intervals", c->threads);
    std::vector<int> seen(c->threads, 0);
    for (int k = 0; k < nb; k++) {
        if (ids[k] < 0 || ids[k] >= c->threads || seen[ids[k]]++)
            return fail(PBB_E_ARG, "interval id %d invalid or repeated", ids[k]);
        for (int i = 0; i < n; i++) {
            const int lim = n - i;
            if (pos[k * n + i] < 0 || pos[k * n + i] >= lim || end[k * n + i] < 0 || end[k * n + i] >= lim)
                return fail(PBB_E_ARG, "interval %d: factoradic digit %d out of range", k, i);
        }
    }
    cpu_bound_root(c);
    cpu_reset_counters(c);
    for (Cpx *x : c->ex) x->has_work = false;
    int holders = 0;
    for (int k = 0; k < nb; k++) {
        Cpx &X = *c->ex[ids[k]];
        X.has_work = X.start(pos + (size_t)k * n, end + (size_t)k * n);
        holders += X.has_work;
    }
    c->S.holders.store(holders);
    c->S.done.store(holders == 0 ? 1 : 0);
    c->initialized = true;
    return PBB_OK;
}

extern "C" int pbb_cpu_init_full(pbb_cpu *c)
{
    if (!c) return fail(PBB_E_ARG, "null co-worker");
    const int n = c->S.n;
    std::vector<int> p(n, 0), e(n);
    for (int i = 0; i < n; i++) e[i] = n - 1 - i;
    const int id = 0;
    return pbb_cpu_init_intervals(c, 1, &id, p.data(), e.data());
}

extern "C" int pbb_cpu_run(pbb_cpu *c, uint64_t node_budget, double time_budget_s, int *all_done)
{
    if (!c) return fail(PBB_E_ARG, "null co-worker");
    if (!c->initialized) return fail(PBB_E_STATE, "pbb_cpu_run before pbb_cpu_init_full/pbb_cpu_init_intervals");
    CpuShared &S = c->S;
    if (S.done.load() || S.holders.load() == 0) {
        if (all_done) *all_done = 1;
        return PBB_OK;
    }
    S.stop.store(0);
    S.interrupt.store(0);
    S.run_dec.store(0);
    S.node_budget = node_budget;
    S.has_deadline = time_budget_s > 0;
    if (S.has_deadline)
        S.deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(time_budget_s));
    S.waiting.clear();
    S.nwaiting.store(0);
    std::vector<std::thread> th;
    th.reserve(c->threads);
    for (Cpx *x : c->ex) th.emplace_back(worker, &S, x);
    for (auto &t : th) t.join();
    if (all_done) *all_done = S.done.load() ? 1 : 0;
    return PBB_OK;
}

extern "C" int pbb_cpu_get_intervals(pbb_cpu *c, int *pos, int *end, int *ids, int *nb, int max, uint64_t *decomposed,
                                     uint64_t *leaves)
{
    if (!c || !pos || !end || !ids || !nb) return fail(PBB_E_ARG, "null argument");
    if (max < c->threads) return fail(PBB_E_CAPACITY, "buffer too small (%d < %d)", max, c->threads);
    const int n = c->S.n;
    int k = 0;
    uint64_t d = 0, l = 0;
    for (Cpx *x : c->ex) {
        d += x->dec;
        l += x->leaves;
        if (!x->has_work || !x->before_end()) continue;
        ids[k] = x->id;
        for (int i = 0; i < n; i++) {
            pos[(size_t)k * n + i] = x->pos[i];
            end[(size_t)k * n + i] = x->end[i];
        }
        k++;
    }
    *nb = k;
    if (decomposed) *decomposed = d - c->exported_dec;
    if (leaves) *leaves = l - c->exported_leaves;
    c->exported_dec = d;
    c->exported_leaves = l;
    return PBB_OK;
}

extern "C" int pbb_cpu_get_incumbent(pbb_cpu *c, int *cost, int *perm)
{
    if (!c || !cost) return fail(PBB_E_ARG, "null argument");
    std::lock_guard<std::mutex> g(c->S.mu);
    *cost = c->S.best_cost;
    if (perm) std::copy(c->S.best_perm.begin(), c->S.best_perm.end(), perm);
    return PBB_OK;
}

extern "C" int pbb_cpu_set_incumbent(pbb_cpu *c, int cost, const int *perm)
{
    if (!c) return fail(PBB_E_ARG, "null co-worker");
    if (cost < 0) return fail(PBB_E_ARG, "incumbent cost %d is negative", cost);
    if (perm && !is_perm(perm, c->S.n)) return fail(PBB_E_ARG, "incumbent order is not a permutation");
    std::vector<int> p(c->S.n);
    {
        std::lock_guard<std::mutex> g(c->S.mu);
        if (cost >= c->S.best_cost) return PBB_OK;  // solution::update keeps strict improvements
        p = c->S.best_perm;
    }
    c->S.offer(cost, perm ? perm : p.data());
    return PBB_OK;
}

extern "C" int pbb_cpu_get_stats(pbb_cpu *c, pbb_stats *st)
{
    if (!c || !st) return fail(PBB_E_ARG, "null argument");
    std::memset(st, 0, sizeof *st);
    for (Cpx *x : c->ex) {
        st->decomposed += x->dec;
        st->leaves += x->leaves;
        st->simple_bounds += x->sb;
        st->johnson_bounds += x->jb;
        st->active += x->has_work && x->before_end();
        for (int d = 0; d <= c->S.n && d <= PBB_MAX_JOBS + 1; d++) {
            st->depth_histo[d] += x->hist[d];
            st->refine_histo[d] += x->jhist[d];
        }
    }
    st->explorers = c->threads;
    return PBB_OK;
}

extern "C" int pbb_cpu_interrupt(pbb_cpu *c)
{
    if (!c) return fail(PBB_E_ARG, "null co-worker");
    c->S.interrupt.store(1);
    return PBB_OK;
}

// ------------------------------------------------------------- hybrid
extern "C" int pbb_hybrid_run(pbb_engine *e, pbb_cpu *c, double period_s, double time_budget_s,
                              uint64_t *gpu_decomposed, uint64_t *cpu_decomposed, int *all_done)
{
    if (!e || !c) return fail(PBB_E_ARG, "null argument");
    if (!(period_s > 0)) return fail(PBB_E_ARG, "period must be positive");
    if (!c->initialized) {
        int rc = pbb_cpu_init_intervals(c, 0, nullptr, nullptr, nullptr);
        if (rc) return rc;
    }
    const int n = c->S.n;
    pbb_stats gs;
    int rc = pbb_get_stats(e, &gs);
    if (rc) return rc;
    const int E = gs.explorers;
    std::vector<int> gpos((size_t)E * n), gend((size_t)E * n), gids(E);
    std::vector<int> cpos((size_t)c->threads * n), cend((size_t)c->threads * n), cids(c->threads);
    uint64_t gdec = 0, cdec = 0;
    const auto t0 = Clock::now();
    int gdone = 0, cdone = 0;
    while (true) {
        const double el = std::chrono::duration<double>(Clock::now() - t0).count();
        if (time_budget_s > 0 && el >= time_budget_s) break;
        // the host explorers ran dry: move part of the engine's intervals to them
        pbb_stats cs;
        pbb_cpu_get_stats(c, &cs);
        if (cs.active == 0 && !gdone) {
            int nb = 0;
            uint64_t dd = 0;
            if ((rc = pbb_get_intervals(e, gpos.data(), gend.data(), gids.data(), &nb, E, &dd, nullptr))) return rc;
            gdec += dd;
            if (nb == 0) {
                gdone = 1;
            } else {
                // the host takes the last intervals, at most one per thread and a
                // quarter of them; the engine keeps the rest
                const int give = std::min(c->threads, std::max(1, nb / 4));
                const int keep = nb - give;
                if (give > 0 && keep > 0) {
                    std::vector<int> hid(give), kid(keep);
                    for (int i = 0; i < give; i++) hid[i] = i;
                    for (int i = 0; i < keep; i++) kid[i] = i;
                    uint64_t cd = 0;
                    int cn = 0;
                    pbb_cpu_get_intervals(c, cpos.data(), cend.data(), cids.data(), &cn, c->threads, &cd, nullptr);
                    cdec += cd;
                    if ((rc = pbb_cpu_init_intervals(c, give, hid.data(), gpos.data() + (size_t)keep * n,
                                                     gend.data() + (size_t)keep * n)))
                        return rc;
                    if ((rc = pbb_init_intervals(e, keep, kid.data(), gpos.data(), gend.data()))) return rc;
                } else {
                    std::vector<int> kid(nb);
                    for (int i = 0; i < nb; i++) kid[i] = i;
                    if ((rc = pbb_init_intervals(e, nb, kid.data(), gpos.data(), gend.data()))) return rc;
                }
            }
        }
        // one period: the engine in this thread, the host explorers beside it
        const double left = time_budget_s > 0 ? time_budget_s - el : period_s;
        const double per = std::min(period_s, std::max(1e-3, left));
        int crc = PBB_OK, cd = 0;
        std::thread ct([&]() { crc = pbb_cpu_run(c, 0, per, &cd); });
        int gd = 0;
        rc = gdone ? PBB_OK : pbb_run(e, 0, per, &gd);
        ct.join();
        if (rc) return rc;
        if (crc) return crc;
        gdone = gdone || gd;
        cdone = cd;
        // incumbents: each side gets the other's improvement
        int gc = 0, ccost = 0;
        std::vector<int> gp(n), cp(n);
        if ((rc = pbb_get_incumbent(e, &gc, gp.data()))) return rc;
        if ((rc = pbb_cpu_get_incumbent(c, &ccost, cp.data()))) return rc;
        if (ccost < gc && (rc = pbb_set_incumbent(e, ccost, cp.data()))) return rc;
        if (gc < ccost && (rc = pbb_cpu_set_incumbent(c, gc, gp.data()))) return rc;
        if (gdone && cdone) break;
    }
    // what the engine and the host decomposed since the call (the engine's
    // counters restart at each re-initialisation, so the exports are added;
    // the last export collects what is left)
    {
        int nb = 0;
        uint64_t dd = 0;
        if ((rc = pbb_get_intervals(e, gpos.data(), gend.data(), gids.data(), &nb, E, &dd, nullptr))) return rc;
        gdec += dd;
    }
    pbb_stats c2;
    pbb_cpu_get_stats(c, &c2);
    cdec += c2.decomposed - c->exported_dec;
    if (gpu_decomposed) *gpu_decomposed = gdec;
    if (cpu_decomposed) *cpu_decomposed = cdec;
    if (all_done) *all_done = gdone && cdone;
    return PBB_OK;
}
