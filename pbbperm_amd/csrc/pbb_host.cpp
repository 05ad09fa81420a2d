// pbb_host.cpp -- host side of the MI355X IVM engine: the C ABI of
// include/pbb_engine.h.  Instance tables are built here and uploaded once;
// everything on the node path runs in the HIP kernels of pbb_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstddef>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pbb_engine.h"
#include "pbb_internal.h"
#include "pbb_kernels.h"

using namespace pbb;

// ------------------------------------------------------------ device tables
void pbb::build_tables(const int32_t *ptm, int n, int m, HostTables &T)
{
    T.n = n;
    T.m = m;
    T.mp = (m + 3) & ~3;
    T.ptm.assign(ptm, ptm + (size_t)n * m);
    T.ptmT.assign((size_t)n * T.mp, 0);
    for (int j = 0; j < n; j++)
        for (int k = 0; k < m; k++) T.ptmT[(size_t)j * T.mp + k] = ptm[k * n + j];
    // bound_fsp_weak::fillMinTimeArrDep (bound_fsp_weak.cpp:248-286)
    T.min_arr.assign(m, INT_MAX);
    T.min_dep.assign(m, INT_MAX);
    T.min_arr[0] = 0;
    T.min_dep[m - 1] = 0;
    for (int j = 0; j < n; j++) {
        int tail = 0;
        for (int k = m - 2; k >= 0; k--) {
            tail += ptm[(k + 1) * n + j];
            T.min_dep[k] = std::min(T.min_dep[k], tail);
        }
        int head = 0;
        for (int k = 1; k < m; k++) {
            head += ptm[(k - 1) * n + j];
            T.min_arr[k] = std::min(T.min_arr[k], head);
        }
    }
    // machine pairs and Johnson/Mitten orders (bound_fsp_strong.cpp:264-436).
    // Ties inside a key leave the two-machine optimum unchanged; a stable
    // order by (set, key) is used.
    T.npairs = m * (m - 1) / 2;
    T.pair_ab.clear();
    T.jtab.assign((size_t)T.npairs * n, int4{0, 0, 0, 0});
    int p = 0;
    for (int a = 0; a < m - 1; a++)
        for (int b = a + 1; b < m; b++, p++) {
            T.pair_ab.push_back(a | (b << 8));
            std::vector<int> lag(n), set(n), key(n), ord(n);
            for (int j = 0; j < n; j++) {
                int l = 0;
                for (int k = a + 1; k < b; k++) l += ptm[k * n + j];
                lag[j] = l;
                bool first = ptm[a * n + j] < ptm[b * n + j];
                set[j] = first ? 1 : 2;
                key[j] = (first ? ptm[a * n + j] : ptm[b * n + j]) + l;
                ord[j] = j;
            }
            std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
                if (set[x] != set[y]) return set[x] < set[y];
                return set[x] == 1 ? key[x] < key[y] : key[x] > key[y];
            });
            for (int i = 0; i < n; i++) {
                int j = ord[i];
                T.jtab[(size_t)p * n + i] = int4{j, ptm[a * n + j], ptm[b * n + j], lag[j]};
            }
        }
}

struct DevTables {
    int *ptm = nullptr, *ptmT = nullptr, *min_arr = nullptr, *min_dep = nullptr, *pair_ab = nullptr;
    int4 *jtab = nullptr;
    DevInst I{};
};

static int upload_tables(const HostTables &T, DevTables &D)
{
    HIPCHK(hipMalloc(&D.ptm, T.ptm.size() * 4));
    HIPCHK(hipMalloc(&D.ptmT, T.ptmT.size() * 4));
    HIPCHK(hipMalloc(&D.min_arr, T.m * 4));
    HIPCHK(hipMalloc(&D.min_dep, T.m * 4));
    HIPCHK(hipMalloc(&D.pair_ab, std::max(1, T.npairs) * 4));
    HIPCHK(hipMalloc(&D.jtab, std::max<size_t>(1, T.jtab.size()) * sizeof(int4)));
    HIPCHK(hipMemcpy(D.ptm, T.ptm.data(), T.ptm.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(D.ptmT, T.ptmT.data(), T.ptmT.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(D.min_arr, T.min_arr.data(), T.m * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(D.min_dep, T.min_dep.data(), T.m * 4, hipMemcpyHostToDevice));
    if (T.npairs) {
        HIPCHK(hipMemcpy(D.pair_ab, T.pair_ab.data(), T.npairs * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(D.jtab, T.jtab.data(), T.jtab.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    D.I.n = T.n;
    D.I.m = T.m;
    D.I.mp = T.mp;
    D.I.ptm = D.ptm;
    D.I.ptmT = D.ptmT;
    D.I.min_arr = D.min_arr;
    D.I.min_dep = D.min_dep;
    D.I.npairs = T.npairs;
    D.I.pair_ab = D.pair_ab;
    D.I.jtab = D.jtab;
    return PBB_OK;
}

static void free_tables(DevTables &D)
{
    hipFree(D.ptm);
    hipFree(D.ptmT);
    hipFree(D.min_arr);
    hipFree(D.min_dep);
    hipFree(D.pair_ab);
    hipFree(D.jtab);
    D = DevTables{};
}

int pbb::check_device()
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count < 1) return fail(PBB_E_NODEVICE, "no HIP device available");
    return PBB_OK;
}

// ------------------------------------------------------------ bound library
struct pbb_bounds {
    HostTables T;
    DevTables D;
};

extern "C" int pbb_bounds_create(const int32_t *ptm, int n, int m, pbb_bounds **out)
{
    if (!out) return fail(PBB_E_ARG, "null out");
    int rc = check_shape(ptm, n, m);
    if (rc) return rc;
    if ((rc = check_device())) return rc;
    pbb_bounds *b = new pbb_bounds();
    build_tables(ptm, n, m, b->T);
    rc = upload_tables(b->T, b->D);
    if (rc) {
        free_tables(b->D);
        delete b;
        return rc;
    }
    *out = b;
    return PBB_OK;
}

extern "C" int pbb_bounds_destroy(pbb_bounds *b)
{
    if (!b) return PBB_OK;
    free_tables(b->D);
    delete b;
    return PBB_OK;
}

static int bounds_call(pbb_bounds *b, int which, int count, const int *perms, const int *l1,
                       const int *l2, int arg, int *o1, int *o2, void *stream)
{
    if (!b || count < 0 || !perms || !o1) return fail(PBB_E_ARG, "bad bound batch arguments");
    if (which != 3 && (!l1 || !l2)) return fail(PBB_E_ARG, "null l1/l2");
    if (which == 0 && arg && !o2) return fail(PBB_E_ARG, "null ce with both=1");
    if (count == 0) return PBB_OK;
    KArgs a{};
    a.I = b->D.I;
    a.rows = b->D.I.n;
    HIPCHK(launch_bounds(which, a, count, perms, l1, l2, arg, o1, o2, (hipStream_t)stream));
    return PBB_OK;
}

extern "C" int pbb_lb1_children_batch(pbb_bounds *b, int count, const int *perms, const int *l1,
                                      const int *l2, int both, int *cb, int *ce, void *stream)
{
    return bounds_call(b, 0, count, perms, l1, l2, both ? 1 : 0, cb, ce, stream);
}

extern "C" int pbb_lb1_node_batch(pbb_bounds *b, int count, const int *perms, const int *l1,
                                  const int *l2, int *lb, void *stream)
{
    return bounds_call(b, 1, count, perms, l1, l2, 0, lb, nullptr, stream);
}

extern "C" int pbb_lb2_node_batch(pbb_bounds *b, int count, const int *perms, const int *l1,
                                  const int *l2, int pairs, int *lb, void *stream)
{
    if (pairs < 0 || pairs > 2) return fail(PBB_E_ARG, "pairs mode %d", pairs);
    return bounds_call(b, 2, count, perms, l1, l2, pairs, lb, nullptr, stream);
}

extern "C" int pbb_makespan_batch(pbb_bounds *b, int count, const int *perms, int *cmax, void *stream)
{
    return bounds_call(b, 3, count, perms, nullptr, nullptr, 0, cmax, nullptr, stream);
}

// ------------------------------------------------------------------ engine
struct pbb_engine {
    pbb_config cfg;
    HostTables T;
    DevTables D;
    int n, m, E, blocks, ring_cap;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // device buffers
    int8_t *mat = nullptr, *pos = nullptr, *end = nullptr, *dir = nullptr;
    int *line = nullptr, *state = nullptr;
    unsigned long long *ring = nullptr;
    Ctl *ctl = nullptr;
    unsigned long long *best = nullptr;
    int *best_perm = nullptr, *root_row = nullptr, *root_dir = nullptr;
    int *host_flag = nullptr, *host_flag_d = nullptr;
    // pinned read-back of the control block and the incumbent word after a launch
    Ctl *ctl_h = nullptr;
    unsigned long long *best_h = nullptr;
    // host state
    std::vector<int> root_perm;
    int inc_cost;
    std::vector<int> inc_perm;
    int initial_ub;
    bool root_bounded = false;
    bool initialized = false;
    unsigned long long exported_dec = 0, exported_leaves = 0;
    // counters carried in from a checkpoint file (works::init_infile restores
    // simpleBounds, distributed/src/works.cpp:81-83; the unit's exploredNodes)
    unsigned long long dec_base = 0, sb_base = 0;
    // the device control block's done flag and node count as of the last
    // set-up or launch (nothing else writes them between launches)
    bool done_host = false;
    unsigned long long dec_dev = 0;
    // the team launch: its explorer count (0: none fits) and whether the
    // next pbb_run starts with it (set by pbb_init_full)
    int team_E = 0;
    bool team_pending = false;
    double last_ms = 0.0;
    int wall_khz = 100000;
};

static KArgs make_args(pbb_engine *e)
{
    KArgs a{};
    a.I = e->D.I;
    a.E = e->E;
    a.rows = e->n;
    a.mat = e->mat;
    a.pos = e->pos;
    a.end = e->end;
    a.dir = e->dir;
    a.line = e->line;
    a.state = e->state;
    a.ring = e->ring;
    a.ring_cap = e->ring_cap;
    a.ctl = e->ctl;
    a.best = e->best;
    a.best_perm = e->best_perm;
    a.root_row = e->root_row;
    a.root_dir = e->root_dir;
    a.root_row_out = e->root_row;
    a.root_dir_out = e->root_dir;
    a.host_flag = e->host_flag_d;
    a.bound_mode = e->cfg.bound_mode;
    a.branching = e->cfg.branching;
    a.refine = e->cfg.refine;
    a.pairs_mode = e->cfg.johnson_pairs;
    a.find_all = e->cfg.find_all;
    a.single_node = e->cfg.single_node;
    a.initial_ub = e->initial_ub;
    a.steal_margin = e->cfg.steal_margin;
    a.ramp = (unsigned)e->cfg.ramp_nodes;
    return a;
}

// Every copy of engine state is ordered on the engine's stream (launches are
// not followed by a wait of their own; the host waits where it reads).
static hipError_t copy_sync(pbb_engine *e, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    hipError_t r = hipMemcpyAsync(dst, src, bytes, kind, e->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    return r;
}

static int capacity(pbb_engine *e, int *explorers)
{
    KArgs a = make_args(e);
    int bpc = 0;
    HIPCHK(explore_occupancy(a, &bpc));
    if (bpc < 1) return fail(PBB_E_CAPACITY, "explorer kernel does not fit one workgroup per CU (LDS)");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    *explorers = bpc * prop.multiProcessorCount * PBB_EXPLORE_WAVES;
    return PBB_OK;
}

extern "C" int pbb_destroy(pbb_engine *e)
{
    if (!e) return PBB_OK;
    hipFree(e->mat);
    hipFree(e->pos);
    hipFree(e->end);
    hipFree(e->dir);
    hipFree(e->line);
    hipFree(e->state);
    hipFree(e->ring);
    hipFree(e->ctl);
    hipFree(e->best);
    hipFree(e->best_perm);
    hipFree(e->root_row);
    hipFree(e->root_dir);
    if (e->host_flag) hipHostFree(e->host_flag);
    if (e->ctl_h) hipHostFree(e->ctl_h);
    if (e->best_h) hipHostFree(e->best_h);
    if (e->ev0) hipEventDestroy(e->ev0);
    if (e->ev1) hipEventDestroy(e->ev1);
    if (e->stream) hipStreamDestroy(e->stream);
    free_tables(e->D);
    delete e;
    return PBB_OK;
}

extern "C" int pbb_create(const int32_t *ptm, int n, int m, const pbb_config *cfg, int ub,
                          const int *ub_perm, pbb_engine **out)
{
    if (!out || !cfg) return fail(PBB_E_ARG, "null argument");
    *out = nullptr;
    int rc = check_shape(ptm, n, m);
    if (rc) return rc;
    if ((rc = validate(cfg))) return rc;
    // an incumbent cost is a makespan: a negative one (PBB_UB_NEH of the pool
    // entry points, say) would prune every root child and report it as optimal
    if (ub < 0) return fail(PBB_E_ARG, "incumbent cost %d is negative (INT_MAX = none)", ub);
    if (ub_perm && !is_perm(ub_perm, n)) return fail(PBB_E_ARG, "incumbent order is not a permutation of 0..%d", n - 1);
    if ((rc = check_device())) return rc;
    pbb_engine *e = new pbb_engine();
    e->cfg = *cfg;
    e->n = n;
    e->m = m;
    e->inc_cost = ub;
    e->initial_ub = ub;
    e->inc_perm.resize(n);
    for (int i = 0; i < n; i++) e->inc_perm[i] = ub_perm ? ub_perm[i] : i;
    e->root_perm.resize(n);
    for (int i = 0; i < n; i++) e->root_perm[i] = i;
    build_tables(ptm, n, m, e->T);
#define TRY(x)                \
    do {                      \
        int r_ = (x);         \
        if (r_) {             \
            pbb_destroy(e);   \
            return r_;        \
        }                     \
    } while (0)
#define HTRY(x) TRY(([&]() -> int { HIPCHK(x); return PBB_OK; })())
    TRY(upload_tables(e->T, e->D));
    HTRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    HTRY(hipEventCreate(&e->ev0));
    HTRY(hipEventCreate(&e->ev1));
    int cap = 0;
    TRY(capacity(e, &cap));
    e->E = std::max(1, std::min(cfg->nb_ivms, cap));
    e->blocks = (e->E + PBB_EXPLORE_WAVES - 1) / PBB_EXPLORE_WAVES;
    {
        // the team: up to team_waves_max explorers whose LDS fits one workgroup
        const bool lb2 = cfg->bound_mode != 0 && cfg->refine == 1;
        const int ref = cfg->bound_mode == 0 ? 0 : (cfg->refine == 1 ? 2 : 1);
        int w = std::min(e->E, team_waves_max(ref));
        if (cfg->team_waves > 0) w = std::min(w, cfg->team_waves);
        while (w > 0 && team_lds_bytes(e->D.I, lb2, w) > 160 * 1024) w--;
        e->team_E = w;
    }
    e->ring_cap = 1;
    while (e->ring_cap < 2 * e->E) e->ring_cap <<= 1;
    size_t E = (size_t)e->E;
    HTRY(hipMalloc(&e->mat, E * n * 64));
    HTRY(hipMalloc(&e->pos, E * 64));
    HTRY(hipMalloc(&e->end, E * 64));
    HTRY(hipMalloc(&e->dir, E * 64));
    HTRY(hipMalloc(&e->line, E * 4));
    HTRY(hipMalloc(&e->state, E * 4));
    HTRY(hipMalloc(&e->ring, (size_t)e->ring_cap * 8));
    HTRY(hipMalloc(&e->ctl, sizeof(Ctl)));
    HTRY(hipMalloc(&e->best, 8));
    HTRY(hipMalloc(&e->best_perm, E * 64 * 4));
    HTRY(hipMalloc(&e->root_row, 64 * 4));
    HTRY(hipMalloc(&e->root_dir, 4));
    HTRY(hipHostMalloc((void **)&e->host_flag, 4, hipHostMallocMapped));
    *e->host_flag = 0;
    HTRY(hipHostGetDevicePointer((void **)&e->host_flag_d, e->host_flag, 0));
    HTRY(hipHostMalloc((void **)&e->ctl_h, sizeof(Ctl), hipHostMallocDefault));
    HTRY(hipHostMalloc((void **)&e->best_h, sizeof(unsigned long long), hipHostMallocDefault));
    int dev = 0;
    HTRY(hipGetDevice(&dev));
    HTRY(hipDeviceGetAttribute(&e->wall_khz, hipDeviceAttributeWallClockRate, dev));
    HTRY(hipMemsetAsync(e->state, 0, E * 4, e->stream));
    unsigned long long packed = ((unsigned long long)(unsigned)ub << 32) | 0xffffffffull;
    HTRY(copy_sync(e, e->best, &packed, 8, hipMemcpyHostToDevice));
#undef HTRY
#undef TRY
    *out = e;
    return PBB_OK;
}

extern "C" int pbb_set_root(pbb_engine *e, const int *root_perm)
{
    if (!e || !root_perm) return fail(PBB_E_ARG, "null argument");
    std::vector<int> seen(e->n, 0);
    for (int i = 0; i < e->n; i++) {
        if (root_perm[i] < 0 || root_perm[i] >= e->n || seen[root_perm[i]]++)
            return fail(PBB_E_ARG, "root order is not a permutation");
    }
    e->root_perm.assign(root_perm, root_perm + e->n);
    e->root_bounded = false;
    return PBB_OK;
}

// take the device incumbent word (read back already) into the host's copy
static int take_incumbent(pbb_engine *e, unsigned long long packed)
{
    int cost = (int)(packed >> 32);
    unsigned who = (unsigned)(packed & 0xffffffffull);
    if (cost < e->inc_cost && who != 0xffffffffu && (int)who < e->E) {
        std::vector<int> p(64);
        HIPCHK(copy_sync(e, p.data(), e->best_perm + (size_t)who * 64, 64 * 4, hipMemcpyDeviceToHost));
        e->inc_cost = cost;
        e->inc_perm.assign(p.begin(), p.begin() + e->n);
    }
    return PBB_OK;
}

static int sync_incumbent(pbb_engine *e)
{
    HIPCHK(copy_sync(e, e->best_h, e->best, 8, hipMemcpyDeviceToHost));
    return take_incumbent(e, *e->best_h);
}

// ivm_bound::boundRoot (ivm_bound.cpp:273-330) under the current incumbent:
// the host's copy is current here (every launch ends by reading the device
// word back, and pbb_set_incumbent writes both).  The launch is only
// enqueued; the set-up and explorer launches after it on the same stream
// read its output.
static int bound_root(pbb_engine *e)
{
    if (e->root_bounded) return PBB_OK;
    e->initial_ub = e->inc_cost;
    RootPerm rp{};
    std::copy(e->root_perm.begin(), e->root_perm.end(), rp.p);
    KArgs a = make_args(e);
    HIPCHK(launch_root(a, rp, e->stream));
    e->root_bounded = true;
    return PBB_OK;
}

// reset the explorer pool: explorers in `active` get (pos, end) and unfold,
// the others are registered idle.
static int setup_pool(pbb_engine *e, const std::vector<int> &ids, const int *pos, const int *end)
{
    const int n = e->n, E = e->E;
    std::vector<int> state(E, ST_IDLE);
    std::vector<int8_t> hp((size_t)E * 64, 0), he((size_t)E * 64, 0);
    for (size_t k = 0; k < ids.size(); k++) {
        int id = ids[k];
        state[id] = ST_UNFOLD;
        for (int i = 0; i < n; i++) {
            hp[(size_t)id * 64 + i] = (int8_t)pos[k * n + i];
            he[(size_t)id * 64 + i] = (int8_t)end[k * n + i];
        }
    }
    // The idle ring starts empty: idle explorers register themselves when
    // their wavefront runs (pbb_run empties it again before every launch).
    Ctl c;
    memset(&c, 0, sizeof c);
    c.work_count = (int)ids.size();
    c.done = ids.empty() ? 1 : 0;
    HIPCHK(copy_sync(e, e->state, state.data(), (size_t)E * 4, hipMemcpyHostToDevice));
    HIPCHK(copy_sync(e, e->pos, hp.data(), hp.size(), hipMemcpyHostToDevice));
    HIPCHK(copy_sync(e, e->end, he.data(), he.size(), hipMemcpyHostToDevice));
    HIPCHK(copy_sync(e, e->ctl, &c, sizeof c, hipMemcpyHostToDevice));
    e->exported_dec = e->exported_leaves = 0;
    e->dec_base = e->sb_base = 0;
    e->done_host = c.done != 0;
    e->dec_dev = 0;
    e->team_pending = false;
    e->initialized = true;
    return PBB_OK;
}

// The whole search space to explorer 0, the others idle: two launches
// enqueued on the engine's stream (root bounding, set-up), no copy and no
// wait; pbb_run's launches follow them on the stream.
extern "C" int pbb_init_full(pbb_engine *e)
{
    if (!e) return fail(PBB_E_ARG, "null engine");
    int rc = bound_root(e);
    if (rc) return rc;
    KArgs a = make_args(e);
    HIPCHK(launch_setup_full(a, e->stream));
    e->exported_dec = e->exported_leaves = 0;
    e->dec_base = e->sb_base = 0;
    e->done_host = false;
    e->dec_dev = 0;
    e->initialized = true;
    e->team_pending = e->team_E > 0 && e->cfg.team_nodes > 0;
    return PBB_OK;
}

extern "C" int pbb_init_intervals(pbb_engine *e, int nb, const int *ids, const int *pos, const int *end)
{
    if (!e || nb < 0 || (nb > 0 && (!ids || !pos || !end))) return fail(PBB_E_ARG, "bad arguments");
    if (nb > e->E) return fail(PBB_E_CAPACITY, "cannot handle more than %d intervals", e->E);
    std::vector<int> seen(e->E, 0), idv(nb);
    for (int k = 0; k < nb; k++) {
        if (ids[k] < 0 || ids[k] >= e->E || seen[ids[k]]++)
            return fail(PBB_E_ARG, "interval id %d invalid or repeated", ids[k]);
        for (int i = 0; i < e->n; i++) {
            int lim = e->n - i;
            if (pos[k * e->n + i] < 0 || pos[k * e->n + i] >= lim || end[k * e->n + i] < 0 ||
                end[k * e->n + i] >= lim)
                return fail(PBB_E_ARG, "interval %d: factoradic digit %d out of range", k, i);
        }
        idv[k] = ids[k];
    }
    int rc = bound_root(e);
    if (rc) return rc;
    return setup_pool(e, idv, pos, end);
}

// One launch: empty the idle ring and clear the stop word (explorers idle at
// the end of the previous launch register again when they run, so the ring
// never names a wavefront that is not resident), then the explorers, either
// the device-wide grid or the one-workgroup team; the control block and the
// incumbent word come back into pinned memory behind them, and the host
// waits once.  Returns the launch's device time in *ms.
static int launch_once(pbb_engine *e, bool team, uint64_t node_budget, double time_budget_s, float *ms)
{
    *e->host_flag = 0;
    KArgs a = make_args(e);
    if (team) a.E = e->team_E;
    a.node_budget = node_budget ? e->dec_dev + node_budget : 0;
    a.deadline = time_budget_s > 0 ? (unsigned long long)(time_budget_s * e->wall_khz * 1000.0) : 0;
    HIPCHK(launch_prelaunch(a, e->stream));
    HIPCHK(hipEventRecord(e->ev0, e->stream));
    if (team) HIPCHK(launch_team(a, e->stream));
    else HIPCHK(launch_explore(a, e->blocks, e->stream));
    HIPCHK(hipEventRecord(e->ev1, e->stream));
    HIPCHK(hipMemcpyAsync(e->ctl_h, e->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(e->best_h, e->best, 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipEventElapsedTime(ms, e->ev0, e->ev1));
    const Ctl &c = *e->ctl_h;
    e->done_host = c.done != 0;
    e->dec_dev = c.dec;
    if (c.err) return fail(PBB_E_HIP, "device error %d in the explorer kernel", c.err);
    return take_incumbent(e, *e->best_h);
}

// A search started by pbb_init_full runs first on the team (one workgroup)
// for up to cfg.team_nodes decomposed nodes; what is left then continues
// on the device-wide grid from the team's saved intervals.
extern "C" int pbb_run(pbb_engine *e, uint64_t node_budget, double time_budget_s, int *all_done)
{
    if (!e) return fail(PBB_E_ARG, "null engine");
    if (!e->initialized) return fail(PBB_E_STATE, "pbb_run before pbb_init_full/pbb_init_intervals");
    if (all_done) *all_done = 0;
    if (e->done_host) {
        if (all_done) *all_done = 1;
        return PBB_OK;
    }
    e->last_ms = 0.0;
    float ms = 0.f;
    if (e->team_pending) {
        e->team_pending = false;
        uint64_t tb = (uint64_t)e->cfg.team_nodes;
        if (node_budget && node_budget < tb) tb = node_budget;
        const uint64_t before = e->dec_dev;
        const auto t0 = std::chrono::steady_clock::now();
        int rc = launch_once(e, true, tb, time_budget_s, &ms);
        if (rc) return rc;
        e->last_ms += ms;
        if (e->done_host) {
            if (all_done) *all_done = 1;
            return PBB_OK;
        }
        const uint64_t used = e->dec_dev - before;
        if (node_budget) {
            if (used >= node_budget) return PBB_OK;
            node_budget -= used;
        }
        if (time_budget_s > 0) {
            time_budget_s -= std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (time_budget_s <= 0) return PBB_OK;
        }
    }
    int rc = launch_once(e, false, node_budget, time_budget_s, &ms);
    if (rc) return rc;
    e->last_ms += ms;
    if (all_done) *all_done = e->done_host ? 1 : 0;
    return PBB_OK;
}

// every active explorer's interval (state != idle)
static int export_intervals(pbb_engine *e, int *pos, int *end, int *ids, int *nb, int max)
{
    const int n = e->n, E = e->E;
    std::vector<int> st(E);
    std::vector<int8_t> hp((size_t)E * 64), he((size_t)E * 64);
    HIPCHK(copy_sync(e, st.data(), e->state, (size_t)E * 4, hipMemcpyDeviceToHost));
    HIPCHK(copy_sync(e, hp.data(), e->pos, hp.size(), hipMemcpyDeviceToHost));
    HIPCHK(copy_sync(e, he.data(), e->end, he.size(), hipMemcpyDeviceToHost));
    memset(pos, 0, sizeof(int) * (size_t)max * n);
    memset(end, 0, sizeof(int) * (size_t)max * n);
    int k = 0;
    for (int i = 0; i < E; i++)
        if (st[i] != ST_IDLE) {
            ids[k] = i;
            for (int d = 0; d < n; d++) {
                pos[(size_t)k * n + d] = hp[(size_t)i * 64 + d];
                end[(size_t)k * n + d] = he[(size_t)i * 64 + d];
            }
            k++;
        }
    *nb = k;
    return PBB_OK;
}

extern "C" int pbb_get_intervals(pbb_engine *e, int *pos, int *end, int *ids, int *nb, int max,
                                 uint64_t *decomposed, uint64_t *leaves)
{
    if (!e || !pos || !end || !ids || !nb) return fail(PBB_E_ARG, "null argument");
    if (max < e->E) return fail(PBB_E_CAPACITY, "buffer too small (%d < %d)", max, e->E);
    int rc = export_intervals(e, pos, end, ids, nb, max);
    if (rc) return rc;
    Ctl c;
    HIPCHK(copy_sync(e, &c, e->ctl, sizeof c, hipMemcpyDeviceToHost));
    if (decomposed) *decomposed = c.dec - e->exported_dec;
    if (leaves) *leaves = c.leaves - e->exported_leaves;
    e->exported_dec = c.dec;
    e->exported_leaves = c.leaves;
    return PBB_OK;
}

extern "C" int pbb_get_incumbent(pbb_engine *e, int *cost, int *perm)
{
    if (!e || !cost) return fail(PBB_E_ARG, "null argument");
    int rc = sync_incumbent(e);
    if (rc) return rc;
    *cost = e->inc_cost;
    if (perm) std::copy(e->inc_perm.begin(), e->inc_perm.end(), perm);
    return PBB_OK;
}

extern "C" int pbb_set_incumbent(pbb_engine *e, int cost, const int *perm)
{
    if (!e) return fail(PBB_E_ARG, "null engine");
    if (cost < 0) return fail(PBB_E_ARG, "incumbent cost %d is negative", cost);
    if (perm && !is_perm(perm, e->n)) return fail(PBB_E_ARG, "incumbent order is not a permutation of 0..%d", e->n - 1);
    int rc = sync_incumbent(e);
    if (rc) return rc;
    if (cost >= e->inc_cost) return PBB_OK;  // solution::update keeps strict improvements
    e->inc_cost = cost;
    if (perm) e->inc_perm.assign(perm, perm + e->n);
    unsigned long long packed = ((unsigned long long)(unsigned)cost << 32) | 0xffffffffull;
    HIPCHK(copy_sync(e, e->best, &packed, 8, hipMemcpyHostToDevice));
    return PBB_OK;
}

extern "C" int pbb_get_stats(pbb_engine *e, pbb_stats *st)
{
    if (!e || !st) return fail(PBB_E_ARG, "null argument");
    Ctl c;
    HIPCHK(copy_sync(e, &c, e->ctl, sizeof c, hipMemcpyDeviceToHost));
    memset(st, 0, sizeof *st);
    st->decomposed = e->dec_base + c.dec;
    st->leaves = c.leaves;
    st->simple_bounds = e->sb_base + c.sb;
    st->johnson_bounds = c.jb;
    st->explorers = e->E;
    std::vector<int> s(e->E);
    HIPCHK(copy_sync(e, s.data(), e->state, (size_t)e->E * 4, hipMemcpyDeviceToHost));
    int act = 0;
    for (int v : s) act += v != ST_IDLE;
    st->active = act;
    st->kernel_ms = e->last_ms;
    for (int i = 0; i <= PBB_MAX_JOBS + 1; i++) {
        st->depth_histo[i] = c.histo[i];
        st->refine_histo[i] = c.jhisto[i];
    }
    return PBB_OK;
}

extern "C" int pbb_get_deep_subproblems(pbb_engine *e, int *perms, int max, int *count)
{
    if (!e || !count || max < 0 || (max > 0 && !perms)) return fail(PBB_E_ARG, "bad arguments");
    *count = 0;
    if (!e->initialized || max == 0) return PBB_OK;
    const int n = e->n, m = e->m, E = e->E;
    std::vector<int> st(E), line(E);
    std::vector<int8_t> pos((size_t)E * 64), dir((size_t)E * 64), mat((size_t)E * n * 64);
    HIPCHK(copy_sync(e, st.data(), e->state, (size_t)E * 4, hipMemcpyDeviceToHost));
    HIPCHK(copy_sync(e, line.data(), e->line, (size_t)E * 4, hipMemcpyDeviceToHost));
    HIPCHK(copy_sync(e, pos.data(), e->pos, pos.size(), hipMemcpyDeviceToHost));
    HIPCHK(copy_sync(e, dir.data(), e->dir, dir.size(), hipMemcpyDeviceToHost));
    HIPCHK(copy_sync(e, mat.data(), e->mat, mat.size(), hipMemcpyDeviceToHost));
    std::vector<std::pair<int, int>> order;  // (cost, explorer)
    std::vector<int> sched((size_t)E * n);
    for (int i = 0; i < E; i++) {
        if (st[i] != ST_ACTIVE) continue;
        const int L = line[i];
        if (L < 0 || L >= n) continue;
        const int8_t *rows = mat.data() + (size_t)i * n * 64;
        int *sc = sched.data() + (size_t)i * n;
        int l1 = 0, l2 = n - 1;
        for (int l = 0; l < L; l++) {
            const int p = pos[(size_t)i * 64 + l];
            const int v = rows[l * 64 + p];
            const int job = v < 0 ? -v - 1 : v;
            if (dir[(size_t)i * 64 + l] == 0) sc[l1++] = job;
            else sc[l2--] = job;
        }
        for (int c = 0; c < n - L; c++) {
            const int v = rows[L * 64 + c];
            sc[l1 + c] = v < 0 ? -v - 1 : v;
        }
        order.emplace_back(host_partial_makespan(e->T.ptm.data(), n, m, sc, n), i);
    }
    std::sort(order.begin(), order.end());
    const int k = std::min((int)order.size(), max);
    for (int j = 0; j < k; j++)
        std::copy(sched.begin() + (size_t)order[j].second * n, sched.begin() + (size_t)(order[j].second + 1) * n,
                  perms + (size_t)j * n);
    *count = k;
    return PBB_OK;
}

extern "C" int pbb_interrupt(pbb_engine *e)
{
    if (!e) return fail(PBB_E_ARG, "null engine");
    __atomic_store_n(e->host_flag, 1, __ATOMIC_SEQ_CST);
    return PBB_OK;
}

extern "C" int pbb_query_explorers(int n, int m, const pbb_config *cfg, int *explorers)
{
    if (!cfg || !explorers) return fail(PBB_E_ARG, "null argument");
    std::vector<int32_t> ptm((size_t)n * m, 1);
    pbb_engine *e = nullptr;
    int rc = pbb_create(ptm.data(), n, m, cfg, INT_MAX, nullptr, &e);
    if (rc) return rc;
    *explorers = e->E;
    return pbb_destroy(e);
}

// ------------------------------------------------- plain stream replay
// The one-instance engine re-pointed at a new instance of the same shape:
// tables re-uploaded into the existing buffers, incumbent and root reset.
static int engine_retarget(pbb_engine *e, const int32_t *ptm, int ub, const int *ub_perm)
{
    build_tables(ptm, e->n, e->m, e->T);
    const HostTables &T = e->T;
    HIPCHK(copy_sync(e, e->D.ptm, T.ptm.data(), T.ptm.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(copy_sync(e, e->D.ptmT, T.ptmT.data(), T.ptmT.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(copy_sync(e, e->D.min_arr, T.min_arr.data(), T.m * 4, hipMemcpyHostToDevice));
    HIPCHK(copy_sync(e, e->D.min_dep, T.min_dep.data(), T.m * 4, hipMemcpyHostToDevice));
    if (T.npairs) {
        HIPCHK(copy_sync(e, e->D.pair_ab, T.pair_ab.data(), T.npairs * 4, hipMemcpyHostToDevice));
        HIPCHK(copy_sync(e, e->D.jtab, T.jtab.data(), T.jtab.size() * sizeof(int4), hipMemcpyHostToDevice));
    }
    e->inc_cost = ub;
    e->initial_ub = ub;
    for (int i = 0; i < e->n; i++) {
        e->root_perm[i] = i;
        e->inc_perm[i] = ub_perm ? ub_perm[i] : i;
    }
    e->root_bounded = false;
    e->initialized = false;
    *e->host_flag = 0;
    unsigned long long packed = ((unsigned long long)(unsigned)ub << 32) | 0xffffffffull;
    HIPCHK(copy_sync(e, e->best, &packed, 8, hipMemcpyHostToDevice));
    return PBB_OK;
}

extern "C" int pbb_plain_replay(const pbb_config *cfg, int count, const int32_t *ptms, const int *n,
                                const int *m, const int *ub, const double *arrival_s, double time_budget_s,
                                double request_budget_s, pbb_result *out, double *latency_ms, double *wall_s)
{
    if (!cfg || count < 0 || (count > 0 && (!ptms || !n || !m || !ub || !arrival_s || !out || !latency_ms)))
        return fail(PBB_E_ARG, "bad replay arguments");
    if (request_budget_s < 0) return fail(PBB_E_ARG, "negative request budget");
    std::vector<size_t> off(count + 1, 0);
    for (int i = 0; i < count; i++) {
        int rc = check_shape(ptms + off[i], n[i], m[i]);
        if (rc) return rc;
        off[i + 1] = off[i] + (size_t)n[i] * m[i];
    }
    // one engine per shape, created before the clock starts
    std::map<std::pair<int, int>, pbb_engine *> eng;
    auto cleanup = [&]() {
        for (auto &kv : eng) pbb_destroy(kv.second);
    };
    for (int i = 0; i < count; i++) {
        auto key = std::make_pair(n[i], m[i]);
        if (eng.count(key)) continue;
        pbb_engine *e = nullptr;
        int rc = pbb_create(ptms + off[i], n[i], m[i], cfg, INT_MAX, nullptr, &e);
        if (rc) {
            cleanup();
            return rc;
        }
        eng[key] = e;
    }
    const double budget = time_budget_s > 0 ? time_budget_s : 600.0;
    const auto t0 = std::chrono::steady_clock::now();
    auto now_s = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    double t_last = 0.0;
    int rc = PBB_OK;
    for (int i = 0; i < count; i++) {
        memset(&out[i], 0, sizeof out[i]);
        out[i].tag = i;
        out[i].status = PBB_RES_ABORTED;
        out[i].best = ub[i];
        latency_ms[i] = -1.0;
    }
    for (int i = 0; i < count && rc == PBB_OK; i++) {
        while (now_s() < arrival_s[i]) std::this_thread::yield();
        double left = budget - now_s();
        if (left <= 0) break;
        if (request_budget_s > 0 && request_budget_s < left) left = request_budget_s;
        pbb_engine *e = eng[std::make_pair(n[i], m[i])];
        // PBB_UB_NEH (any negative ub): the incumbent is the host NEH's, as the
        // pool computes it on the device at admission; its time counts in the
        // request's latency there too
        int u = ub[i];
        std::vector<int> neh_perm;
        if (u < 0) {
            neh_perm.resize(n[i]);
            for (int j = 0; j < n[i]; j++) neh_perm[j] = j;
            if ((rc = pbb_neh(ptms + off[i], n[i], m[i], neh_perm.data(), &u))) break;
        }
        if ((rc = engine_retarget(e, ptms + off[i], u, u == ub[i] ? nullptr : neh_perm.data()))) break;
        const double t_admit = now_s();
        int all_done = 0;
        if ((rc = pbb_init_full(e))) break;
        if ((rc = pbb_run(e, 0, left, &all_done))) break;
        int cost = 0;
        std::vector<int> perm(e->n);
        if ((rc = pbb_get_incumbent(e, &cost, perm.data()))) break;
        pbb_stats st;
        if ((rc = pbb_get_stats(e, &st))) break;
        const double tc = now_s();
        pbb_result &o = out[i];
        o.status = all_done ? PBB_RES_SOLVED : PBB_RES_ABORTED;
        o.best = cost;
        // improved: perm holds a schedule of cost best (the NEH order counts)
        o.improved = (cost < u || u != ub[i]) ? 1 : 0;
        o.decomposed = st.decomposed;
        o.leaves = st.leaves;
        o.simple_bounds = st.simple_bounds;
        o.johnson_bounds = st.johnson_bounds;
        o.admit_ms = t_admit * 1e3;
        o.done_ms = tc * 1e3;
        for (int j = 0; j < PBB_MAX_JOBS; j++) o.perm[j] = (o.improved && j < e->n) ? perm[j] : -1;
        latency_ms[i] = (tc - arrival_s[i]) * 1e3;
        t_last = tc;
        replay_progress("plain replay", i, i + 1, count);
    }
    cleanup();
    if (wall_s) *wall_s = count ? t_last - arrival_s[0] : 0.0;
    return rc;
}

// ----------------------------------------------------------- checkpoint
extern "C" int pbb_checkpoint_save(pbb_engine *e, const char *path)
{
    if (!e || !path) return fail(PBB_E_ARG, "null argument");
    if (!e->initialized) return fail(PBB_E_STATE, "checkpoint before pbb_init_full/pbb_init_intervals");
    const int n = e->n, E = e->E;
    std::vector<int> pos((size_t)E * n), end((size_t)E * n), ids(E);
    int nb = 0;
    int rc = export_intervals(e, pos.data(), end.data(), ids.data(), &nb, E);
    if (rc) return rc;
    if ((rc = sync_incumbent(e))) return rc;
    pbb_stats st;
    if ((rc = pbb_get_stats(e, &st))) return rc;
    pbb_checkpoint c;
    memset(&c, 0, sizeof c);
    c.n = n;
    c.root_cost = e->initial_ub;
    for (int i = 0; i < n; i++) {
        c.root_perm[i] = e->root_perm[i];
        c.best_perm[i] = e->inc_perm[i];
    }
    c.simple_bounds = st.simple_bounds;
    c.best_cost = e->inc_cost;
    c.work_id = 0;
    c.max_intervals = E;
    c.explored = st.decomposed;
    c.nb = nb;
    int len = 0;
    if ((rc = pbb_checkpoint_format(&c, ids.data(), pos.data(), end.data(), nullptr, 0, &len))) return rc;
    std::vector<char> buf(len + 1);
    if ((rc = pbb_checkpoint_format(&c, ids.data(), pos.data(), end.data(), buf.data(), len + 1, &len))) return rc;
    FILE *f = fopen(path, "w");
    if (!f) return fail(PBB_E_ARG, "cannot write %s", path);
    size_t w = fwrite(buf.data(), 1, len, f);
    fclose(f);
    if (w != (size_t)len) return fail(PBB_E_ARG, "short write to %s", path);
    return PBB_OK;
}

extern "C" int pbb_checkpoint_load(pbb_engine *e, const char *path, pbb_checkpoint *info)
{
    if (!e || !path) return fail(PBB_E_ARG, "null argument");
    std::ifstream f(path);
    if (!f) return fail(PBB_E_ARG, "cannot read %s", path);
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    const int n = e->n, E = e->E;
    std::vector<int> pos((size_t)E * n), end((size_t)E * n), ids(E);
    pbb_checkpoint c;
    int rc = pbb_checkpoint_parse(text.c_str(), n, &c, ids.data(), pos.data(), end.data(), E);
    if (rc) return rc;
    if (!is_perm(c.best_perm, n)) return fail(PBB_E_ARG, "%s: incumbent is not a permutation of 0..%d", path, n - 1);
    if ((rc = pbb_set_root(e, c.root_perm))) return rc;
    if ((rc = pbb_set_incumbent(e, c.best_cost, c.best_perm))) return rc;
    for (int k = 0; k < c.nb; k++) ids[k] = k;  // interval k -> explorer k
    if ((rc = pbb_init_intervals(e, c.nb, ids.data(), pos.data(), end.data()))) return rc;
    // the file's counters carry on (works::init_infile restores simpleBounds,
    // distributed/src/works.cpp:81-83): pbb_get_stats and the next save
    // report the whole run, not the segment since this load
    e->dec_base = c.explored;
    e->sb_base = c.simple_bounds;
    if (info) *info = c;
    return PBB_OK;
}
