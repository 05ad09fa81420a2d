// pbb_pool.cpp -- host side of the multi-instance pool (include/pbb_engine.h,
// "multi-instance pool"): batched solves as one HIP-graph launch, a serving
// loop over host-mapped request/result records, stream replay, autotuner.
// The device work is k_pool (pbb_pool.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <deque>
#include <iterator>
#include <thread>
#include <vector>

#include "../../include/pbb_engine.h"
#include "pbb_internal.h"
#include "pbb_kernels.h"

using namespace pbb;

namespace {

using Clock = std::chrono::steady_clock;

double secs_since(Clock::time_point t0)
{
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

int next_pow2(int x)
{
    int r = 1;
    while (r < x) r <<= 1;
    return r;
}

}  // namespace

struct pbb_pool {
    pbb_config cfg;
    pbb_pool_params prm;
    int M = 0, nmax = 0, pmax = 0, ring_cap = 0, jt_lds = 0, W = 0, epi = 0, margin = 0;
    int ig = 0;  // IG helper rounds per instance (0 = none)
    int groups = 0, E = 0, qcap = 0;
    size_t lds = 0;
    PoolLayout lay{};
    int wall_khz = 100000;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // device
    char *arena = nullptr;
    int8_t *mat = nullptr, *pos = nullptr, *end = nullptr, *dir = nullptr;
    int *line = nullptr, *state = nullptr, *best_perm = nullptr;
    Ctl *run = nullptr;
    unsigned *q_head = nullptr;
    PoolReq *req_d = nullptr;
    PoolRes *res_d = nullptr;
    PoolQueue *q_d = nullptr;
    // pinned batch staging
    PoolReq *req_h = nullptr;
    PoolRes *res_h = nullptr;
    // host-mapped serving records (host pointer / device pointer)
    PoolReq *sreq = nullptr, *sreq_d = nullptr;
    PoolRes *sres = nullptr, *sres_d = nullptr;
    PoolQueue *sq = nullptr, *sq_d = nullptr;
    int *host_flag = nullptr, *host_flag_d = nullptr;
    // batch graph
    hipGraphExec_t gexec = nullptr;
    int g_count = -1, g_epi = -1, g_margin = -1;
    unsigned long long g_deadline = 0, g_req = 0;
    int last_count = 0;  // instances of the last pbb_pool_solve batch
    // serving state
    bool serving = false;
    std::vector<int> free_slots;
    std::vector<int> slot_ticket, slot_tag;
    std::vector<uint64_t> slot_fp;
    std::vector<char> slot_busy;
    uint64_t fp_used = 0, fp_budget = 0;
    int qtail = 0;
    int ticket = 0;
};

static size_t queue_bytes(int qcap) { return sizeof(PoolQueue) + (size_t)qcap * sizeof(int); }

static KArgs pool_kargs(const pbb_pool *p, double budget_s)
{
    KArgs a{};
    a.I.n = p->nmax;
    a.I.m = p->M;
    a.I.mp = (p->M + 3) & ~3;
    a.E = p->E;
    a.rows = p->nmax;
    a.mat = p->mat;
    a.pos = p->pos;
    a.end = p->end;
    a.dir = p->dir;
    a.line = p->line;
    a.state = p->state;
    a.best_perm = p->best_perm;
    a.ctl = p->run;
    a.host_flag = p->host_flag_d;
    a.node_budget = 0;
    a.deadline = budget_s > 0 ? (unsigned long long)(budget_s * p->wall_khz * 1000.0) : 0ull;
    a.bound_mode = p->cfg.bound_mode;
    a.branching = p->cfg.branching;
    a.refine = p->cfg.refine;
    a.pairs_mode = p->cfg.johnson_pairs;
    a.find_all = p->cfg.find_all;
    a.single_node = p->cfg.single_node;
    a.steal_margin = p->margin;
    return a;
}

static PoolArgs pool_pargs(const pbb_pool *p, bool serve)
{
    PoolArgs q{};
    q.arena = p->arena;
    q.lay = p->lay;
    q.nslots = p->prm.slots;
    q.W = p->W;
    q.epi = p->epi;
    q.nmax = p->nmax;
    q.pmax = p->pmax;
    q.ring_cap = p->ring_cap;
    q.ig_iters = p->ig;
    q.jt_lds = p->jt_lds;
    q.pairs_mode = p->cfg.johnson_pairs;
    q.qcap = p->qcap;
    q.sys = serve ? 1 : 0;
    q.req = serve ? p->sreq_d : p->req_d;
    q.res = serve ? p->sres_d : p->res_d;
    q.q = serve ? p->sq_d : p->q_d;
    q.q_head = p->q_head;
    q.run = p->run;
    q.req_ticks = p->prm.request_budget_s > 0 ? (unsigned long long)(p->prm.request_budget_s * p->wall_khz * 1000.0)
                                              : 0ull;
    return q;
}

extern "C" int pbb_pool_params_default(pbb_pool_params *prm)
{
    if (!prm) return fail(PBB_E_ARG, "null params");
    prm->max_jobs = 30;
    prm->machines = 20;
    prm->slots = 512;
    prm->waves_per_group = 4;
    prm->explorers_per_instance = 256;
    prm->steal_margin = 0;
    prm->groups = 0;
    prm->hbm_budget_gb = 0.0;
    prm->request_budget_s = 0.0;
    prm->ig_iterations = 0;
    return PBB_OK;
}

extern "C" int pbb_pool_destroy(pbb_pool *p)
{
    if (!p) return PBB_OK;
    if (p->serving) {
        __atomic_store_n(p->host_flag, 1, __ATOMIC_SEQ_CST);
        hipStreamSynchronize(p->stream);
    }
    if (p->gexec) hipGraphExecDestroy(p->gexec);
    hipFree(p->arena);
    hipFree(p->mat);
    hipFree(p->pos);
    hipFree(p->end);
    hipFree(p->dir);
    hipFree(p->line);
    hipFree(p->state);
    hipFree(p->best_perm);
    hipFree(p->run);
    hipFree(p->q_head);
    hipFree(p->req_d);
    hipFree(p->res_d);
    hipFree(p->q_d);
    if (p->req_h) hipHostFree(p->req_h);
    if (p->res_h) hipHostFree(p->res_h);
    if (p->sreq) hipHostFree(p->sreq);
    if (p->sres) hipHostFree(p->sres);
    if (p->sq) hipHostFree(p->sq);
    if (p->host_flag) hipHostFree(p->host_flag);
    if (p->ev0) hipEventDestroy(p->ev0);
    if (p->ev1) hipEventDestroy(p->ev1);
    if (p->stream) hipStreamDestroy(p->stream);
    delete p;
    return PBB_OK;
}

extern "C" int pbb_pool_create(const pbb_config *cfg, const pbb_pool_params *prm, pbb_pool **out)
{
    if (!cfg || !prm || !out) return fail(PBB_E_ARG, "null argument");
    *out = nullptr;
    int rc = validate(cfg);
    if (rc) return rc;
    if (cfg->branching == -2 || cfg->branching == -3)
        return fail(PBB_E_ARG, "pool: branching %d not supported (use -1, 1, 2, 3)", cfg->branching);
    if (prm->max_jobs < 3 || prm->max_jobs > PBB_MAX_JOBS)
        return fail(PBB_E_ARG, "max_jobs %d outside [3, %d]", prm->max_jobs, PBB_MAX_JOBS);
    if (prm->machines < 2 || prm->machines > PBB_MAX_MACHINES)
        return fail(PBB_E_ARG, "machines %d outside [2, %d]", prm->machines, PBB_MAX_MACHINES);
    if (prm->slots < 1 || prm->slots > (1 << 20)) return fail(PBB_E_ARG, "slots %d", prm->slots);
    if (prm->waves_per_group != 1 && prm->waves_per_group != 2 && prm->waves_per_group != 4)
        return fail(PBB_E_ARG, "waves_per_group must be 1, 2 or 4");
    if (prm->explorers_per_instance < prm->waves_per_group)
        return fail(PBB_E_ARG, "explorers_per_instance below waves_per_group");
    if (prm->steal_margin < 0 || prm->groups < 0 || prm->hbm_budget_gb < 0 || prm->request_budget_s < 0 ||
        prm->ig_iterations < 0)
        return fail(PBB_E_ARG, "negative pool parameter");
    if ((rc = check_device())) return rc;

    pbb_pool *p = new pbb_pool();
    p->cfg = *cfg;
    p->prm = *prm;
    p->M = prm->machines;
    p->nmax = prm->max_jobs;
    p->W = prm->waves_per_group;
    p->epi = prm->explorers_per_instance - prm->explorers_per_instance % p->W;
    p->margin = prm->steal_margin ? prm->steal_margin : cfg->steal_margin;
    const bool lb2 = cfg->bound_mode != 0 && cfg->refine == 1;
    p->pmax = lb2 ? (cfg->johnson_pairs == 0 ? p->M * (p->M - 1) / 2 : p->M - 1) : 0;
    p->jt_lds = (lb2 && jtab_in_lds(p->nmax, p->pmax)) ? 1 : 0;
    p->ig = p->W >= 2 ? prm->ig_iterations : 0;
    p->lds = pool_lds_bytes(p->nmax, p->M, p->pmax, p->jt_lds != 0, p->W) + (p->ig ? pool_ig_bytes(p->M) : 0);
    p->qcap = next_pow2(std::max(64, prm->slots));
#define TRY(x)              \
    do {                    \
        int r_ = (x);       \
        if (r_) {           \
            pbb_pool_destroy(p); \
            return r_;      \
        }                   \
    } while (0)
#define HTRY(x) TRY(([&]() -> int { HIPCHK(x); return PBB_OK; })())
    HTRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    HTRY(hipEventCreate(&p->ev0));
    HTRY(hipEventCreate(&p->ev1));
    int dev = 0;
    HTRY(hipGetDevice(&dev));
    HTRY(hipDeviceGetAttribute(&p->wall_khz, hipDeviceAttributeWallClockRate, dev));
    hipDeviceProp_t prop;
    HTRY(hipGetDeviceProperties(&prop, dev));
    // explorers: resident workgroups x W
    {
        PoolArgs q{};
        q.nmax = p->nmax;
        q.pmax = p->pmax;
        q.jt_lds = p->jt_lds;
        q.W = p->W;
        KArgs a{};
        a.bound_mode = cfg->bound_mode;
        a.refine = cfg->refine;
        int gpc = 0;
        HTRY(pool_occupancy(p->M, a, q, &gpc));
        if (gpc < 1) TRY(fail(PBB_E_CAPACITY, "pool workgroup does not fit a CU (LDS %zu bytes)", p->lds));
        const int cap = gpc * prop.multiProcessorCount;
        p->groups = prm->groups ? std::min(prm->groups, cap) : cap;
    }
    p->E = p->groups * p->W;
    p->epi = std::min(p->epi, p->E);
    p->ring_cap = next_pow2(std::max(64, 2 * p->epi));
    p->lay = pool_layout(p->nmax, p->M, p->pmax, p->ring_cap);
    const size_t S = (size_t)prm->slots, E = (size_t)p->E;
    HTRY(hipMalloc(&p->arena, S * p->lay.stride));
    HTRY(hipMalloc(&p->mat, E * p->nmax * 64));
    HTRY(hipMalloc(&p->pos, E * 64));
    HTRY(hipMalloc(&p->end, E * 64));
    HTRY(hipMalloc(&p->dir, E * 64));
    HTRY(hipMalloc(&p->line, E * 4));
    HTRY(hipMalloc(&p->state, E * 4));
    HTRY(hipMalloc(&p->best_perm, E * 64 * 4));
    HTRY(hipMalloc(&p->run, sizeof(Ctl)));
    HTRY(hipMalloc(&p->q_head, 64));
    HTRY(hipMalloc(&p->req_d, S * sizeof(PoolReq)));
    HTRY(hipMalloc(&p->res_d, S * sizeof(PoolRes)));
    HTRY(hipMalloc(&p->q_d, queue_bytes(p->qcap)));
    HTRY(hipMemset(p->arena, 0, S * p->lay.stride));
    HTRY(hipMemset(p->state, 0, E * 4));
    HTRY(hipMemset(p->run, 0, sizeof(Ctl)));
    HTRY(hipHostMalloc((void **)&p->req_h, S * sizeof(PoolReq), hipHostMallocDefault));
    HTRY(hipHostMalloc((void **)&p->res_h, S * sizeof(PoolRes), hipHostMallocDefault));
    const unsigned mapped = hipHostMallocMapped | hipHostMallocCoherent;
    HTRY(hipHostMalloc((void **)&p->sreq, S * sizeof(PoolReq), mapped));
    HTRY(hipHostMalloc((void **)&p->sres, S * sizeof(PoolRes), mapped));
    HTRY(hipHostMalloc((void **)&p->sq, queue_bytes(p->qcap), mapped));
    HTRY(hipHostMalloc((void **)&p->host_flag, 64, mapped));
    HTRY(hipHostGetDevicePointer((void **)&p->sreq_d, p->sreq, 0));
    HTRY(hipHostGetDevicePointer((void **)&p->sres_d, p->sres, 0));
    HTRY(hipHostGetDevicePointer((void **)&p->sq_d, p->sq, 0));
    HTRY(hipHostGetDevicePointer((void **)&p->host_flag_d, p->host_flag, 0));
    memset(p->sreq, 0, S * sizeof(PoolReq));
    memset(p->sres, 0, S * sizeof(PoolRes));
    memset(p->sq, 0, queue_bytes(p->qcap));
    *p->host_flag = 0;
    size_t freeb = 0, totalb = 0;
    HTRY(hipMemGetInfo(&freeb, &totalb));
    p->fp_budget = prm->hbm_budget_gb > 0 ? (uint64_t)(prm->hbm_budget_gb * 1e9) : (uint64_t)freeb;
#undef HTRY
#undef TRY
    *out = p;
    return PBB_OK;
}

extern "C" int pbb_pool_info(pbb_pool *p, int *explorers, int *groups, int *lds_bytes)
{
    if (!p) return fail(PBB_E_ARG, "null pool");
    if (explorers) *explorers = p->E;
    if (groups) *groups = p->groups;
    if (lds_bytes) *lds_bytes = (int)p->lds;
    return PBB_OK;
}

extern "C" int pbb_pool_retune(pbb_pool *p, int explorers_per_instance, int steal_margin)
{
    if (!p) return fail(PBB_E_ARG, "null pool");
    if (p->serving) return fail(PBB_E_STATE, "pool is serving");
    if (explorers_per_instance < p->W || steal_margin < 0)
        return fail(PBB_E_ARG, "explorers_per_instance %d / steal_margin %d", explorers_per_instance, steal_margin);
    int epi = std::min(explorers_per_instance - explorers_per_instance % p->W, p->E);
    if (2 * epi > p->ring_cap)
        return fail(PBB_E_CAPACITY, "explorers_per_instance %d above the pool's idle-ring capacity %d", epi,
                    p->ring_cap / 2);
    p->epi = epi;
    p->margin = steal_margin ? steal_margin : p->cfg.steal_margin;
    return PBB_OK;
}

// validate one instance of a request list and copy it into a request record
static int fill_req(const pbb_pool *p, PoolReq *r, const int32_t *ptm, int n, int m, int ub, int tag)
{
    int rc = check_shape(ptm, n, m);
    if (rc) return rc;
    if (n > p->nmax) return fail(PBB_E_ARG, "n=%d above the pool's max_jobs %d", n, p->nmax);
    if (m > p->M) return fail(PBB_E_ARG, "m=%d above the pool's machines %d", m, p->M);
    r->n = n;
    r->m = m;
    r->ub = ub;
    r->tag = tag;
    memcpy(r->ptm, ptm, sizeof(int32_t) * (size_t)n * m);
    return PBB_OK;
}

static void to_result(const pbb_pool *p, const PoolRes &r, int n, int tag, unsigned long long t0, pbb_result *o)
{
    memset(o, 0, sizeof *o);
    o->tag = tag;
    o->status = r.status;
    o->best = r.best;
    o->improved = r.improved;
    o->decomposed = r.dec;
    o->leaves = r.leaves;
    o->simple_bounds = r.sb;
    o->johnson_bounds = r.jb;
    const double k = 1.0 / p->wall_khz;
    o->admit_ms = r.t_admit >= t0 ? (double)(r.t_admit - t0) * k : 0.0;
    o->done_ms = r.t_done >= t0 ? (double)(r.t_done - t0) * k : 0.0;
    for (int i = 0; i < PBB_MAX_JOBS; i++) o->perm[i] = (r.improved && i < n) ? r.perm[i] : -1;
}

static int check_run_error(pbb_pool *p)
{
    Ctl c;
    HIPCHK(hipMemcpy(&c, p->run, sizeof c, hipMemcpyDeviceToHost));
    if (c.err) return fail(PBB_E_HIP, "device error %d in the pool kernel", c.err);
    return PBB_OK;
}

extern "C" int pbb_pool_solve(pbb_pool *p, int count, const int32_t *ptms, const int *n, const int *m,
                              const int *ub, const int *tags, double time_budget_s, pbb_result *out,
                              double *kernel_ms)
{
    if (!p || count < 0 || (count > 0 && (!ptms || !n || !m || !ub || !out)))
        return fail(PBB_E_ARG, "bad pool_solve arguments");
    if (p->serving) return fail(PBB_E_STATE, "pool is serving");
    if (count > p->prm.slots) return fail(PBB_E_CAPACITY, "%d instances > %d slots", count, p->prm.slots);
    if (count == 0) return PBB_OK;
    // The device record's tag is the request's index, which the kernel stores
    // as the last word of a retired result (k_pool_init resets it to -1), so
    // any caller tag -- -1 included -- leaves the retirement test exact.
    size_t off = 0;
    for (int i = 0; i < count; i++) {
        int rc = fill_req(p, &p->req_h[i], ptms + off, n[i], m[i], ub[i], i);
        if (rc) return rc;
        off += (size_t)n[i] * m[i];
    }
    const double budget = time_budget_s > 0 ? time_budget_s : 600.0;
    KArgs a = pool_kargs(p, budget);
    PoolArgs q = pool_pargs(p, false);
    if (!p->gexec || p->g_count != count || p->g_deadline != a.deadline || p->g_epi != p->epi ||
        p->g_margin != p->margin || p->g_req != q.req_ticks) {
        if (p->gexec) {
            hipGraphExecDestroy(p->gexec);
            p->gexec = nullptr;
        }
        int gpc = 0;
        HIPCHK(pool_occupancy(p->M, a, q, &gpc));  // sets the kernel's LDS attribute outside the capture
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeRelaxed));
        hipError_t e1 = hipMemcpyAsync(p->req_d, p->req_h, (size_t)count * sizeof(PoolReq), hipMemcpyHostToDevice,
                                       p->stream);
        hipError_t e2 = e1 == hipSuccess ? launch_pool(p->M, a, q, p->groups, count, p->stream) : e1;
        hipError_t e3 = e2 == hipSuccess ? hipMemcpyAsync(p->res_h, p->res_d, (size_t)count * sizeof(PoolRes),
                                                          hipMemcpyDeviceToHost, p->stream)
                                         : e2;
        hipError_t e4 = hipStreamEndCapture(p->stream, &g);
        if (e3 != hipSuccess) {
            if (g) hipGraphDestroy(g);
            HIPCHK(e3);
        }
        HIPCHK(e4);
        hipError_t e5 = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        HIPCHK(e5);
        p->g_count = count;
        p->g_deadline = a.deadline;
        p->g_req = q.req_ticks;
        p->g_epi = p->epi;
        p->g_margin = p->margin;
    }
    *p->host_flag = 0;
    HIPCHK(hipEventRecord(p->ev0, p->stream));
    HIPCHK(hipGraphLaunch(p->gexec, p->stream));
    HIPCHK(hipEventRecord(p->ev1, p->stream));
    HIPCHK(hipStreamSynchronize(p->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, p->ev0, p->ev1));
    if (kernel_ms) *kernel_ms = ms;
    p->last_count = count;
    int rc = check_run_error(p);
    if (rc) return rc;
    unsigned long long t0 = ULLONG_MAX;
    for (int i = 0; i < count; i++)
        if (p->res_h[i].tag == i) t0 = std::min(t0, p->res_h[i].t_admit);
    for (int i = 0; i < count; i++) {
        const int tag = tags ? tags[i] : i;
        if (p->res_h[i].tag != i) {
            // never retired (the launch ended first)
            PoolRes r{};
            r.status = RES_ABORTED;
            r.best = ub[i];
            to_result(p, r, n[i], tag, t0, &out[i]);
        } else {
            to_result(p, p->res_h[i], n[i], tag, t0, &out[i]);
        }
    }
    return PBB_OK;
}

extern "C" int pbb_pool_batch_histograms(pbb_pool *p, int count, uint64_t *decomposed, uint64_t *refined)
{
    if (!p || count < 0 || (count > 0 && (!decomposed || !refined))) return fail(PBB_E_ARG, "bad arguments");
    if (p->serving) return fail(PBB_E_STATE, "pool is serving");
    if (count > p->last_count) return fail(PBB_E_ARG, "%d instances > the last batch's %d", count, p->last_count);
    if (count == 0) return PBB_OK;
    // a batch admits request i into slot i; each slot's control block keeps
    // its histograms until the slot is admitted again
    // one linear copy per slot (the strided 2D copy read zeros under a
    // profiler's counter collection)
    std::vector<Ctl> c(count);
    for (int i = 0; i < count; i++)
        HIPCHK(hipMemcpy(&c[i], p->arena + (size_t)i * p->lay.stride + p->lay.hdr, sizeof(Ctl),
                         hipMemcpyDeviceToHost));
    for (int i = 0; i < count; i++)
        for (int d = 0; d < 64; d++) {
            decomposed[(size_t)i * 64 + d] = c[i].histo[d];
            refined[(size_t)i * 64 + d] = c[i].jhisto[d];
        }
    return PBB_OK;
}

// ---------------------------------------------------------------- serving
extern "C" int pbb_serve_footprint(pbb_pool *p, int n, uint64_t *bytes)
{
    if (!p || !bytes) return fail(PBB_E_ARG, "null argument");
    *bytes = (uint64_t)p->epi * ((uint64_t)n * n + 2ull * n) * 4ull + p->lay.stride;
    return PBB_OK;
}

extern "C" int pbb_serve_start(pbb_pool *p, double time_budget_s)
{
    if (!p) return fail(PBB_E_ARG, "null pool");
    if (p->serving) return fail(PBB_E_STATE, "already serving");
    if (!(time_budget_s > 0)) return fail(PBB_E_ARG, "serving needs a positive time budget");
    const int S = p->prm.slots;
    memset(p->sq, 0, queue_bytes(p->qcap));
    for (int s = 0; s < S; s++) p->sres[s].tag = -1;
    p->free_slots.clear();
    for (int s = S - 1; s >= 0; s--) p->free_slots.push_back(s);
    p->slot_ticket.assign(S, -1);
    p->slot_tag.assign(S, 0);
    p->slot_fp.assign(S, 0);
    p->slot_busy.assign(S, 0);
    p->fp_used = 0;
    p->qtail = 0;
    p->ticket = 0;
    __atomic_store_n(p->host_flag, 0, __ATOMIC_SEQ_CST);
    std::atomic_thread_fence(std::memory_order_seq_cst);
    KArgs a = pool_kargs(p, time_budget_s);
    PoolArgs q = pool_pargs(p, true);
    HIPCHK(launch_pool(p->M, a, q, p->groups, 0, p->stream));
    p->serving = true;
    return PBB_OK;
}

extern "C" int pbb_serve_submit(pbb_pool *p, const int32_t *ptm, int n, int m, int ub, int tag)
{
    if (!p) return fail(PBB_E_ARG, "null pool");
    if (!p->serving) return fail(PBB_E_STATE, "pool is not serving");
    uint64_t fp = 0;
    pbb_serve_footprint(p, n, &fp);
    if (p->free_slots.empty() || p->fp_used + fp > p->fp_budget) return PBB_E_BUSY;
    const int s = p->free_slots.back();
    const int t = p->ticket;
    int rc = fill_req(p, &p->sreq[s], ptm, n, m, ub, t);
    if (rc) return rc;
    p->free_slots.pop_back();
    p->ticket = (p->ticket + 1) & 0x3fffffff;
    p->slot_ticket[s] = t;
    p->slot_tag[s] = tag;
    p->slot_fp[s] = fp;
    p->slot_busy[s] = 1;
    p->fp_used += fp;
    // publish: record, then the queue cell, then the tail
    std::atomic_thread_fence(std::memory_order_release);
    __atomic_store_n(&p->sq->ids[p->qtail & (p->qcap - 1)], s, __ATOMIC_RELEASE);
    p->qtail++;
    __atomic_store_n(&p->sq->tail, p->qtail, __ATOMIC_RELEASE);
    return PBB_OK;
}

extern "C" int pbb_serve_poll(pbb_pool *p, pbb_result *out, int max, int *count)
{
    if (!p || !count || (max > 0 && !out)) return fail(PBB_E_ARG, "bad poll arguments");
    int k = 0;
    const int S = p->prm.slots;
    for (int s = 0; s < S && k < max; s++) {
        if (!p->slot_busy[s]) continue;
        if (__atomic_load_n(&p->sres[s].tag, __ATOMIC_ACQUIRE) != p->slot_ticket[s]) continue;
        PoolRes r;
        memcpy(&r, &p->sres[s], sizeof r);
        to_result(p, r, p->sreq[s].n, p->slot_tag[s], 0ull, &out[k++]);
        p->slot_busy[s] = 0;
        p->fp_used -= p->slot_fp[s];
        p->free_slots.push_back(s);
    }
    *count = k;
    return PBB_OK;
}

extern "C" int pbb_serve_stop(pbb_pool *p)
{
    if (!p) return fail(PBB_E_ARG, "null pool");
    if (!p->serving) return PBB_OK;
    __atomic_store_n(&p->sq->closed, 1, __ATOMIC_RELEASE);
    HIPCHK(hipStreamSynchronize(p->stream));
    p->serving = false;
    return check_run_error(p);
}

// Stop serving on several pools: close every queue first, then wait for
// each kernel.  A pool whose kernel still has groups waiting for a CU can
// only finish once the other pools' kernels have exited, so waiting for one
// pool before the others are closed would hold until their time budgets.
static int stop_pools(pbb_pool *const *pools, const int *which, int count)
{
    for (int j = 0; j < count; j++)
        if (pools[which[j]]->serving) __atomic_store_n(&pools[which[j]]->sq->closed, 1, __ATOMIC_RELEASE);
    int rc = PBB_OK;
    for (int j = 0; j < count; j++) {
        const int r = pbb_serve_stop(pools[which[j]]);
        if (r && !rc) rc = r;
    }
    return rc;
}

// The serving loop of one or more pools sharing the device: request i goes
// to pools[route[i]]; every pool keeps its own pending queue.
static int replay_pools(pbb_pool *const *pools, int np, const std::vector<int> &route, int count,
                        const int32_t *ptms, const int *n, const int *m, const int *ub, const double *arrival_s,
                        double time_budget_s, pbb_result *out, double *latency_ms, double *wall_s)
{
    std::vector<size_t> off(count + 1, 0);
    for (int i = 0; i < count; i++) off[i + 1] = off[i] + (size_t)n[i] * m[i];
    const double budget = time_budget_s > 0 ? time_budget_s : 600.0;
    // Start the pools with the most machines (the most registers per wave)
    // first.  Their serving kernels stay resident until stopped, so a pool
    // started later only gets the register space the earlier ones leave:
    // 512 groups of a 10-machine pool started first hold 2 waves of 168
    // VGPRs on every SIMD and leave no room for one 230-VGPR wave of a
    // 20-machine pool, whose requests then wait for the time budget.
    std::vector<int> order(np);
    for (int k = 0; k < np; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pools[a]->M > pools[b]->M; });
    for (int j = 0; j < np; j++) {
        int rc = pbb_serve_start(pools[order[j]], budget + 5.0);
        if (rc) {
            stop_pools(pools, order.data(), j);
            return rc;
        }
    }
    // Every pool needs a resident workgroup: the first group of a serving
    // kernel to start sets its queue's `started` word.  A pool with none
    // after a few seconds has no room beside the pools started before it.
    const auto ts = Clock::now();
    for (int j = 0; j < np; j++) {
        const int *flag = &pools[order[j]]->sq->started;
        while (!__atomic_load_n(flag, __ATOMIC_ACQUIRE) && secs_since(ts) < 5.0) std::this_thread::yield();
        if (!__atomic_load_n(flag, __ATOMIC_ACQUIRE)) {
            stop_pools(pools, order.data(), np);
            return fail(PBB_E_CAPACITY,
                        "pool %d (%d machines): no workgroup became resident beside the pools started before it; "
                        "give those pools fewer groups", order[j], pools[order[j]]->M);
        }
    }
    for (int i = 0; i < count; i++) {
        latency_ms[i] = -1.0;
        memset(&out[i], 0, sizeof out[i]);
        out[i].tag = i;
        out[i].status = PBB_RES_ABORTED;
        out[i].best = ub[i];
    }
    std::vector<std::deque<int>> pending(np);
    int slots = 1;
    for (int k = 0; k < np; k++) slots = std::max(slots, pools[k]->prm.slots);
    std::vector<pbb_result> got(slots);
    int next = 0, done = 0, rc = PBB_OK;
    double t_last = 0.0;
    const auto t0 = Clock::now();
    while (done < count && rc == PBB_OK) {
        const double t = secs_since(t0);
        if (t > budget) break;
        bool progress = false;
        while (next < count && arrival_s[next] <= t) {
            pending[route[next]].push_back(next);
            next++;
        }
        for (int k = 0; k < np && rc == PBB_OK; k++) {
            while (!pending[k].empty()) {
                const int i = pending[k].front();
                const int r = pbb_serve_submit(pools[k], ptms + off[i], n[i], m[i], ub[i], i);
                if (r == PBB_E_BUSY) break;
                if (r) {
                    rc = r;
                    break;
                }
                pending[k].pop_front();
                progress = true;
            }
            int got_k = 0;
            pbb_serve_poll(pools[k], got.data(), (int)got.size(), &got_k);
            if (got_k) {
                const double tc = secs_since(t0);
                for (int j = 0; j < got_k; j++) {
                    const int i = got[j].tag;
                    out[i] = got[j];
                    latency_ms[i] = (tc - arrival_s[i]) * 1e3;
                    t_last = tc;
                }
                replay_progress("pool replay", done, done + got_k, count);
                done += got_k;
                progress = true;
            }
        }
        if (!progress) std::this_thread::yield();
    }
    const int rc2 = stop_pools(pools, order.data(), np);
    for (int k = 0; k < np; k++) {
        // results that completed while stopping keep latency -1
        int got_k = 0;
        pbb_serve_poll(pools[k], got.data(), (int)got.size(), &got_k);
        for (int j = 0; j < got_k; j++) out[got[j].tag] = got[j];
    }
    if (wall_s) *wall_s = done > 0 ? t_last - arrival_s[0] : 0.0;
    return rc ? rc : rc2;
}

extern "C" int pbb_serve_replay(pbb_pool *p, int count, const int32_t *ptms, const int *n, const int *m,
                                const int *ub, const double *arrival_s, double time_budget_s, pbb_result *out,
                                double *latency_ms, double *wall_s)
{
    if (!p || count < 0 || (count > 0 && (!ptms || !n || !m || !ub || !arrival_s || !out || !latency_ms)))
        return fail(PBB_E_ARG, "bad replay arguments");
    size_t off = 0;
    for (int i = 0; i < count; i++) {
        int rc = check_shape(ptms + off, n[i], m[i]);
        if (rc) return rc;
        if (n[i] > p->nmax || m[i] > p->M) return fail(PBB_E_ARG, "instance %d exceeds the pool shape", i);
        off += (size_t)n[i] * m[i];
    }
    pbb_pool *const pools[1] = {p};
    return replay_pools(pools, 1, std::vector<int>(count, 0), count, ptms, n, m, ub, arrival_s, time_budget_s,
                        out, latency_ms, wall_s);
}

extern "C" int pbb_serve_replay_routed(pbb_pool *const *pools, int npools, int count, const int32_t *ptms,
                                       const int *n, const int *m, const int *ub, const double *arrival_s,
                                       double time_budget_s, pbb_result *out, double *latency_ms, double *wall_s)
{
    if (!pools || npools < 1 || npools > 8 || count < 0 ||
        (count > 0 && (!ptms || !n || !m || !ub || !arrival_s || !out || !latency_ms)))
        return fail(PBB_E_ARG, "bad replay arguments");
    for (int k = 0; k < npools; k++) {
        if (!pools[k]) return fail(PBB_E_ARG, "null pool %d", k);
        for (int j = 0; j < k; j++)
            if (pools[j] == pools[k]) return fail(PBB_E_ARG, "pool %d repeated", k);
    }
    std::vector<int> route(count, -1);
    size_t off = 0;
    for (int i = 0; i < count; i++) {
        int rc = check_shape(ptms + off, n[i], m[i]);
        if (rc) return rc;
        off += (size_t)n[i] * m[i];
        // fewest machines first, then the smallest job limit
        for (int k = 0; k < npools; k++) {
            if (n[i] > pools[k]->nmax || m[i] > pools[k]->M) continue;
            const pbb_pool *r = route[i] < 0 ? nullptr : pools[route[i]];
            if (!r || pools[k]->M < r->M || (pools[k]->M == r->M && pools[k]->nmax < r->nmax)) route[i] = k;
        }
        if (route[i] < 0) return fail(PBB_E_ARG, "instance %d fits no pool", i);
    }
    return replay_pools(pools, npools, route, count, ptms, n, m, ub, arrival_s, time_budget_s, out, latency_ms,
                        wall_s);
}

// -------------------------------------------------------------- autotuner
extern "C" int pbb_pool_autotune(const pbb_config *cfg, const pbb_pool_params *base, int count,
                                 const int32_t *ptms, const int *n, const int *m, const int *ub,
                                 pbb_pool_params *best, double *trials, int max_trials, int *ntrials)
{
    if (!cfg || !base || !best || count < 1 || !ptms || !n || !m || !ub)
        return fail(PBB_E_ARG, "bad autotune arguments");
    // most-promising first, so the time cap tightens early
    const int Ws[] = {4, 2, 1};
    const int epis[] = {256, 1024, 3072, 64};
    const int margins[] = {2, 4, 8};
    std::vector<pbb_result> res(count);
    double best_ms = 1e300;
    int nt = 0;
    *best = *base;
    for (int W : Ws) {
        pbb_pool_params prm = *base;
        prm.waves_per_group = W;
        prm.explorers_per_instance = *std::max_element(std::begin(epis), std::end(epis));
        prm.slots = std::max(base->slots, count);
        pbb_pool *p = nullptr;
        int rc = pbb_pool_create(cfg, &prm, &p);
        if (rc) return rc;
        for (int epi : epis) {
            if (epi < W || epi > p->E) continue;
            for (int mg : margins) {
                rc = pbb_pool_retune(p, epi, mg);
                double ms = 0.0;
                // a trial may not run longer than twice the best so far (+0.5 s)
                const double cap = best_ms < 1e299 ? 2.0 * best_ms / 1e3 + 0.5 : 120.0;
                if (!rc) rc = pbb_pool_solve(p, count, ptms, n, m, ub, nullptr, cap, res.data(), &ms);  // warm
                ms = 1e300;
                for (int r = 0; r < 2 && !rc; r++) {  // best of two (the search order varies)
                    double t = 0.0;
                    rc = pbb_pool_solve(p, count, ptms, n, m, ub, nullptr, cap, res.data(), &t);
                    bool all = true;
                    for (int i = 0; i < count; i++) all = all && res[i].status == PBB_RES_SOLVED;
                    if (!rc && all) ms = std::min(ms, t);
                }
                if (rc) {
                    pbb_pool_destroy(p);
                    return rc;
                }
                if (trials && nt < max_trials) {
                    trials[4 * nt + 0] = W;
                    trials[4 * nt + 1] = epi;
                    trials[4 * nt + 2] = mg;
                    trials[4 * nt + 3] = ms;
                }
                nt++;
                if (ms < best_ms) {
                    best_ms = ms;
                    best->waves_per_group = W;
                    best->explorers_per_instance = epi;
                    best->steal_margin = mg;
                }
            }
        }
        pbb_pool_destroy(p);
    }
    if (ntrials) *ntrials = nt;
    return PBB_OK;
}

// ------------------------------------------------------------ device NEH
extern "C" int pbb_neh_device(int count, const int32_t *ptms, const int *n, const int *m, int *perms, int *costs)
{
    if (count < 0 || (count > 0 && (!ptms || !n || !m || !perms || !costs))) return fail(PBB_E_ARG, "bad arguments");
    if (count == 0) return PBB_OK;
    int rc = check_device();
    if (rc) return rc;
    std::vector<long long> off(count);
    long long tot = 0;
    int M = 2, nmax = 1;
    for (int i = 0; i < count; i++) {
        if ((rc = check_shape(ptms + tot, n[i], m[i]))) return rc;
        off[i] = tot;
        tot += (long long)n[i] * m[i];
        M = std::max(M, m[i]);
        nmax = std::max(nmax, n[i]);
    }
    int32_t *dp = nullptr;
    long long *doff = nullptr;
    int *dn = nullptr, *dm = nullptr, *dperm = nullptr, *dcost = nullptr;
    auto cleanup = [&]() {
        hipFree(dp);
        hipFree(doff);
        hipFree(dn);
        hipFree(dm);
        hipFree(dperm);
        hipFree(dcost);
    };
    auto run = [&]() -> int {
        HIPCHK(hipMalloc(&dp, (size_t)tot * 4));
        HIPCHK(hipMalloc(&doff, (size_t)count * 8));
        HIPCHK(hipMalloc(&dn, (size_t)count * 4));
        HIPCHK(hipMalloc(&dm, (size_t)count * 4));
        HIPCHK(hipMalloc(&dperm, (size_t)count * 64 * 4));
        HIPCHK(hipMalloc(&dcost, (size_t)count * 4));
        HIPCHK(hipMemcpy(dp, ptms, (size_t)tot * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(doff, off.data(), (size_t)count * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dn, n, (size_t)count * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dm, m, (size_t)count * 4, hipMemcpyHostToDevice));
        HIPCHK(launch_neh(M, dp, doff, dn, dm, count, nmax, dperm, dcost, nullptr));
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(perms, dperm, (size_t)count * 64 * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(costs, dcost, (size_t)count * 4, hipMemcpyDeviceToHost));
        return PBB_OK;
    };
    rc = run();
    cleanup();
    return rc;
}

// ---------------------------------------------------------- iterated greedy
// IG::acceptance (heuristic/src/IG.cpp:319-338): thr[d] = floor(2^32 exp(-d/T)),
// T = 0.2 * (mean processing time) / 10 (IG.cpp:24-32), thr[0] = 2^32.
static void ig_thresholds(const int32_t *ptm, int n, int m, unsigned long long *thr)
{
    double sum = 0.0;
    for (int i = 0; i < n * m; i++) sum += ptm[i];
    const double T = 0.2 * (sum / ((double)n * m)) / 10.0;
    for (int d = 0; d < IG_THR; d++) {
        const double p = T > 0.0 ? std::exp(-(double)d / T) : (d == 0 ? 1.0 : 0.0);
        const double v = std::floor(p * 4294967296.0);
        thr[d] = v >= 4294967296.0 ? (1ull << 32) : (unsigned long long)v;
    }
}

extern "C" int pbb_ig_device(int count, const int32_t *ptms, const int *n, const int *m, int chains, int iterations,
                             uint64_t seed, int *perms, int *costs, int *chain_costs, int *chain_perms)
{
    if (count < 0 || chains < 1 || iterations < 0 || (count > 0 && (!ptms || !n || !m || !perms || !costs)))
        return fail(PBB_E_ARG, "bad arguments");
    if (count == 0) return PBB_OK;
    if ((long long)count * chains > (1 << 24)) return fail(PBB_E_CAPACITY, "count x chains above 2^24");
    int rc = check_device();
    if (rc) return rc;
    std::vector<long long> off(count);
    std::vector<unsigned long long> thr((size_t)count * IG_THR);
    long long tot = 0;
    int M = 2;
    for (int i = 0; i < count; i++) {
        if ((rc = check_shape(ptms + tot, n[i], m[i]))) return rc;
        if (n[i] < 3 || n[i] > PBB_POOL_ROW - 1) return fail(PBB_E_ARG, "instance %d: n = %d outside [3, 63]", i, n[i]);
        off[i] = tot;
        ig_thresholds(ptms + tot, n[i], m[i], thr.data() + (size_t)i * IG_THR);
        tot += (long long)n[i] * m[i];
        M = std::max(M, m[i]);
    }
    const size_t outs = (size_t)count * chains;
    int32_t *dp = nullptr;
    long long *doff = nullptr;
    int *dn = nullptr, *dm = nullptr, *dperm = nullptr, *dcost = nullptr;
    unsigned long long *dthr = nullptr;
    auto cleanup = [&]() {
        hipFree(dp);
        hipFree(doff);
        hipFree(dn);
        hipFree(dm);
        hipFree(dperm);
        hipFree(dcost);
        hipFree(dthr);
    };
    std::vector<int> cperm(outs * 64), ccost(outs);
    auto run = [&]() -> int {
        HIPCHK(hipMalloc(&dp, (size_t)tot * 4));
        HIPCHK(hipMalloc(&doff, (size_t)count * 8));
        HIPCHK(hipMalloc(&dn, (size_t)count * 4));
        HIPCHK(hipMalloc(&dm, (size_t)count * 4));
        HIPCHK(hipMalloc(&dperm, outs * 64 * 4));
        HIPCHK(hipMalloc(&dcost, outs * 4));
        HIPCHK(hipMalloc(&dthr, thr.size() * 8));
        HIPCHK(hipMemcpy(dp, ptms, (size_t)tot * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(doff, off.data(), (size_t)count * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dn, n, (size_t)count * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dm, m, (size_t)count * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dthr, thr.data(), thr.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(launch_ig(M, dp, doff, dn, dm, count, chains, iterations, (unsigned long long)seed, dthr, dperm, dcost,
                         nullptr));
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemcpy(cperm.data(), dperm, outs * 64 * 4, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(ccost.data(), dcost, outs * 4, hipMemcpyDeviceToHost));
        return PBB_OK;
    };
    rc = run();
    cleanup();
    if (rc) return rc;
    for (int i = 0; i < count; i++) {
        int bc = 0;
        for (int c = 1; c < chains; c++)
            if (ccost[(size_t)i * chains + c] < ccost[(size_t)i * chains + bc]) bc = c;
        costs[i] = ccost[(size_t)i * chains + bc];
        std::copy(cperm.begin() + ((size_t)i * chains + bc) * 64, cperm.begin() + ((size_t)i * chains + bc + 1) * 64,
                  perms + (size_t)i * 64);
    }
    if (chain_costs) std::copy(ccost.begin(), ccost.end(), chain_costs);
    if (chain_perms) std::copy(cperm.begin(), cperm.end(), chain_perms);
    return PBB_OK;
}
