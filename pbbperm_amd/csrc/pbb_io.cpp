// pbb_io.cpp -- the host-only parts of the C ABI that touch no device:
// error reporting, the configuration (defaults, checks, gpuconfig.ini
// reader), instance sources (Taillard generator, seeded synthetic instances,
// VRF reader), host makespan and NEH.  No HIP header is included, so these
// parsers also build on their own (for instance with the CPU sanitizers).
#include <algorithm>
#include <cctype>
#include <climits>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "../../include/pbb_engine.h"
#include "pbb_util.h"

using namespace pbb;

// ------------------------------------------------------------------ errors
static thread_local std::string g_err;

int pbb::fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char *pbb_last_error(void) { return g_err.c_str(); }
extern "C" const char *pbb_version(void) { return "pbbperm_amd 0.1 (gfx950)"; }

// ------------------------------------------------------------------ config
extern "C" int pbb_config_default(pbb_config *c)
{
    if (!c) return fail(PBB_E_ARG, "null config");
    c->bound_mode = 2;          // arguments.cpp:135
    c->branching = 3;           // arguments.cpp:136
    c->refine = 0;              // as shipped: ivm_bound.cpp:170
    c->johnson_pairs = 1;       // arguments.cpp:158
    c->early_stop_johnson = 1;  // arguments.cpp:157
    c->find_all = 0;
    c->single_node = 1;         // standalone solver (gpu/src/main.cpp:22)
    c->sorted_dfs = 0;
    c->nb_ivms = 4096;          // arguments.cpp:148
    c->steal_margin = 6;        // gpu/include/loadbalance.cuh:90 (size_d - 6)
    c->init_mode = -1;
    c->init_solution = 0;
    c->checkpoint_s = 1;
    c->balance_s = 1;
    c->ramp_nodes = 0;
    c->team_nodes = 8192;
    c->team_waves = 0;
    return PBB_OK;
}

int pbb::validate(const pbb_config *c)
{
    if (c->bound_mode < 0 || c->bound_mode > 2)
        return fail(PBB_E_ARG, "boundingMode %d not in {0,1,2}", c->bound_mode);
    if (!(c->branching == -3 || c->branching == -2 || c->branching == -1 || (c->branching >= 1 && c->branching <= 3)))
        return fail(PBB_E_ARG, "adaptiveBranchingMode %d not in {-3,-2,-1,1,2,3}", c->branching);
    if (c->refine != 0 && c->refine != 1) return fail(PBB_E_ARG, "refine must be 0 or 1");
    if (c->johnson_pairs < 0 || c->johnson_pairs > 2)
        return fail(PBB_E_ARG, "JohnsonMode %d unsupported (0, 1, 2; learned pairs are stateful)",
                    c->johnson_pairs);
    if (c->sorted_dfs != 0)
        return fail(PBB_E_ARG, "sortedDFS=%d unsupported: the engine explores siblings in IVM row order (sortedDFS=0)",
                    c->sorted_dfs);
    if (c->nb_ivms < 1) return fail(PBB_E_ARG, "nbIVMs must be >= 1");
    if (c->steal_margin < 0) return fail(PBB_E_ARG, "steal margin %d is negative", c->steal_margin);
    if (c->ramp_nodes < 0) return fail(PBB_E_ARG, "ramp %d is negative", c->ramp_nodes);
    if (c->team_nodes < 0) return fail(PBB_E_ARG, "team node budget %d is negative", c->team_nodes);
    if (c->team_waves < 0) return fail(PBB_E_ARG, "team size %d is negative", c->team_waves);
    // earlyStopJohnson (bound_fsp_strong.cpp:114-123, :150-154): the Johnson
    // bound stops at the first pair whose value exceeds the incumbent and
    // reorders its pairs by past success, state of each bound object.  With
    // boundingMode 2 the refined value only prunes, and a value above the
    // incumbent prunes whether or not the loop stopped early, so the search
    // is unchanged and the option is accepted.  With boundingMode 1 the
    // Johnson values also choose the branching direction
    // (ivm_bound.cpp:668-676), so the early values would steer the search;
    // the engine computes every pair, and refuses the combination rather than
    // return different counts silently.
    if (c->early_stop_johnson && c->bound_mode == 1 && c->refine == 1)
        return fail(PBB_E_ARG,
                    "earlyStopJohnson=true with boundingMode 1 and the Johnson refinement is unsupported "
                    "(its early values choose the branching direction); set earlyStopJohnson=false");
    return PBB_OK;
}

static std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && isspace((unsigned char)s[a])) a++;
    while (b > a && isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

// INI reader with inih's conventions: '[section]', 'key = value', full-line
// '#'/';' comments, inline ';' comments preceded by whitespace.
static bool read_ini(const char *path, std::map<std::string, std::string> &kv)
{
    std::ifstream f(path);
    if (!f) return false;
    std::string line, section;
    while (std::getline(f, line)) {
        std::string t = trim(line);
        if (t.empty() || t[0] == '#' || t[0] == ';') continue;
        for (size_t i = 1; i < t.size(); i++)
            if (t[i] == ';' && isspace((unsigned char)t[i - 1])) {
                t = trim(t.substr(0, i));
                break;
            }
        if (t[0] == '[') {
            size_t r = t.find(']');
            section = trim(t.substr(1, r == std::string::npos ? std::string::npos : r - 1));
            continue;
        }
        size_t eq = t.find_first_of("=:");
        if (eq == std::string::npos) continue;
        std::string key = trim(t.substr(0, eq)), val = trim(t.substr(eq + 1));
        std::string k = section + "." + key;
        for (auto &ch : k) ch = (char)tolower((unsigned char)ch);
        kv[k] = val;
    }
    return true;
}

static long ini_int(const std::map<std::string, std::string> &kv, const char *k, long dflt)
{
    auto it = kv.find(k);
    if (it == kv.end()) return dflt;
    char *endp = nullptr;
    long v = strtol(it->second.c_str(), &endp, 0);
    return (endp && *endp == '\0' && !it->second.empty()) ? v : dflt;
}

static bool ini_bool(const std::map<std::string, std::string> &kv, const char *k, bool dflt)
{
    auto it = kv.find(k);
    if (it == kv.end()) return dflt;
    std::string v = it->second;
    for (auto &ch : v) ch = (char)tolower((unsigned char)ch);
    if (v == "true" || v == "yes" || v == "on" || v == "1") return true;
    if (v == "false" || v == "no" || v == "off" || v == "0") return false;
    return dflt;
}

extern "C" int pbb_config_load_ini(const char *path, pbb_config *c, char *problem, char *instance)
{
    if (!path || !c) return fail(PBB_E_ARG, "null argument");
    std::map<std::string, std::string> kv;
    if (!read_ini(path, kv)) return fail(PBB_E_ARG, "cannot load %s", path);
    pbb_config_default(c);
    auto gs = [&](const char *k) {
        auto it = kv.find(k);
        return it == kv.end() ? std::string("UNKNOWN") : it->second;
    };
    if (problem) snprintf(problem, 64, "%s", gs("problem.problem").c_str());
    if (instance) snprintf(instance, 64, "%s", gs("problem.instance").c_str());
    // keys and defaults of arguments::readIniFile (common/src/arguments.cpp:68-133)
    c->init_mode = (int)ini_int(kv, "initial.ub", -1);
    c->init_solution = (int)ini_int(kv, "initial.sol", 0);
    c->nb_ivms = (int)ini_int(kv, "gpu.nbivms", 4096);
    c->sorted_dfs = (int)ini_int(kv, "bb.sorteddfs", 1);
    c->early_stop_johnson = ini_bool(kv, "bb.earlystopjohnson", true) ? 1 : 0;
    c->bound_mode = (int)ini_int(kv, "bb.boundingmode", 2);
    c->johnson_pairs = (int)ini_int(kv, "bb.johnsonmode", 1);
    c->find_all = ini_bool(kv, "bb.findall", false) ? 1 : 0;
    c->single_node = ini_bool(kv, "bb.singlenode", false) ? 1 : 0;
    c->branching = (int)ini_int(kv, "bb.adaptivebranchingmode", 3);
    c->checkpoint_s = (int)ini_int(kv, "time.checkpoint", 1);
    c->balance_s = (int)ini_int(kv, "time.balance", 1);
    // engine-specific keys (not in the reference): refinement bound, split
    // margin, donation ramp, team phase
    c->refine = (int)ini_int(kv, "bb.refinementbound", 0);
    c->steal_margin = (int)ini_int(kv, "gpu.stealmargin", 6);
    c->ramp_nodes = (int)ini_int(kv, "gpu.rampnodes", 0);
    c->team_nodes = (int)ini_int(kv, "gpu.teamnodes", 8192);
    c->team_waves = (int)ini_int(kv, "gpu.teamwaves", 0);
    return validate(c);
}

// --------------------------------------------------------------- instances
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

extern "C" int pbb_instance_taillard(int id, int *n_out, int *m_out, int32_t *ptm)
{
    if (id < 1 || id > 120 || !n_out || !m_out || !ptm) return fail(PBB_E_ARG, "taillard id %d", id);
    // shapes of instance_flowshop.cpp:104-144
    int n = id > 110 ? 500 : id > 90 ? 200 : id > 60 ? 100 : id > 30 ? 50 : 20;
    static const int lim[] = {110, 100, 90, 80, 70, 60, 50, 40, 30, 20, 10};
    static const int val[] = {20, 20, 10, 20, 10, 5, 20, 10, 5, 20, 10};
    int m = 5;
    for (int i = 0; i < 11; i++)
        if (id > lim[i]) {
            m = val[i];
            break;
        }
    // Park-Miller via Schrage, single-precision scaling (instance_flowshop.cpp:146-157)
    long seed = ta_seeds[id - 1];
    for (int k = 0; k < m; k++)
        for (int j = 0; j < n; j++) {
            const long M = 2147483647, A = 16807, Q = 127773, R = 2836;
            long q = seed / Q;
            seed = A * (seed % Q) - q * R;
            if (seed < 0) seed += M;
            float u = (float)seed / (float)M;
            ptm[k * n + j] = (int32_t)(1 + (long)((double)u * 99.0));
        }
    *n_out = n;
    *m_out = m;
    return PBB_OK;
}

static uint64_t sm_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint64_t sm_uniform(uint64_t *s, uint64_t k)
{
    uint64_t limit = UINT64_MAX - (UINT64_MAX % k + 1) % k;
    uint64_t x;
    do x = sm_next(s);
    while (x > limit);
    return x % k;
}

extern "C" int pbb_instance_synth(uint64_t seed, int n, int m, int32_t *ptm)
{
    if (n < 2 || m < 2 || !ptm) return fail(PBB_E_ARG, "bad synth shape");
    uint64_t s = seed;
    for (int k = 0; k < m; k++)
        for (int j = 0; j < n; j++) ptm[k * n + j] = 1 + (int32_t)sm_uniform(&s, 99);
    return PBB_OK;
}

extern "C" int pbb_instance_synth_shape(uint64_t seed, int nlo, int nhi, const int *mchoices,
                                        int nchoices, int *n, int *m)
{
    if (nhi < nlo || nchoices < 1 || !mchoices || !n || !m) return fail(PBB_E_ARG, "bad shape range");
    uint64_t s = seed ^ 0x5348415045ull;
    *n = nlo + (int)sm_uniform(&s, (uint64_t)(nhi - nlo + 1));
    *m = mchoices[sm_uniform(&s, (uint64_t)nchoices)];
    return PBB_OK;
}

extern "C" int pbb_instance_read_vrf(const char *path, int *n_out, int *m_out, int32_t *ptm, int cap)
{
    if (!path || !n_out || !m_out || !ptm || cap < 0) return fail(PBB_E_ARG, "bad VRF arguments");
    std::ifstream f(path);
    if (!f) return fail(PBB_E_ARG, "cannot open %s", path);
    int n = 0, m = 0;
    f >> n >> m;
    if (n < 1 || m < 1) return fail(PBB_E_ARG, "bad VRF header in %s", path);
    // the product in 64 bits: a header such as "65536 65536" wraps an int product
    if ((long long)n * m > cap)
        return fail(PBB_E_CAPACITY, "ptm buffer too small (%d < %lld)", cap, (long long)n * m);
    // job-major (machine, time) pairs, transposed on read (instance_vrf.cpp:50-64)
    for (int j = 0; j < n; j++)
        for (int k = 0; k < m; k++) {
            int mach = 0, t = 0;
            if (!(f >> mach >> t)) return fail(PBB_E_ARG, "truncated VRF file %s", path);
            ptm[k * n + j] = t;
        }
    *n_out = n;
    *m_out = m;
    return PBB_OK;
}

// every job id 0..n-1 exactly once (the ids index the time matrix)
bool pbb::is_perm(const int *p, int n)
{
    std::vector<char> seen(n, 0);
    for (int i = 0; i < n; i++)
        if (p[i] < 0 || p[i] >= n || seen[p[i]]++) return false;
    return true;
}

int pbb::host_partial_makespan(const int32_t *ptm, int n, int m, const int *seq, int len)
{
    std::vector<int> cv(m, 0);
    for (int i = 0; i < len; i++) {
        int j = seq[i];
        cv[0] += ptm[j];
        for (int k = 1; k < m; k++) cv[k] = std::max(cv[k - 1], cv[k]) + ptm[k * n + j];
    }
    return cv[m - 1];
}

extern "C" int pbb_makespan_host(const int32_t *ptm, int n, int m, const int *perm, int *cost)
{
    if (!ptm || !perm || !cost || n < 1 || m < 1) return fail(PBB_E_ARG, "bad makespan args");
    if (!is_perm(perm, n)) return fail(PBB_E_ARG, "makespan: not a permutation of 0..%d", n - 1);
    *cost = host_partial_makespan(ptm, n, m, perm, n);
    return PBB_OK;
}

// NEH (fastNEH.cpp:151-243): stable sort by decreasing total time, the first
// two jobs kept in sorted order only if strictly better, then insertion of
// each job at the first position of minimum partial makespan.
extern "C" int pbb_neh(const int32_t *ptm, int n, int m, int *perm, int *cost)
{
    if (!ptm || !perm || !cost || n < 2 || m < 1) return fail(PBB_E_ARG, "bad NEH args");
    if (!is_perm(perm, n)) return fail(PBB_E_ARG, "NEH: start order is not a permutation of 0..%d", n - 1);
    std::vector<int> tot(n, 0);
    for (int j = 0; j < n; j++)
        for (int k = 0; k < m; k++) tot[j] += ptm[k * n + j];
    std::stable_sort(perm, perm + n, [&](int a, int b) { return tot[a] > tot[b]; });
    int two[2] = {perm[1], perm[0]};
    if (!(host_partial_makespan(ptm, n, m, perm, 2) < host_partial_makespan(ptm, n, m, two, 2))) {
        perm[0] = two[0];
        perm[1] = two[1];
    }
    std::vector<int> tmp(n);
    for (int k = 2; k < n; k++) {
        int job = perm[k], bestc = INT_MAX, bestp = 0;
        for (int p = 0; p <= k; p++) {
            int t = 0;
            for (int i = 0; i < p; i++) tmp[t++] = perm[i];
            tmp[t++] = job;
            for (int i = p; i < k; i++) tmp[t++] = perm[i];
            int c = host_partial_makespan(ptm, n, m, tmp.data(), k + 1);
            if (c < bestc) {
                bestc = c;
                bestp = p;
            }
        }
        for (int p = k; p > bestp; p--) perm[p] = perm[p - 1];
        perm[bestp] = job;
    }
    *cost = host_partial_makespan(ptm, n, m, perm, n);
    return PBB_OK;
}


int pbb::check_shape(const int32_t *ptm, int n, int m)
{
    if (!ptm) return fail(PBB_E_ARG, "null ptm");
    if (n < 3 || n > PBB_MAX_JOBS) return fail(PBB_E_ARG, "n=%d outside [3, %d]", n, PBB_MAX_JOBS);
    if (m < 2 || m > PBB_MAX_MACHINES)
        return fail(PBB_E_ARG, "m=%d outside [2, %d]", m, PBB_MAX_MACHINES);
    for (int i = 0; i < n * m; i++)
        if (ptm[i] < 0 || ptm[i] > 1000000) return fail(PBB_E_ARG, "processing time %d out of range", ptm[i]);
    return PBB_OK;
}

void pbb::replay_progress(const char *what, int before, int after, int count)
{
    static const bool on = getenv("PBB_REPLAY_PROGRESS") != nullptr;
    if (on && (after / 500 > before / 500 || (after == count && before < count))) {
        fprintf(stderr, "[%s] %d / %d requests served\n", what, after, count);
        fflush(stderr);
    }
}

