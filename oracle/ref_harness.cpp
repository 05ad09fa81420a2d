// ref_harness.cpp -- drives the reference's own compiled sources (bounds/,
// multicore/src/ivm.cpp, multicore/src/ivm_bound.cpp, common/src/{solution,
// subproblem,arguments}.cpp, heuristic/src/{fastNEH,IG,fspnhoods,
// fastinsertremove}.cpp), built unmodified from
// the jangmys/PBBPerm tree by oracle/Makefile into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: it writes golden vectors for tests/golden and is
// the "reference" CPU baseline of bench.py.  Two pieces of the reference do not
// compile as shipped and are NOT linked:
//   * common/src/pbab.cpp (uses the undeclared arguments::init_solution):
//     the pbab object is created here field by field (size, instance, sltn,
//     root_sltn, stats) -- the only members ivm_bound reads.
//   * multicore/src/sequentialbb.cpp (declares next(bool&), defines next()):
//     its driver loop (setRoot / initAtInterval / unfold / next) is restated
//     below in terms of the reference's ivm and ivm_bound methods.
// The Johnson-refinement variant (ivm_bound.cpp:171 instead of :170) is made by
// composition: a bound object whose boundChildren is the reference LB1 and
// whose bornes_calculer forwards to the reference bound_fsp_strong.
//
// Usage:
//   ref_harness bounds  <ptmfile> <nsamples> <seed> <pairs>   -> per-sample lines
//   ref_harness solve   <ptmfile> <ub> <boundMode> <branching> <refine> <pairs> <findAll>
//                       [max_nodes] [time_limit_s] [pos... end...]
//   ref_harness neh     <ptmfile>
//   ref_harness taillard <id>
//   ref_harness moves   <ptmfile> <nsamples> <seed>
// ptmfile: "n m" followed by the m x n machine-major processing times.
#include <chrono>
#include <fcntl.h>
#include <unistd.h>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "arguments.h"
#include "IG.h"
#include "fastNEH.h"
#include "ivm.h"
#include "ivm_bound.h"
#include "libbounds.h"
#include "pbab.h"
#include "solution.h"
#include "subproblem.h"

struct ptm_instance : public instance_abstract {
    explicit ptm_instance(const std::string &text)
    {
        data = new std::stringstream(text);
        *data >> size;
        data->seekg(0);
    }
};

static std::string read_file(const char *path)
{
    std::ifstream f(path);
    if (!f) {
        fprintf(stderr, "cannot open %s\n", path);
        exit(2);
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// bound object for the Johnson-refinement variant: LB1 children from the
// reference bound_fsp_weak, node bound from the reference bound_fsp_strong.
struct weak_children_strong_node : public bound_fsp_weak {
    bound_fsp_strong *strong = nullptr;
    void bornes_calculer(int permutation[], int l1, int l2, int *couts, int best) override
    {
        strong->bornes_calculer(permutation, l1, l2, couts, best);
    }
};

// splitmix64 + rejection, the same stream as the oracle's generator
static uint64_t sm_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint64_t sm_unif(uint64_t *s, uint64_t k)
{
    uint64_t limit = UINT64_MAX - (UINT64_MAX % k + 1) % k;
    uint64_t x;
    do x = sm_next(s);
    while (x > limit);
    return x % k;
}

static int cmd_bounds(int argc, char **argv)
{
    ptm_instance inst(read_file(argv[2]));
    long ns = atol(argv[3]);
    uint64_t seed = strtoull(argv[4], nullptr, 10);
    int pairs = atoi(argv[5]);
    int n = inst.size;

    bound_fsp_weak weak;
    weak.set_instance(&inst);
    weak.init();
    weak.branchingMode = 1;
    bound_fsp_strong strong;
    strong.set_instance(&inst);
    strong.init();
    strong.branchingMode = 1;
    strong.earlyExit = 0;
    strong.machinePairs = pairs;

    std::vector<int> perm(n), cb(n), ce(n), pb(n), pe(n);
    uint64_t s = seed;
    for (long t = 0; t < ns; t++) {
        for (int i = 0; i < n; i++) perm[i] = i;
        for (int i = n - 1; i > 0; i--) {
            int j = (int)sm_unif(&s, (uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        int l1 = -1 + (int)sm_unif(&s, (uint64_t)n);            // [-1, n-2]
        int l2 = l1 + 1 + (int)sm_unif(&s, (uint64_t)(n - l1)); // [l1+1, n]
        int c[2] = {0, 0};
        printf("%d %d", l1, l2);
        for (int i = 0; i < n; i++) printf(" %d", perm[i]);
        printf(" |");
        weak.boundChildren(perm.data(), l1, l2, cb.data(), ce.data(), pb.data(), pe.data());
        for (int i = 0; i < n; i++) printf(" %d", cb[i]);
        printf(" |");
        for (int i = 0; i < n; i++) printf(" %d", ce[i]);
        printf(" |");
        weak.bornes_calculer(perm.data(), l1, l2, c, -1);
        printf(" %d", c[0]);
        c[0] = c[1] = -7;
        strong.bornes_calculer(perm.data(), l1, l2, c, -1);
        printf(" %d %d", c[0], c[1]);
        printf(" %d\n", weak.evalSolution(perm.data()));
    }
    return 0;
}

// fastNEH::findBestPos reads tail[k][len] (heuristic/src/fastNEH.cpp:159),
// a cell computeTails(perm, len) never writes (:104-124): as shipped the
// value at the last insertion position depends on whatever malloc returned.
// The harness zeroes every tail row after construction (tail is a public
// member, heuristic/include/fastNEH.h:17-19).  Step k writes cells 0..k-1
// only, so cell k still holds that zero when step k reads it, and the
// position-k candidate becomes the makespan of appending the job -- the
// reference's own code with its undefined read made defined.
// Input: one or more instances, each "n m" + the m x n times, read until EOF;
// one output line per instance: the cost then the order.
static int cmd_neh(int argc, char **argv)
{
    std::stringstream all(read_file(argv[2]));
    while (true) {
        int n = 0, m = 0;
        if (!(all >> n >> m) || n < 2 || m < 1) break;
        std::stringstream one;
        one << n << " " << m << "\n";
        for (int i = 0; i < n * m; i++) {
            int v = 0;
            all >> v;
            one << v << (i % n == n - 1 ? "\n" : " ");
        }
        ptm_instance inst(one.str());
        fastNEH neh(&inst);
        for (int k = 0; k < neh.nbMachines; k++)
            for (int j = 0; j < neh.nbJob; j++) neh.tail[k][j] = 0;
        std::vector<int> perm(n);
        for (int i = 0; i < n; i++) perm[i] = i;
        int cost = 0;
        neh.initialSort(perm.data());
        neh.runNEH(perm.data(), cost);
        printf("%d", cost);
        for (int i = 0; i < n; i++) printf(" %d", perm[i]);
        printf("\n");
    }
    return 0;
}

// The moves of the reference's iterated greedy (heuristic/src/IG.cpp,
// fspnhoods.cpp, fastinsertremove.cpp): the IG object's visit order, then
// for nsamples seeded partial sequences (len uniform in [2, n-1], a
// Fisher-Yates order from the splitmix64 stream of `seed`, the job after the
// sequence as the one to insert) fastInsertRemove::insertMakespans and
// removeMakespans.  The tail table is zeroed before each call: insertMakespans
// reads tail[k][len], a cell computeTails(perm, len) does not write, so the
// insertion at the end adds zero as intended (the reference reads whatever
// an earlier, longer sequence left there).
// Output: "vo v0 v1 ...", then per sample
// "len job | seq0 .. seq_len-1 | ins0 .. ins_len | rem0 .. rem_len-1".
static int cmd_moves(int argc, char **argv)
{
    ptm_instance inst(read_file(argv[2]));
    const int ns = atoi(argv[3]);
    uint64_t st = strtoull(argv[4], nullptr, 10);
    IG ig(&inst);
    fastInsertRemove *mv = ig.nhood->m;
    const int n = ig.nbJob;
    printf("vo");
    for (int i = 0; i < n; i++) printf(" %d", ig.visitOrder[i]);
    printf("\n");
    std::vector<int> perm(n), ms(n + 1);
    std::vector<int> msv(n + 1);
    for (int s = 0; s < ns; s++) {
        for (int i = 0; i < n; i++) perm[i] = i;
        for (int i = n - 1; i > 0; i--) {
            int j = (int)sm_unif(&st, (uint64_t)(i + 1));
            std::swap(perm[i], perm[j]);
        }
        int len = 2 + (int)sm_unif(&st, (uint64_t)(n - 2));
        int job = perm[len];
        for (int k = 0; k < mv->nbMachines; k++)
            for (int j = 0; j < n; j++) mv->tail[k][j] = 0;
        mv->insertMakespans(perm.data(), len, job, msv);
        printf("%d %d |", len, job);
        for (int i = 0; i < len; i++) printf(" %d", perm[i]);
        printf(" |");
        for (int i = 0; i <= len; i++) printf(" %d", msv[i]);
        mv->removeMakespans(perm.data(), len, ms.data());
        printf(" |");
        for (int i = 0; i < len; i++) printf(" %d", ms[i]);
        printf("\n");
    }
    return 0;
}

static int cmd_taillard(int argc, char **argv)
{
    std::string name = std::string("ta") + argv[2];
    instance_flowshop inst(name.c_str());
    inst.data->seekg(0);
    std::string all = inst.data->str();
    printf("%s\n", all.c_str());
    return 0;
}

// ---- the driver loop of multicore/src/sequentialbb.cpp, restated ----
struct seqdriver {
    pbab *pbb;
    ivm *IVM;
    ivm_bound *bd;
    int n;
    int mode;

    void bound_current()
    {
        switch (mode) {
        case 0: bd->weakBoundPrune(IVM); break;
        case 1: bd->strongBoundPrune(IVM); break;
        case 2: bd->mixedBoundPrune(IVM); break;
        }
    }
    void set_root(int *order)
    {
        IVM->clearInterval();
        for (int i = 0; i < n; i++) IVM->jobMat[i] = order[i];
        IVM->line = 0;
        bd->prepareSchedule(IVM);
        bd->boundRoot(IVM);
    }
    void unfold()
    {
        while (IVM->line < n - 2) {
            if (IVM->pruningCellState()) {
                for (int i = IVM->line + 1; i < n; i++) IVM->posVect[i] = 0;
                break;
            }
            IVM->line++;
            IVM->generateLine(IVM->line, false);
            bd->prepareSchedule(IVM);
            bound_current();
        }
    }
    bool init_interval(const int *pos, const int *end)
    {
        IVM->line = 0;
        for (int i = 0; i < n; i++) {
            IVM->posVect[i] = pos[i];
            IVM->endVect[i] = end[i];
        }
        pbb->stats.simpleBounds = 0;
        if (IVM->beforeEnd()) {
            unfold();
            return true;
        }
        return false;
    }
    bool next()
    {
        bd->local_best = INT_MAX;
        pbb->sltn->getBest(bd->local_best);
        int state = 0;
        while (IVM->beforeEnd()) {
            if (IVM->lineEndState()) {
                IVM->goUp();
                continue;
            }
            if (IVM->pruningCellState()) {
                IVM->goRight();
                continue;
            }
            state = 1;
            pbb->stats.totDecomposed++;
            IVM->goDown();
            bd->prepareSchedule(IVM);
            if (IVM->isLastLine()) {
                bd->boundLeaf(IVM);
                state = 0;
                continue;
            }
            break;
        }
        if (state == 1) bound_current();
        return state == 1;
    }
};

static int cmd_solve(int argc, char **argv)
{
    ptm_instance inst(read_file(argv[2]));
    int n = inst.size;
    int ub = atoi(argv[3]);
    int boundMode = atoi(argv[4]);
    int branching = atoi(argv[5]);
    int refine = atoi(argv[6]);
    int pairs = atoi(argv[7]);
    int findAll = atoi(argv[8]);
    unsigned long long max_nodes = argc > 9 ? strtoull(argv[9], nullptr, 10) : 0;
    double time_limit = argc > 10 ? atof(argv[10]) : 0.0;
    std::vector<int> pos(n, 0), end(n, 0);
    for (int i = 0; i < n; i++) end[i] = n - i - 1;
    if (argc > 11 + 2 * n - 1) {
        for (int i = 0; i < n; i++) pos[i] = atoi(argv[11 + i]);
        for (int i = 0; i < n; i++) end[i] = atoi(argv[11 + n + i]);
    }

    strcpy(arguments::problem, "flowshop");
    strcpy(arguments::inst_name, "Vharness");
    arguments::boundMode = boundMode;
    arguments::branchingMode = branching;
    arguments::sortNodes = 0;
    arguments::nodePriority = 1;
    arguments::earlyStopJohnson = false;
    arguments::johnsonPairs = pairs;
    arguments::findAll = findAll != 0;
    arguments::printSolutions = false;
    arguments::singleNode = true;

    // pbab without pbab.cpp: only the fields ivm_bound reads are set
    void *raw = ::operator new(sizeof(pbab));
    memset(raw, 0, sizeof(pbab));
    pbab *pbb = static_cast<pbab *>(raw);
    pbb->size = n;
    pbb->instance = &inst;
    pbb->sltn = new solution(n);
    pbb->sltn->cost = ub;
    pbb->root_sltn = new solution(n);
    pbb->root_sltn->cost = ub;
    pbb->foundSolution = false;

    seqdriver d;
    d.pbb = pbb;
    d.n = n;
    d.mode = boundMode;
    d.IVM = new ivm(n);
    d.bd = new ivm_bound(pbb);
    if (refine == 1) {
        auto *comp = new weak_children_strong_node();
        comp->set_instance(&inst);
        comp->init();
        comp->branchingMode = branching;
        comp->strong = static_cast<bound_fsp_strong *>(d.bd->bound[1]);
        d.bd->bound[0] = comp;
    }

    // leaf improvements print "leaf <cost>" from ivm_bound::boundLeaf; keep them off stdout
    fflush(stdout);
    int saved = dup(1);
    int nul = open("/dev/null", O_WRONLY);
    dup2(nul, 1);
    close(nul);

    auto t0 = std::chrono::steady_clock::now();
    d.set_root(pbb->root_sltn->perm);
    d.init_interval(pos.data(), end.data());
    bool finished = true;
    unsigned long long iter = 0;
    while (d.next()) {
        if (max_nodes && pbb->stats.totDecomposed >= max_nodes) {
            finished = false;
            break;
        }
        if (time_limit > 0 && (++iter & 1023) == 0) {
            double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (el >= time_limit) {
                finished = false;
                break;
            }
        }
    }
    double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fflush(stdout);
    dup2(saved, 1);
    close(saved);

    printf("{\"finished\": %s, \"best\": %d, \"decomposed\": %llu, \"leaves\": %llu, "
           "\"simple_bounds\": %llu, \"johnson_bounds\": %llu, \"wall_s\": %.6f, \"perm\": [",
           finished ? "true" : "false", pbb->sltn->cost,
           (unsigned long long)pbb->stats.totDecomposed.load(), (unsigned long long)pbb->stats.leaves.load(),
           (unsigned long long)pbb->stats.simpleBounds.load(),
           (unsigned long long)pbb->stats.johnsonBounds.load(), el);
    for (int i = 0; i < n; i++) printf(i ? ", %d" : "%d", pbb->sltn->perm[i]);
    printf("]}\n");
    fflush(stdout);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 3) {
        fprintf(stderr, "usage: see header\n");
        return 2;
    }
    std::string cmd = argv[1];
    if (cmd == "bounds" && argc >= 6) return cmd_bounds(argc, argv);
    if (cmd == "solve" && argc >= 9) return cmd_solve(argc, argv);
    if (cmd == "neh") return cmd_neh(argc, argv);
    if (cmd == "taillard") return cmd_taillard(argc, argv);
    if (cmd == "moves" && argc >= 5) return cmd_moves(argc, argv);
    fprintf(stderr, "bad command\n");
    return 2;
}
