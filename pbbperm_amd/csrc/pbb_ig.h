// pbb_ig.h -- iterated greedy chains on one wavefront (heuristic/src/IG.cpp
// runIG, with the k-insertion moves of heuristic/src/fspnhoods.cpp:46-100 and
// the insertion/removal evaluations of heuristic/src/fastinsertremove.cpp),
// shared by the IG kernel (pbb_ig.hip) and the pool's helper wave
// (pbb_pool.hip).
//
// A chain's sequences are lane vectors (lane = position).  Every move is
// evaluated with the heads and tails of the current sequence
// (fastInsertRemove's computeHeads / computeTails), computed by a systolic
// pass over the lanes: at step t lane j advances position j to machine t - j
// (heads) taking its left neighbour's value of the step before through DPP,
// and likewise from the right for the tails, so one pass costs n + M - 1
// steps.  The makespans of inserting a job before every position, or of
// removing every position, then take M steps with one lane per position.
//
// The reference seeds its generators from std::random_device and accepts
// worse solutions through a float exp; a chain here draws from its own
// SplitMix64 stream and compares against integer thresholds
// floor(2^32 exp(-delta / T)).
#pragma once

#include "pbb_explorer.h"

namespace pbb {

__device__ __forceinline__ uint64_t ig_next(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// uniform in [0, k), k >= 1: high 32 bits of the stream, multiply-shift
// with rejection of the biased low products (every lane draws the same)
__device__ __forceinline__ uint32_t ig_below(uint64_t &s, uint32_t k)
{
    for (;;) {
        const uint32_t x = (uint32_t)(ig_next(s) >> 32);
        const uint64_t mm = (uint64_t)x * k;
        const uint32_t lo = (uint32_t)mm;
        if (lo < k) {
            const uint32_t t = (0u - k) % k;
            if (lo < t) continue;
        }
        return (uint32_t)(mm >> 32);
    }
}

// IG::IG visitOrder (IG.cpp:38-52) at index i: the middle position, then
// alternately below and above it
__device__ __forceinline__ int ig_visit(int n, int i)
{
    const int start = n / 2;
    return (i & 1) ? start - 1 - (i >> 1) : start + (i >> 1);
}

// inclusive prefix sum over all 64 lanes
__device__ __forceinline__ int scan_add64(int x)
{
    x = scan_add32(x);
    const int lo = rdlane(x, 31);
    return lane_id() >= 32 ? x + lo : x;
}

// lane i takes lane i+1 (lane 63: 0) / lane i-1 (lane 0: 0)
__device__ __forceinline__ int from_right(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, false); }
__device__ __forceinline__ int from_left(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, false); }

__device__ __forceinline__ int seq_remove(int v, int pos)
{
    const int r = from_right(v);
    return lane_id() >= pos ? r : v;
}

__device__ __forceinline__ int seq_insert(int v, int pos, int job)
{
    const int l = from_left(v);
    const int lane = lane_id();
    return lane > pos ? l : (lane == pos ? job : v);
}

template <int M>
struct IgWave {
    const int *pm;   // [M][PM_STRIDE] machine-major times (zero rows beyond the instance's m)
    const int *spt;  // [64] total time of each job
    int *hd, *tl;    // [M][64] heads / tails of the current sequence
    int *scr;        // 64 ints of scratch
    int n, lane;

    // computeHeads + computeTails (fastinsertremove.cpp:75-116) of the
    // sequence in v, positions 0..len-1; returns its makespan
    __device__ __forceinline__ int heads_tails(int v, int len) const
    {
        const bool in = lane < len;
        const int job = in ? v : 0;
        int h = 0, q = 0;
        const int steps = len + M - 1;
        for (int t = 0; t < steps; t++) {
            const int k = t - lane;                       // heads: machine of this position
            const int kt = M - 1 - (t - (len - 1 - lane));  // tails: machine of this position
            const int left = from_left(h);
            const int right = from_right(q);
            const bool dh = in && k >= 0 && k < M;
            const bool dt = in && kt >= 0 && kt < M;
            const int ph = dh ? pm[k * PM_STRIDE + job] : 0;
            const int pt = dt ? pm[kt * PM_STRIDE + job] : 0;
            if (dh) {
                h = max(h, left) + ph;
                hd[k * 64 + lane] = h;
            }
            if (dt) {
                q = max(q, right) + pt;
                tl[kt * 64 + lane] = q;
            }
        }
        __builtin_amdgcn_wave_barrier();
        return uni(hd[(M - 1) * 64 + len - 1]);
    }

    // insertMakespans (fastinsertremove.cpp:196-214): lane i <= len, the
    // makespan with job inserted before position i (zero tail at the end)
    __device__ __forceinline__ int insert_eval(int len, int job) const
    {
        const int i = lane;
        int e = 0, mk = 0;
#pragma unroll
        for (int k = 0; k < M; k++) {
            const int hp = (i > 0 && i <= len) ? hd[k * 64 + i - 1] : 0;
            const int tv = i < len ? tl[k * 64 + i] : 0;
            e = max(e, hp) + pm[k * PM_STRIDE + job];
            mk = max(mk, e + tv);
        }
        return mk;
    }

    // removeMakespans (fastinsertremove.cpp:243-279): lane i < len, the
    // makespan without position i
    __device__ __forceinline__ int remove_eval(int len) const
    {
        const int i = lane;
        int mk = 0;
#pragma unroll
        for (int k = 0; k < M; k++) {
            const int hp = (i > 0 && i < len) ? hd[k * 64 + i - 1] : 0;
            const int tv = (i + 1 < len) ? tl[k * 64 + i + 1] : 0;
            mk = max(mk, hp + tv);
        }
        return mk;
    }

    // first position of least makespan among lanes 0..len, except `tabu`
    // (bestInsert's strict-less scan, fastinsertremove.cpp:326-363)
    __device__ __forceinline__ int best_insert(int &v, int &len, int job, int tabu) const
    {
        heads_tails(v, len);
        const int mk = insert_eval(len, job);
        const bool ok = lane <= len && lane != tabu;
        const int mn = wave_min(ok ? mk : 0x7fffffff);
        const int bp = __ffsll((unsigned long long)ballot(ok && mk == mn)) - 1;
        v = seq_insert(v, bp, job);
        len++;
        return mn;
    }

    // fspnhood::kImove (fspnhoods.cpp:46-100); returns the makespan of v on exit
    __device__ __forceinline__ int ki_move(int &v, int pos, int kmax, int cmax0, uint64_t &rng) const
    {
        int len = n;
        int rjob = rdlane(v, pos);
        v = seq_remove(v, pos);
        len--;
        uint64_t tabu_jobs = 1ull << rjob;
        int tabu_pos = pos;
        int cmax1 = 0;
        for (int k = 0;;) {
            cmax1 = best_insert(v, len, rjob, tabu_pos);
            if (cmax1 <= cmax0) break;
            if (++k == kmax) break;
            // bestRemove2 (fastinsertremove.cpp:463-496): a non-tabu job
            // drawn with weight (gain / total time)^2, as q^2 with
            // q = floor(256 gain / total time)
            const int old = heads_tails(v, len);
            const int rm = remove_eval(len);
            const int job = lane < len ? v : 0;
            const bool free_ = lane < len && !((tabu_jobs >> job) & 1ull);
            uint32_t w = 0;
            if (free_) {
                const uint32_t s = (uint32_t)spt[job];
                if (s > 0) {
                    const uint64_t q = ((uint64_t)(uint32_t)(old - rm) << 8) / s;
                    w = (uint32_t)(q * q);
                }
            }
            const uint32_t W = (uint32_t)wave_sum((int)w);
            const uint64_t fm = ballot(free_);
            int pick = fm ? __ffsll((unsigned long long)fm) - 1 : 0;
            if (W > 0) {
                const uint32_t u = ig_below(rng, W);
                const uint32_t acc = (uint32_t)scan_add64((int)w);
                pick = __ffsll((unsigned long long)ballot(lane < len && u < acc)) - 1;
            }
            cmax1 = rdlane(rm, pick);
            rjob = rdlane(v, pick);
            v = seq_remove(v, pick);
            len--;
            tabu_jobs |= 1ull << rjob;
            tabu_pos = pick;
        }
        return cmax1;
    }

    // IG::localSearchKI (IG.cpp:566-610): first improvement over positions in
    // visit order, restarted after each improvement.  stop() is asked before
    // every move; once it answers true the search returns at once (the
    // sequence in v is then still a permutation of cost `cost`).
    template <class Stop>
    __device__ __forceinline__ int local_search(int &v, int cost, int kmax, uint64_t &rng, const Stop &stop) const
    {
        for (int it = 0; it < 10000; it++) {
            bool found = false;
            for (int j = 0; j < n; j++) {
                if (stop()) return cost;
                int t = v;
                const int c = ki_move(t, ig_visit(n, j), kmax, cost, rng);
                if (c < cost) {
                    cost = c;
                    v = t;
                    found = true;
                    break;
                }
            }
            if (!found) break;
        }
        return cost;
    }
};

// IG::runIG (IG.cpp:341-404) as one chain on this wavefront: NEH start, then
// `iters` rounds of destruction, greedy re-insertion, local search and
// acceptance.  rng: the chain's stream; th: the instance's acceptance
// thresholds (IG_THR entries).  publish(cost, seq) is called with the NEH
// start and with every later improvement of the chain's best (seq: lane i =
// the job at position i); stop() is asked before every round and every
// local-search move, and ends the chain early when it answers true.
// Returns the chain's best cost; *best_seq receives its sequence.
template <int M, class Stop, class Publish>
__device__ __forceinline__ int ig_chain(const IgWave<M> &G, uint64_t &rng, const unsigned long long *th, int iters,
                                        int *best_seq, const Stop &stop, const Publish &publish)
{
    const int n = G.n, lane = G.lane;
    // NEH start (runIG calls runNEH on its subproblem, IG.cpp:357)
    int curcost = neh_wave<M>(G.pm, PM_STRIDE, n, G.scr);
    int cur = lane < n ? G.scr[lane] : 0;
    int best = cur, bestcost = curcost;
    publish(bestcost, best);
    int kmax = 0;
    while ((kmax + 1) * (kmax + 1) <= n) kmax++;
    int d = 2;
    for (int it = 0; it < iters && !stop(); it++) {
        int tmp = cur;
        // destruction (IG.cpp:94-150): d positions from a shuffle of the
        // first r of the visit order, r uniform in [d, n]
        const int dd = d < n - 1 ? d : n - 1;
        const int r = dd + (int)ig_below(rng, (uint32_t)(n - dd + 1));
        int vv = lane < r ? ig_visit(n, lane) : 0;
        for (int i = r - 1; i > 0; i--) {
            const int j = (int)ig_below(rng, (uint32_t)(i + 1));
            const int vi = rdlane(vv, i), vj = rdlane(vv, j);
            vv = lane == i ? vj : (lane == j ? vi : vv);
        }
        uint64_t gone = 0ull;
        int removed = 0;  // lane t < dd: the t-th removed job
        for (int t = 0; t < dd; t++) {
            const int p = rdlane(vv, t);
            gone |= 1ull << p;
            const int job = rdlane(tmp, p);
            removed = lane == t ? job : removed;
        }
        // compact the kept positions
        const bool keep = lane < n && !((gone >> lane) & 1ull);
        const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
        const int np = lane - __popcll(gone & below);
        __builtin_amdgcn_wave_barrier();
        if (keep) G.scr[np] = tmp;
        __builtin_amdgcn_wave_barrier();
        int len = n - dd;
        tmp = lane < len ? G.scr[lane] : 0;
        __builtin_amdgcn_wave_barrier();
        // construction (IG.cpp:295-317): each removed job at its first best position
        for (int t = 0; t < dd; t++) G.best_insert(tmp, len, rdlane(removed, t), -1);
        // local search (IG.cpp:378)
        const int tc = G.local_search(tmp, G.heads_tails(tmp, n), kmax, rng, stop);
        // acceptance (IG.cpp:319-338, thresholds from the host)
        bool acc;
        if (tc < curcost) {
            acc = true;
        } else {
            const int delta = tc - curcost;
            const uint64_t x = ig_next(rng) >> 32;
            acc = delta < IG_THR && x < th[delta];
        }
        if (acc) {
            cur = tmp;
            curcost = tc;
            d = 2;
        } else if (++d > 5) {
            d = 2;
        }
        if (tc < bestcost) {
            bestcost = tc;
            best = tmp;
            publish(bestcost, best);
        }
    }
    *best_seq = best;
    return bestcost;
}

}  // namespace pbb
