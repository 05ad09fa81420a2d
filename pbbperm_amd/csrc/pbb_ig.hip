// pbb_ig.hip -- iterated greedy on the device (heuristic/src/IG.cpp runIG):
// one IG chain per 64-lane wavefront, four chains of one instance per
// workgroup sharing the instance's processing times in LDS.  The chain itself
// (moves, evaluations, acceptance) is in pbb_ig.h.
//
// Translation unit compiled once per machine count M (-DPBB_M=<m>); an
// instance with fewer machines is padded with zero-time machines, which
// changes no makespan.
#include "pbb_ig.h"

namespace pbb {

namespace {

// IG::runIG (IG.cpp:341-404) for chain `chain` of instance `s`.
// thr: the instance's acceptance thresholds (IG_THR entries).
template <int M>
__global__ __launch_bounds__(256) void k_ig(const int *ptms, const long long *off, const int *ns, const int *ms,
                                            int count, int chains, int iters, unsigned long long seed,
                                            const unsigned long long *thr, int *perms, int *costs)
{
    extern __shared__ __align__(16) char smem[];
    const int groups = (chains + 3) / 4;
    const int s = blockIdx.x / groups;
    const int wave = threadIdx.x >> 6;
    const int chain = (blockIdx.x % groups) * 4 + wave;
    const int lane = lane_id();
    if (s >= count) return;
    const int n = ns[s], m = ms[s];
    const bool bad = n < 3 || n > PBB_POOL_ROW - 1 || m < 2 || m > M;
    int *pm = (int *)smem;
    int *spt = pm + M * PM_STRIDE;
    if (!bad) {
        const int *src = ptms + off[s];
        for (int i = threadIdx.x; i < M * n; i += blockDim.x) {
            const int q = i / n, j = i - q * n;
            pm[q * PM_STRIDE + j] = q < m ? src[q * n + j] : 0;
        }
    }
    __syncthreads();
    if (!bad && threadIdx.x < 64) {
        int t = 0;
        if (threadIdx.x < n)
            for (int q = 0; q < M; q++) t += pm[q * PM_STRIDE + threadIdx.x];
        spt[threadIdx.x] = t;
    }
    __syncthreads();
    if (chain >= chains) return;
    const size_t out = (size_t)s * chains + chain;
    if (bad) {
        perms[out * 64 + lane] = -1;
        if (lane == 0) costs[out] = -1;
        return;
    }
    int *wbase = spt + 64 + wave * (2 * M * 64 + 64);
    IgWave<M> G{pm, spt, wbase, wbase + M * 64, wbase + 2 * M * 64, n, lane};
    const unsigned long long *th = thr + (size_t)s * IG_THR;

    // the chain's stream: the seed mixed with the instance and chain numbers
    uint64_t rng = seed ^ (((uint64_t)(uint32_t)s << 32) | (uint32_t)chain);
    rng = ig_next(rng);

    int best = 0;
    const int bestcost = ig_chain<M>(G, rng, th, iters, &best, [] { return false; }, [](int, int) {});
    if (lane < n) perms[out * 64 + lane] = best;
    else perms[out * 64 + lane] = -1;
    if (lane == 0) costs[out] = bestcost;
}

}  // namespace

template <int M>
size_t ig_lds_bytes_m()
{
    return ((size_t)M * PM_STRIDE + 64 + 4 * (2 * (size_t)M * 64 + 64)) * 4;
}

template <int M>
hipError_t launch_ig_m(const int *ptms, const long long *off, const int *ns, const int *ms, int count, int chains,
                       int iters, unsigned long long seed, const unsigned long long *thr, int *perms, int *costs,
                       hipStream_t s)
{
    const size_t lds = ig_lds_bytes_m<M>();
    if (const hipError_t err = hipFuncSetAttribute((const void *)k_ig<M>,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds))
        return err;
    const int groups = (chains + 3) / 4;
    hipLaunchKernelGGL(k_ig<M>, dim3((unsigned)count * groups), dim3(256), lds, s, ptms, off, ns, ms, count, chains,
                       iters, seed, thr, perms, costs);
    return hipGetLastError();
}

#ifdef PBB_M
template hipError_t launch_ig_m<PBB_M>(const int *, const long long *, const int *, const int *, int, int, int,
                                       unsigned long long, const unsigned long long *, int *, int *, hipStream_t);
#endif

}  // namespace pbb
