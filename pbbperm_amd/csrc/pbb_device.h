// pbb_device.h -- wave64 primitives and the per-wavefront bound routines of
// the gfx950 IVM engine.  One IVM explorer = one 64-lane wavefront:
//   * lane i holds row cell i of the current IVM row (job id, negative when
//     pruned/visited, the flag encoding of common/include/misc.h:19-29);
//   * lane l holds pos[l], end[l], dir[l] of the factoradic interval;
//   * lane s holds entry s of three register stacks (one VGPR per machine):
//       F[s] = completion times of the first s front jobs,
//       B[s] = reverse completion times of the last s back jobs,
//       R[s] = per-machine sum of the jobs still unscheduled at depth s,
//     so the front/back/remain vectors of bound_fsp_weak::computePartial
//     (bounds/src/bound_fsp_weak.cpp:61-130) are read with one v_readlane per
//     machine instead of being recomputed from scratch at every node.
// All arithmetic is int32 max-plus; the recurrences are those of the
// reference, so every value is bit-identical.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PBB_WAVE 64
#define PBB_INF 0x7fffffff

namespace pbb {

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ unsigned long long uni64(unsigned long long v)
{
    unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ int rdlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// write a uniform value into lane l (v_cmp + v_cndmask)
__device__ __forceinline__ int wrlane(int old, int v, int l)
{
    return __lane_id() == l ? v : old;
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// lanes 0..k-1 as a mask, computed in scalar registers (0 <= k <= 64).  A
// ballot of `lane < k && cond` compiles to a VGPR 0/1 select and a second
// compare; lanes_below(k) & ballot(cond) keeps cond's compare mask as is.
__device__ __forceinline__ uint64_t lanes_below(int k) { return k >= 64 ? ~0ull : ((1ull << k) - 1ull); }

__device__ __forceinline__ int shfl(int v, int src) { return __shfl(v, src, PBB_WAVE); }

// Wavefront reductions on DPP lanes moves (row_shr 1,2,4,8 inside each
// 16-lane row, row_bcast:15 and row_bcast:31 across rows): the result is
// complete in lane 63 and read back with v_readlane, without an LDS round
// trip per step.
#define PBB_DPP_REDUCE(NAME, OP, IDENT)                                              \
    __device__ __forceinline__ int NAME(int x)                                       \
    {                                                                                \
        x = OP(x, __builtin_amdgcn_update_dpp(IDENT, x, 0x111, 0xF, 0xF, false));   \
        x = OP(x, __builtin_amdgcn_update_dpp(IDENT, x, 0x112, 0xF, 0xF, false));   \
        x = OP(x, __builtin_amdgcn_update_dpp(IDENT, x, 0x114, 0xF, 0xF, false));   \
        x = OP(x, __builtin_amdgcn_update_dpp(IDENT, x, 0x118, 0xF, 0xF, false));   \
        x = OP(x, __builtin_amdgcn_update_dpp(IDENT, x, 0x142, 0xA, 0xF, false));   \
        x = OP(x, __builtin_amdgcn_update_dpp(IDENT, x, 0x143, 0xC, 0xF, false));   \
        return __builtin_amdgcn_readlane(x, 63);                                     \
    }

__device__ __forceinline__ int op_add(int a, int b) { return a + b; }
__device__ __forceinline__ int op_min(int a, int b) { return min(a, b); }
__device__ __forceinline__ int op_max(int a, int b) { return max(a, b); }
PBB_DPP_REDUCE(wave_sum, op_add, 0)
PBB_DPP_REDUCE(wave_min, op_min, 0x7fffffff)
PBB_DPP_REDUCE(wave_max, op_max, (int)0x80000000)
#undef PBB_DPP_REDUCE

// lanes i and 32+i both receive x[i] (gfx950 v_permlane32_swap)
__device__ __forceinline__ int bcast_low_half(int x)
{
    return __builtin_amdgcn_permlane32_swap(x, x, false, false)[0];
}
// lanes i and 32+i both receive x[32+i]
__device__ __forceinline__ int bcast_high_half(int x)
{
    return __builtin_amdgcn_permlane32_swap(x, x, false, false)[1];
}

__device__ __forceinline__ int flag_neg(int j) { return j < 0 ? j : -j - 1; }  // misc.h negative()
__device__ __forceinline__ int flag_abs(int j) { return j >= 0 ? j : -j - 1; }  // misc.h absolute()

// Instance tables as seen by the device.
struct DevInst {
    int n, m;
    int mp;               // padded stride of ptmT (multiple of 4)
    const int *ptm;       // [m][n] machine-major
    const int *ptmT;      // [n][mp] job-major
    const int *min_arr;   // [m] bound_fsp_weak.cpp:270-283
    const int *min_dep;   // [m] bound_fsp_weak.cpp:253-267
    int npairs;
    const int *pair_ab;   // [npairs] a | b << 8, (i<j) lexicographic (bound_fsp_strong.cpp:282-297)
    const int4 *jtab;     // [npairs][n] {job, p_a, p_b, lag} in Johnson order
};

// ---------------------------------------------------------------------------
// LB1 of one child per lane (bound_fsp_weak.cpp:133-191).
// fv/bv/rv: uniform front/back/remain vectors.  pj: the lane's job's
// processing times.  Writes the front-direction bound to lbF and, when both,
// the back-direction bound to lbB; chF/chB receive the child's own front
// (resp. back) completion chain (the values scheduleFront/scheduleBack would
// compute with the child job appended), which the refinement bounds reuse.
template <int M>
__device__ __forceinline__ void lb1_child(const int (&fv)[M], const int (&bv)[M], const int (&rv)[M],
                                          const int (&pj)[M], bool both, int &lbF, int &lbB,
                                          int (&chF)[M], int (&chB)[M])
{
    int t0 = fv[0] + pj[0];
    int lb = fv[0] + rv[0] + bv[0];
    chF[0] = t0;
#pragma unroll
    for (int k = 1; k < M; k++) {
        int t1 = max(fv[k], t0);
        lb = max(lb, t1 + rv[k] + bv[k]);
        t0 = t1 + pj[k];
        chF[k] = t0;
    }
    lbF = lb;
    if (both) {
        t0 = bv[M - 1] + pj[M - 1];
        lb = fv[M - 1] + rv[M - 1] + bv[M - 1];
        chB[M - 1] = t0;
#pragma unroll
        for (int k = M - 2; k >= 0; k--) {
            int t1 = max(t0, bv[k]);
            lb = max(lb, fv[k] + rv[k] + t1);
            t0 = t1 + pj[k];
            chB[k] = t0;
        }
        lbB = lb;
    } else {
        lbB = 0;
    }
}

// LB1' node bound (bound_fsp_weak.cpp:288-311) from explicit vectors.
template <int M>
__device__ __forceinline__ int lb1_node_vec(const int (&f)[M], const int (&r)[M], const int (&b)[M])
{
    int t0 = f[0] + r[0];
    int lb = t0 + b[0];
#pragma unroll
    for (int k = 1; k < M; k++) {
        int t1 = max(f[k] + r[k], t0);
        lb = max(lb, t1 + b[k]);
        t0 = t1;
    }
    return lb;
}

// Machine pairs of the Johnson bound, one per lane and chunk: chunk c of
// lane l holds pair c*64 + l of the list as a | b << 8, or -1 when there is
// no such pair or the JohnsonMode drops it (pairs_mode 1: pairs ending on the
// last machine; 2: adjacent machines; bound_fsp_strong.cpp:282-297).  Built
// once per explorer, so the pair list is not re-read for every child.
template <int PCH>
__device__ __forceinline__ void pair_lanes(const int *pair_ab, int npairs, int m, int pairs_mode, int (&pinfo)[PCH])
{
    const int lane = lane_id();
#pragma unroll
    for (int c = 0; c < PCH; c++) {
        const int p = c * PBB_WAVE + lane;
        int v = -1;
        if (p < npairs) {
            const int ab = pair_ab[p];
            const int a = ab & 0xff, b = ab >> 8;
            bool use = true;
            if (pairs_mode == 1 && b < m - 1) use = false;
            if (pairs_mode == 2 && b != a + 1) use = false;
            if (use) v = a | (b << 8);
        }
        pinfo[c] = v;
    }
}

// Johnson LB2 of one partial schedule over the machine pairs of pinfo
// (bound_fsp_strong.cpp:50-100, :102-162 without early exit, as
// bornes_calculer calls them, :540-560): fs / bs are its front and back completion
// times (LDS, indexed by machine), unsched the bit set of its unscheduled
// jobs, jtab the per-pair Johnson order {job, p_a, p_b, lag} of all n jobs.
// Each lane runs its pair's two-machine recurrence over the unscheduled jobs
// in that order; the loop body is predicated, so a lane never branches on a
// job, and jtab is read with one 16-byte load per job (ds_read_b128 when the
// caller passes the LDS copy: a pointer that may be either would make every
// access a flat load).  N32: n <= 32, the membership test is one bit
// extract of the low word.
template <int PCH, bool N32>
__device__ __forceinline__ int lb2_pairs(const int (&pinfo)[PCH], const int4 *jtab, int n, const int *fs,
                                         const int *bs, uint64_t unsched)
{
    const unsigned unsched32 = (unsigned)unsched;
    const int lane = lane_id();
    int best = 0;
#pragma unroll
    for (int c = 0; c < PCH; c++) {
        const int info = pinfo[c];
        if (info >= 0) {
            const int a = info & 0xff, b = info >> 8;
            int t0 = fs[a], t1 = fs[b];
            const int4 *row = jtab + (size_t)(c * PBB_WAVE + lane) * n;
#pragma unroll 4
            for (int i = 0; i < n; i++) {
                const int4 e = row[i];
                bool use;
                if constexpr (N32)
                    use = (unsched32 >> e.x) & 1u;
                else
                    use = (unsched >> e.x) & 1ull;
                const int u0 = t0 + e.y;
                const int u1 = max(t1, u0 + e.w) + e.z;
                t0 = use ? u0 : t0;
                t1 = use ? u1 : t1;
            }
            best = max(best, max(t1 + bs[b], t0 + bs[a]));
        }
    }
    return wave_max(best);
}

// lb2_pairs with the table in LDS (jl) or in global memory, for any n
template <int PCH>
__device__ __forceinline__ int lb2_pairs_any(const int (&pinfo)[PCH], bool jl, const int4 *jtabL,
                                             const int4 *jtabG, int n, const int *fs, const int *bs,
                                             uint64_t unsched)
{
    if (n <= 32)
        return jl ? lb2_pairs<PCH, true>(pinfo, jtabL, n, fs, bs, unsched)
                  : lb2_pairs<PCH, true>(pinfo, jtabG, n, fs, bs, unsched);
    return jl ? lb2_pairs<PCH, false>(pinfo, jtabL, n, fs, bs, unsched)
              : lb2_pairs<PCH, false>(pinfo, jtabG, n, fs, bs, unsched);
}

}  // namespace pbb
