// pbb_setup.hip -- set-up kernels of the one-instance engine (compiled once,
// independent of the machine count).  They replace host-to-device copies of
// explorer-sized arrays, so pbb_init_full and pbb_run issue launches on the
// engine's stream and the host waits once, after the explorer kernel.
#include <stddef.h>

#include "pbb_kernels.h"

namespace pbb {

// sequentialbb::initFullInterval (multicore/src/sequentialbb.cpp:66-81): the
// interval [0...0, (n-1)...0] goes to explorer 0, every other explorer is
// idle; the control block is cleared with one explorer holding work.  One
// thread per explorer (pos/end rows of 64 bytes) and per control-block word.
__global__ __launch_bounds__(256) void k_setup_full(KArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.E) {
        a.state[t] = t == 0 ? ST_UNFOLD : ST_IDLE;
        int *pw = (int *)(a.pos + (size_t)t * 64);
        int *ew = (int *)(a.end + (size_t)t * 64);
        for (int w = 0; w < 16; w++) {
            pw[w] = 0;
            int packed = 0;
            if (t == 0)
                for (int b = 0; b < 4; b++) {
                    const int i = 4 * w + b;
                    const int d = i < a.I.n ? a.I.n - 1 - i : 0;
                    packed |= (d & 0xff) << (8 * b);
                }
            ew[w] = packed;
        }
    }
    constexpr int words = (int)(sizeof(Ctl) / sizeof(int));
    if (t < words) {
        constexpr int wc = (int)(offsetof(Ctl, work_count) / sizeof(int));
        ((int *)a.ctl)[t] = t == wc ? 1 : 0;
    }
}

// before every explorer launch: the idle ring empty (cell i waits for ticket
// i and names no explorer, both ticket counters 0) and the stop word clear
__global__ __launch_bounds__(256) void k_prelaunch(KArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.ring_cap) a.ring[t] = (unsigned long long)(unsigned)t << 32;
    if (t == 0) {
        a.ctl->idle_head = 0u;
        a.ctl->idle_tail = 0u;
        a.ctl->stop = 0;
    }
}

hipError_t launch_setup_full(const KArgs &a, hipStream_t s)
{
    constexpr int words = (int)(sizeof(Ctl) / sizeof(int));
    const int threads = a.E > words ? a.E : words;
    hipLaunchKernelGGL(k_setup_full, dim3((threads + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_prelaunch(const KArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(k_prelaunch, dim3((a.ring_cap + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace pbb
