// pbb_noalloc.hip -- NEH on the device with no allocation, for callers that
// cannot allocate device memory (code under HIP graph capture, serving
// loops).  The inputs and outputs are the caller's device buffers; the only
// other device memory the call uses -- the instances' offsets into the
// packed time array -- lives in a fixed array of this library, one per
// device, held until the process ends.
//
// Compiled once (not per machine count): the array below must have exactly
// one definition, and its device address is taken in this file.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pbb_internal.h"
#include "pbb_kernels.h"

namespace pbb {

// Offsets of up to PBB_NEH_NOALLOC_MAX instances (256 KB per device).  Every
// call on a device overwrites the entries it uses before reading them, and
// calls that overlap on one device race on them.
static __device__ __align__(256) long long scratch[PBB_NEH_NOALLOC_MAX];

// Exclusive prefix sum of the instance sizes n[i] * m[i] into off[0..count).
// Sizes with 1 <= n <= 63 and 1 <= m <= 20 count as given, whether or not the
// NEH kernel then runs the instance (it refuses n < 3, m < 2, m > max_m), so
// the caller's packing is followed; a size outside that range counts as 0,
// so one garbage entry cannot push the offsets of the others past
// count x 63 x 20 entries.
// One workgroup: thread t sums a contiguous run of entries, the per-thread
// sums are scanned in LDS, then each thread writes its run.
__global__ __launch_bounds__(1024) void k_neh_offsets(const int *ns, const int *ms, int count, long long *off)
{
    __shared__ long long part[1024];
    const int t = threadIdx.x, nt = blockDim.x;
    const int per = (count + nt - 1) / nt;
    const int lo = min(count, t * per), hi = min(count, lo + per);
    auto size_of = [&](int i) -> long long {
        const int n = ns[i], m = ms[i];
        return (n >= 1 && n <= PBB_MAX_JOBS && m >= 1 && m <= PBB_MAX_MACHINES) ? (long long)n * m : 0ll;
    };
    long long sum = 0;
    for (int i = lo; i < hi; i++) sum += size_of(i);
    part[t] = sum;
    __syncthreads();
    // inclusive Hillis-Steele scan over the nt partial sums
    for (int d = 1; d < nt; d <<= 1) {
        const long long v = t >= d ? part[t - d] : 0ll;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long run = part[t] - sum;  // exclusive prefix of this thread's run
    for (int i = lo; i < hi; i++) {
        off[i] = run;
        run += size_of(i);
    }
}

// The device address of `scratch`, looked up once per device with that
// device current (the array's name is not a device address in host code).
static int scratch_ptr(long long **out)
{
    static std::mutex mu;
    static long long *ptr[64] = {};
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(PBB_E_NODEVICE, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(PBB_E_HIP, "no device address for the NEH offset array");
        ptr[dev] = (long long *)p;
    }
    *out = ptr[dev];
    return PBB_OK;
}

}  // namespace pbb

using namespace pbb;

extern "C" int pbb_neh_device_noalloc(int count, const int32_t *d_ptms, const int *d_n, const int *d_m, int max_m,
                                      int *d_perms, int *d_costs, void *stream)
{
    if (count < 0) return fail(PBB_E_ARG, "negative instance count");
    if (count > PBB_NEH_NOALLOC_MAX)
        return fail(PBB_E_CAPACITY, "%d instances exceed the allocation-free NEH limit of %d", count,
                    PBB_NEH_NOALLOC_MAX);
    if (max_m < 2 || max_m > PBB_MAX_MACHINES) return fail(PBB_E_ARG, "max_m=%d outside [2, %d]", max_m, PBB_MAX_MACHINES);
    if (count == 0) return PBB_OK;
    if (!d_ptms || !d_n || !d_m || !d_perms || !d_costs) return fail(PBB_E_ARG, "null device pointer");
    long long *off = nullptr;
    int rc = scratch_ptr(&off);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_neh_offsets, dim3(1), dim3(1024), 0, s, d_n, d_m, count, off);
    HIPCHK(hipGetLastError());
    HIPCHK(launch_neh(max_m, d_ptms, off, d_n, d_m, count, PBB_MAX_JOBS, d_perms, d_costs, s));
    return PBB_OK;
}
