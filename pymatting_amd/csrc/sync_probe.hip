// Cost of the two ways a multi-step solve can separate its steps on gfx950:
// a kernel boundary (the next launch on the stream) or a grid-wide barrier
// inside one persistent kernel. hm_time_sync measures both for a grid of G
// workgroups, which is the measurement behind running the batched and
// single-image solves as one launch per step instead of a persistent kernel.
//
// The barriers: cooperative groups' grid.sync() (cooperative launch), and a
// counter barrier (one agent-scope atomic add per workgroup and a bounded
// spin on the counter; a spin that runs out sets a failure flag and leaves,
// so the kernel always ends).
#include <hip/hip_cooperative_groups.h>

#include "common.h"
#include "internal.h"

namespace hm {
namespace {

namespace cg = cooperative_groups;

constexpr int kProbeBlock = 256;
constexpr long long kSpinLimit = 1ll << 24;

__global__ __launch_bounds__(kProbeBlock) void k_probe_empty(int* sink) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && sink) sink[1] += 1;
}

// reps grid-wide barriers with a neighbour read between them.
__global__ __launch_bounds__(kProbeBlock) void k_probe_coop(int reps, double* data) {
    cg::grid_group g = cg::this_grid();
    double v = 0.0;
    for (int i = 0; i < reps; ++i) {
        if (threadIdx.x == 0) data[blockIdx.x] = v + i;
        g.sync();
        v += data[(blockIdx.x + 1) % gridDim.x];
    }
    if (threadIdx.x == 0) data[gridDim.x + blockIdx.x] = v;
}

__global__ __launch_bounds__(kProbeBlock) void k_probe_counter(int reps, unsigned* counter,
                                                               int* fail, double* data) {
    const unsigned nb = gridDim.x;
    double v = 0.0;
    bool failed = false;
    for (int i = 0; i < reps && !failed; ++i) {
        if (threadIdx.x == 0) data[blockIdx.x] = v + i;
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned target = (unsigned)(i + 1) * nb;
            long long spin = 0;
            while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                   target) {
                if (++spin > kSpinLimit) {
                    atomicOr(fail, 1);
                    break;
                }
            }
            __threadfence();
        }
        __syncthreads();
        // Every workgroup leaves once any spin has run out.
        failed = __hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
        v += data[(blockIdx.x + 1) % nb];
    }
    if (threadIdx.x == 0) data[nb + blockIdx.x] = v;
}

}  // namespace
}  // namespace hm

using namespace hm;

extern "C" int hm_time_sync(int workgroups, int reps, double* launch_us, double* coop_us,
                            double* counter_us, int* ok, void* stream) {
    if (workgroups < 1 || reps < 1 || reps > 100000 || !launch_us || !coop_us || !counter_us ||
        !ok) {
        set_error("hm_time_sync: invalid argument");
        return HM_ERR_ARG;
    }
    int dev = 0, cus = 0;
    HM_HIP(hipGetDevice(&dev));
    HM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // Every workgroup of a barrier must be resident at once: at most one per
    // compute unit.
    if (workgroups > cus) {
        set_error("hm_time_sync: at most %d workgroups (one per compute unit)", cus);
        return HM_ERR_ARG;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t data_bytes = sizeof(double) * 2 * (size_t)workgroups;
    void* raw = nullptr;
    HM_TRY(dev_alloc(&raw, data_bytes + 256));
    struct BufGuard {
        void* p;
        hipStream_t s;
        ~BufGuard() {
            (void)hipStreamSynchronize(s);
            dev_free(p);
        }
    } bg{raw, s};
    double* data = static_cast<double*>(raw);
    unsigned* counter = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(data) + data_bytes);
    int* flags = reinterpret_cast<int*>(counter + 16);
    HM_HIP(hipMemsetAsync(counter, 0, 256, s));
    hipEvent_t e0, e1;
    HM_HIP(hipEventCreate(&e0));
    HM_HIP(hipEventCreate(&e1));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{e0, e1};
    float ms = 0.f;
    // Kernel boundaries: reps back-to-back launches of the grid.
    for (int i = 0; i < 10; ++i) k_probe_empty<<<workgroups, kProbeBlock, 0, s>>>(flags);
    HM_LAUNCH_CHECK();
    HM_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) k_probe_empty<<<workgroups, kProbeBlock, 0, s>>>(flags);
    HM_LAUNCH_CHECK();
    HM_HIP(hipEventRecord(e1, s));
    HM_HIP(hipEventSynchronize(e1));
    HM_HIP(hipEventElapsedTime(&ms, e0, e1));
    *launch_us = 1e3 * ms / reps;
    // Cooperative grid barrier: one launch with reps barriers (after a
    // warm-up launch).
    {
        int r = reps;
        void* args[] = {&r, &data};
        int warm = 10;
        void* wargs[] = {&warm, &data};
        HM_HIP(hipLaunchCooperativeKernel((const void*)k_probe_coop, dim3(workgroups),
                                          dim3(kProbeBlock), wargs, 0, s));
        HM_HIP(hipEventRecord(e0, s));
        HM_HIP(hipLaunchCooperativeKernel((const void*)k_probe_coop, dim3(workgroups),
                                          dim3(kProbeBlock), args, 0, s));
        HM_HIP(hipEventRecord(e1, s));
        HM_HIP(hipEventSynchronize(e1));
        HM_HIP(hipEventElapsedTime(&ms, e0, e1));
        *coop_us = 1e3 * ms / reps;
    }
    // Counter barrier (a warm-up launch of 10 barriers, then reps).
    k_probe_counter<<<workgroups, kProbeBlock, 0, s>>>(10, counter, flags, data);
    HM_LAUNCH_CHECK();
    HM_HIP(hipMemsetAsync(counter, 0, sizeof(unsigned), s));
    HM_HIP(hipEventRecord(e0, s));
    k_probe_counter<<<workgroups, kProbeBlock, 0, s>>>(reps, counter, flags, data);
    HM_LAUNCH_CHECK();
    HM_HIP(hipEventRecord(e1, s));
    HM_HIP(hipEventSynchronize(e1));
    HM_HIP(hipEventElapsedTime(&ms, e0, e1));
    *counter_us = 1e3 * ms / reps;
    int fail = 0;
    HM_HIP(hipMemcpyAsync(&fail, flags, sizeof(int), hipMemcpyDeviceToHost, s));
    HM_HIP(hipStreamSynchronize(s));
    *ok = fail ? 0 : 1;
    return HM_OK;
}
