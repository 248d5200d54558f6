// Prefix sums of int64 counts (CSR row pointers, ranks of unknown pixels).
//
// Three launches: every 1024-thread workgroup scans a chunk of kScanChunk
// elements (eight consecutive ones per thread, then the threads' totals by
// wavefront shuffles and an LDS pass over the 16 wavefronts) and records the
// chunk's total; one workgroup scans the chunk totals the same way; every
// chunk then adds the total of the chunks before it. Integer sums, so the
// result does not depend on the order.
#include "common.h"
#include "internal.h"

namespace hm {
namespace {

constexpr int kScanBlock = 1024;
constexpr int kScanPer = 8;
constexpr long long kScanChunk = (long long)kScanBlock * kScanPer;

// Inclusive scan of v over the workgroup (v is this thread's value); returns
// the thread's inclusive prefix and sets *block_total.
__device__ __forceinline__ long long block_inclusive_scan(long long v, long long* block_total) {
    __shared__ long long wsum[kScanBlock / kWave];
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const long long u = __shfl_up(v, off, kWave);
        if (lane >= off) v += u;
    }
    if (lane == kWave - 1) wsum[wid] = v;
    __syncthreads();
    if (wid == 0) {
        constexpr int NW = kScanBlock / kWave;
        long long w = lane < NW ? wsum[lane] : 0;
#pragma unroll
        for (int off = 1; off < NW; off <<= 1) {
            const long long u = __shfl_up(w, off, kWave);
            if (lane >= off) w += u;
        }
        if (lane < NW) wsum[lane] = w;  // inclusive prefix of wavefront totals
    }
    __syncthreads();
    const long long before = wid > 0 ? wsum[wid - 1] : 0;
    *block_total = wsum[kScanBlock / kWave - 1];
    return v + before;
}

// out[i + 1] = in[0] + ... + in[i] within each chunk; sums[b] = chunk b's total.
__global__ __launch_bounds__(kScanBlock) void k_scan_chunks(const int64_t* __restrict__ in,
                                                            int64_t* __restrict__ out,
                                                            long long n, int64_t* __restrict__ sums) {
    const long long base = (long long)blockIdx.x * kScanChunk + (long long)threadIdx.x * kScanPer;
    int64_t v[kScanPer];
    long long t = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const long long i = base + k;
        v[k] = i < n ? in[i] : 0;
        t += v[k];
    }
    long long total;
    const long long incl = block_inclusive_scan(t, &total);
    long long run = incl - t;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        run += v[k];
        const long long i = base + k;
        if (i < n) out[i + 1] = run;
    }
    if (threadIdx.x == 0) {
        sums[blockIdx.x] = total;
        if (blockIdx.x == 0) out[0] = 0;
    }
}

// Exclusive scan of the nb <= kScanChunk chunk totals, in place.
__global__ __launch_bounds__(kScanBlock) void k_scan_totals(int64_t* __restrict__ sums,
                                                            long long nb) {
    const long long base = (long long)threadIdx.x * kScanPer;
    int64_t v[kScanPer];
    long long t = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const long long i = base + k;
        v[k] = i < nb ? sums[i] : 0;
        t += v[k];
    }
    long long total;
    const long long incl = block_inclusive_scan(t, &total);
    long long run = incl - t;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const long long i = base + k;
        if (i < nb) sums[i] = run;
        run += v[k];
    }
}

// Adds the chunks before it to every element of chunk b >= 1.
__global__ __launch_bounds__(kScanBlock) void k_scan_add(int64_t* __restrict__ out, long long n,
                                                         const int64_t* __restrict__ sums) {
    const long long b = (long long)blockIdx.x + 1;
    const int64_t add = sums[b];
    const long long base = b * kScanChunk + (long long)threadIdx.x * kScanPer;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const long long i = base + k;
        if (i < n) out[i + 1] += add;
    }
}

}  // namespace

int exclusive_offsets(const int64_t* in, int64_t* out, long long n, int64_t* total,
                      hipStream_t s) {
    if (n < 0 || n > kScanChunk * kScanChunk) {
        set_error("prefix sum of %lld elements (at most %lld)", n, kScanChunk * kScanChunk);
        return HM_ERR_ARG;
    }
    const long long nb = n > 0 ? (n + kScanChunk - 1) / kScanChunk : 1;
    void* sums = nullptr;
    HM_TRY(dev_alloc(&sums, sizeof(int64_t) * (size_t)nb));
    int64_t* sm = static_cast<int64_t*>(sums);
    k_scan_chunks<<<(unsigned)nb, kScanBlock, 0, s>>>(in, out, n, sm);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && nb > 1) {
        k_scan_totals<<<1, kScanBlock, 0, s>>>(sm, nb);
        e = hipGetLastError();
        if (e == hipSuccess) {
            k_scan_add<<<(unsigned)(nb - 1), kScanBlock, 0, s>>>(out, n, sm);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipMemcpyAsync(total, out + n, sizeof(int64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dev_free(sums);
    if (e != hipSuccess) {
        set_error("prefix sum failed: %s", hipGetErrorString(e));
        return HM_ERR_HIP;
    }
    return HM_OK;
}

}  // namespace hm
