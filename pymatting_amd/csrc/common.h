// Shared definitions for the pymatting_amd HIP library (gfx950 / CDNA4).
//
// Everything here is device-agnostic plumbing: status codes, the thread-local
// error message, deterministic FP64 block reductions and the device-resident
// state of a conjugate-gradient solve.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../include/pymatting_amd.h"

namespace hm {

void set_error(const char* fmt, ...);

constexpr int kWave = 64;

// Status propagation for HIP runtime calls inside functions that return int.
#define HM_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::hm::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),        \
                            __FILE__, __LINE__);                                          \
            return HM_ERR_HIP;                                                            \
        }                                                                                 \
    } while (0)

#define HM_TRY(expr)                  \
    do {                              \
        int s_ = (expr);              \
        if (s_ != HM_OK) return s_;   \
    } while (0)

#define HM_LAUNCH_CHECK() HM_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// Device-resident state of one PCG solve. Kernels read the scalars they need
// (alpha, beta) from here and the final block of each reduction writes them,
// so an iteration needs no host round trip. `done` freezes every later kernel
// of the solve once the residual test passed.
struct CgState {
    double rz;      // <r, z> of the current iterate
    double pq;      // <p, A p>
    double alpha;   // rz / pq
    double beta;    // (1 / rz_old) * rz_new
    double rr;      // ||r||^2
    double norm_b;  // ||b||
    double atol;
    double rtol;
    double aux;     // scratch scalar for one-off reductions
    double rr0;     // ||r||^2 of the initial residual (kDotInitRes)
    int done;       // 1 once ||r|| < atol or ||r|| < rtol * ||b||
    int iters;      // completed iterations
    int nonfinite;  // set when a reduction produced NaN/Inf
    int pad;
};

// ---------------------------------------------------------------------------
// Deterministic reductions. Each block reduces its values with wavefront
// shuffles and an LDS stage and writes one partial per value; a one-block
// finalize kernel sums the partials in a fixed order.
struct Reducer {
    double* partials;   // [N][max_blocks]
    int max_blocks;
    unsigned upd_blocks;  // partials of a deferred CG update (slot 2)
};

template <int N>
__device__ __forceinline__ void wave_sum(double (&v)[N]) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, kWave);
    }
}

// Sum over the block; the result is valid in thread 0.
template <int BS, int N>
__device__ __forceinline__ void block_sum(double (&v)[N]) {
    static_assert(BS % kWave == 0 && BS <= 1024, "block size");
    constexpr int NW = BS / kWave;
    __shared__ double red_lds[N][NW];
    wave_sum<N>(v);
    const int lane = threadIdx.x % kWave;
    const int wid = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) red_lds[k][wid] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) s += red_lds[k][w];
            v[k] = s;
        }
    }
}

// Stage one of a reduction: thread 0 writes this block's sums to
// partials[k * max_blocks + block]. Stage two (k_finalize) runs as its own
// one-workgroup kernel after the producing kernel, so the partials are
// visible through the kernel boundary without a device-scope fence per block
// (on gfx950 such a fence writes back the XCD's L2).
template <int BS, int N>
__device__ __forceinline__ void block_partials_at(double (&v)[N], const Reducer& red,
                                                  unsigned slot) {
    block_sum<BS, N>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) red.partials[k * red.max_blocks + slot] = v[k];
    }
}

template <int BS, int N>
__device__ __forceinline__ void block_partials(double (&v)[N], const Reducer& red) {
    block_partials_at<BS, N>(v, red, blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
}

__device__ __forceinline__ bool solve_done(const CgState* st) {
    return st != nullptr && __hip_atomic_load(&st->done, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// Optional per-kernel-class timing with HIP events (hm_prof_*). When enabled,
// each launch of a class is bracketed by events on its own stream; the
// events are read after the measured region.
enum ProfKind {
    kProfCfApply = 0, kProfCfRestrict, kProfCfSmooth, kProfCfResidual,
    kProfDiaApply, kProfDiaRestrict, kProfDiaSmooth, kProfDiaResidual,
    kProfCgUpdate, kProfCgPupdate, kProfDense, kProfFgSweep,
    kProfCoarseApply, kProfCoarseRestrict, kProfCoarseSmooth, kProfCoarseResidual,
    kProfCgUpdateX64,  // float32 CG update with the solution in float64
    kProfCgUpdateR,    // CG update of r alone (x updated beside the V-cycle tail)
    kProfLevel1Apply, kProfLevel1Restrict, kProfLevel1Smooth, kProfLevel1Residual,
    kProfCfRecApply, kProfCfRecRestrict, kProfCfRecSmooth, kProfCfRecResidual,
    kProfPairDown,  // two coarse levels of the descent in one launch
    kProfPairUp,    // two coarse levels of the ascent in one launch
    kProfCount
};

// Whether per-launch profiling (hm_prof_set) is on.
bool prof_on();

struct ProfScope {
    ProfScope(int kind, long long points, int prec, hipStream_t s);
    ~ProfScope();
    int slot;
    hipStream_t stream;
};

inline unsigned div_up(long long a, long long b) { return static_cast<unsigned>((a + b - 1) / b); }

}  // namespace hm
